"""Accuracy report of the HIP path (not collected by pytest; run on a GPU box).

    python tests/accuracy_report.py [--c3] [--c3-oracle] [--c0] [--c1] [--c2]

Prints one JSON object per section:

c3   10,077,696-bead DPD fluid.  States of the FP32-force run at steps 0, 100
     and 500; at each, the force of the FP32 mode and of the FP64 mode on that
     same state, compared as the global relative L2 error ||F - F_ref|| / ||F_ref||
     and the per-bead max |dF| / F_rms.  With --c3-oracle the FP64 oracle is
     also evaluated on the step-0 state (about a minute of CPU time).
c0   24,000-bead DPD fluid.
       energy drift: gamma = 0 (conservative only), 1,000 steps, FP32 and FP64
       modes and the oracle: |E(t) - E(0)| / (|E(0)| * t) at the end, and its
       maximum over the logged steps;
       thermostatted run (gamma = 4.5), 1,000 steps: <T> and <P> over the last
       800 steps with block-average standard errors, FP32, FP64 and the oracle.
c1   8,000-bead mDPD liquid-vapour slab, 5,000 steps at dt = 0.01; the density
     profile along x (bins of 0.5) averaged over the last 2,000 steps, FP32
     mode against FP64 mode.
c2   diblock melt (375k beads): FP32 vs FP64 force on the initial state.
full each config's whole run on the fused path (C0 1,000 steps FP64, C1 5,000
     FP32, C2 10,000 FP32, C3 500 FP32 and 500 FP64): the LDS capacities at
     the start and at the end (they grow only if a brick outgrew them), and T
     and P of the last thermo row.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import gpumdpd_amd as g  # noqa: E402
from gpumdpd_amd import configs  # noqa: E402


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def max_rel(a, b):
    frms = np.sqrt(np.mean(np.sum(b * b, axis=1)))
    return float(np.max(np.linalg.norm(a - b, axis=1)) / frms)


def force_at(ctx, cfg, pos, vel, step, prec):
    s = g.system_from_config(ctx, cfg, precision=prec, upload=False)
    s.upload(cfg["id"], cfg["type"], pos, vel, cfg["masses"])
    s.pair_compute(step=step, log=False)
    if "bonds" in cfg:
        s.bond_compute(step=step, log=False)
    f = s.download()["force"]
    s.destroy()
    return f


def section_c3(ctx, with_oracle):
    cfg = configs.config_c3()
    out = {"section": "c3", "n_beads": cfg["N"], "states": []}
    s = g.system_from_config(ctx, cfg, precision=g.PREC_F32, log_every=10 ** 6)
    s.zero_step()
    done = 0
    for step in (0, 100, 500):
        if step > 0:
            s.advance(step - done)
            done = step
        d = s.download()
        # the state after the drift of step `step` is the one whose force is
        # evaluated at Saru step `step + 1`; for step 0 use the initial state
        pos, vel, t = (cfg["pos"], cfg["vel"], 0) if step == 0 else (d["pos"], d["vel"], step + 1)
        f32 = force_at(ctx, cfg, pos, vel, t, g.PREC_F32)
        f64 = force_at(ctx, cfg, pos, vel, t, g.PREC_F64)
        rec = {"step": t, "fp32_vs_fp64_rel_l2": rel_l2(f32, f64), "fp32_vs_fp64_max_rel": max_rel(f32, f64)}
        if with_oracle and step == 0:
            from oracle import pyoracle
            c2 = dict(cfg, pos=pos, vel=vel)
            t0 = time.perf_counter()
            fo = pyoracle.forces(c2, step=t)[0]
            rec.update(oracle_s=time.perf_counter() - t0, fp64_vs_oracle_rel_l2=rel_l2(f64, fo),
                       fp32_vs_oracle_rel_l2=rel_l2(f32, fo), fp32_vs_oracle_max_rel=max_rel(f32, fo))
        out["states"].append(rec)
        print(f"c3 state {t}: {rec}", file=sys.stderr, flush=True)
    s.finish()
    s.destroy()
    return out


def thermo_stats(rows, first_step):
    """<T>, <P> and block-average standard errors over rows with step >= first_step."""
    r = rows[rows[:, -1] >= first_step] if rows.shape[1] == 12 else rows
    out = {}
    for name, col in (("T", 0), ("P", 3)):
        x = r[:, col]
        nb = 8
        blocks = np.array_split(x, nb)
        means = np.array([b.mean() for b in blocks])
        out[name] = float(x.mean())
        out[name + "_stderr"] = float(means.std(ddof=1) / np.sqrt(nb))
    return out


def energy_drift(e, steps):
    e0 = e[0]
    d = np.abs(e - e0) / (abs(e0) * np.maximum(steps, 1))
    return float(d[-1]), float(np.max(d[1:]))


def section_c0(ctx, with_oracle):
    out = {"section": "c0"}
    nsteps, every = 1000, 10
    # conservative only: gamma = 0, so sigma = 0 and E = KE + PE is conserved
    cfg = configs.config_c0()
    cons = dict(cfg, coeff=configs.dpd_coeff(1, {(1, 1): (25.0, 0.0, 1.0)}))
    for name, prec in (("fp32", g.PREC_F32), ("fp64", g.PREC_F64)):
        s = g.system_from_config(ctx, cons, precision=prec, log_every=every)
        s.run(nsteps)
        rows = s.thermo_rows()
        s.destroy()
        e = rows[:, 10] + rows[:, 1]
        end, mx = energy_drift(e, rows[:, 11])
        out[f"drift_{name}_end"] = end
        out[f"drift_{name}_max"] = mx
        print(f"c0 drift {name}: end {end:.3e} max {mx:.3e}", file=sys.stderr, flush=True)
    if with_oracle:
        from oracle import pyoracle
        rows, _, _ = pyoracle.run(cons, nsteps, log_every=every)
        e = rows[:, 10] + rows[:, 1]
        st = np.arange(len(rows)) * every
        end, mx = energy_drift(e, st)
        out["drift_oracle_end"] = end
        out["drift_oracle_max"] = mx
        print(f"c0 drift oracle: end {end:.3e}", file=sys.stderr, flush=True)
    # thermostatted run: <T>, <P>
    for name, prec in (("fp32", g.PREC_F32), ("fp64", g.PREC_F64)):
        s = g.system_from_config(ctx, cfg, precision=prec, log_every=every)
        s.run(nsteps)
        rows = s.thermo_rows()
        s.destroy()
        out[f"thermo_{name}"] = thermo_stats(rows[rows[:, 11] >= 200], 200)
    if with_oracle:
        from oracle import pyoracle
        rows, _, _ = pyoracle.run(cfg, nsteps, log_every=every)
        st = np.arange(len(rows)) * every
        out["thermo_oracle"] = thermo_stats(rows[st >= 200], 200)
    print(f"c0 thermo: {out}", file=sys.stderr, flush=True)
    return out


def density_profile(ctx, cfg, prec, nsteps, avg_from, every=20, bin_w=0.5):
    s = g.system_from_config(ctx, cfg, precision=prec, log_every=10 ** 6)
    lx = cfg["hi"][0] - cfg["lo"][0]
    nb = int(round(lx / bin_w))
    hist = np.zeros(nb)
    n = 0
    s.zero_step()
    done = 1
    while done < nsteps:
        k = min(every, nsteps - done)
        s.advance(k)
        done += k
        if done >= avg_from:
            x = s.download()["pos"][:, 0] - cfg["lo"][0]
            hist += np.bincount(np.minimum((x / bin_w).astype(int), nb - 1), minlength=nb)
            n += 1
    s.finish()
    s.destroy()
    vol = bin_w * (cfg["hi"][1] - cfg["lo"][1]) * (cfg["hi"][2] - cfg["lo"][2])
    return hist / (n * vol)


def section_c1(ctx):
    cfg = configs.config_c1()
    p32 = density_profile(ctx, cfg, g.PREC_F32, 5000, 3000)
    p64 = density_profile(ctx, cfg, g.PREC_F64, 5000, 3000)
    out = {"section": "c1", "bins": len(p32), "liquid_density_fp32": float(np.max(p32)),
           "liquid_density_fp64": float(np.max(p64)), "profile_max_abs_diff": float(np.max(np.abs(p32 - p64))),
           "profile_rel_l2": rel_l2(p32, p64), "profile_fp32": [round(float(v), 4) for v in p32],
           "profile_fp64": [round(float(v), 4) for v in p64]}
    print(f"c1: liquid {out['liquid_density_fp32']:.3f} / {out['liquid_density_fp64']:.3f}", file=sys.stderr,
          flush=True)
    return out


def section_c2(ctx):
    cfg = configs.config_c2()
    f32 = force_at(ctx, cfg, cfg["pos"], cfg["vel"], 0, g.PREC_F32)
    f64 = force_at(ctx, cfg, cfg["pos"], cfg["vel"], 0, g.PREC_F64)
    return {"section": "c2", "n_beads": cfg["N"], "fp32_vs_fp64_rel_l2": rel_l2(f32, f64),
            "fp32_vs_fp64_max_rel": max_rel(f32, f64)}


def section_full(ctx):
    out = {"section": "full", "runs": []}
    runs = (("c0", configs.config_c0, 1000, g.PREC_F64, 100), ("c1", configs.config_c1, 5000, g.PREC_F32, 1000),
            ("c2", configs.config_c2, 10000, g.PREC_F32, 1000), ("c3", configs.config_c3, 500, g.PREC_F32, 100),
            ("c3", configs.config_c3, 500, g.PREC_F64, 100))
    for name, make, nsteps, prec, every in runs:
        cfg = make()
        s = g.system_from_config(ctx, cfg, precision=prec, log_every=every)
        t0 = s.tuning()
        t = time.perf_counter()
        s.run(nsteps)
        wall = time.perf_counter() - t
        t1 = s.tuning()
        rows = s.thermo_rows()
        s.destroy()
        r = {"config": name, "n_beads": cfg["N"], "steps": nsteps, "precision": "fp64" if prec == g.PREC_F64 else "fp32",
             "brick": [t1["bx"], t1["by"], t1["bz"]], "smax_omax_start": [t0["smax"], t0["omax"]],
             "smax_omax_end": [t1["smax"], t1["omax"]], "T_end": float(rows[-1, 0]), "P_end": float(rows[-1, 3]),
             "wall_s": round(wall, 2)}
        out["runs"].append(r)
        print(f"full: {r}", file=sys.stderr, flush=True)
    return out


def main():
    ap = argparse.ArgumentParser()
    for k in ("c3", "c3-oracle", "c0", "c0-oracle", "c1", "c2", "full"):
        ap.add_argument("--" + k, action="store_true")
    a = ap.parse_args()
    ctx = g.Context(0)
    if a.c3 or a.c3_oracle:
        print(json.dumps(section_c3(ctx, a.c3_oracle)), flush=True)
    if a.c0 or a.c0_oracle:
        print(json.dumps(section_c0(ctx, a.c0_oracle)), flush=True)
    if a.c1:
        print(json.dumps(section_c1(ctx)), flush=True)
    if a.c2:
        print(json.dumps(section_c2(ctx)), flush=True)
    if a.full:
        print(json.dumps(section_full(ctx)), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
