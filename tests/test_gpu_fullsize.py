"""The 10,077,696-bead DPD fluid (216^3 jittered lattice in a 150^3 box) at full
size, where the oracle is too slow to run in a test: properties that hold at
any size instead.

- The FP32-force mode agrees with the FP64 mode on the same state to the
  stated force tolerance (global relative L2 < 1e-5; per-bead max below 1e-3
  of the RMS force).
- Forces are pairwise: the FP64 total force vanishes to round-off, and the
  fused run conserves momentum to the FP32 pair round-off.
- Re-sort interval and graph length leave the trajectory unchanged bit for bit.
- The per-hook step equals the fused step bit for bit.
"""
import numpy as np
import pytest

import gpumdpd_amd as g
from gpumdpd_amd import configs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = g.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def cfg():
    return configs.config_c3()


def force(ctx, cfg, prec, step):
    s = g.system_from_config(ctx, cfg, precision=prec, log_every=10 ** 6)
    s.pair_compute(step, log=False)
    f = s.download()["force"]
    s.destroy()
    return f


def test_fp32_force_within_tolerance_of_fp64(ctx, cfg):
    print(f"case: C3 force FP32 vs FP64, N={cfg['N']}", flush=True)
    f32 = force(ctx, cfg, g.PREC_F32, 3)
    f64 = force(ctx, cfg, g.PREC_F64, 3)
    rel = np.linalg.norm(f32 - f64) / np.linalg.norm(f64)
    frms = np.sqrt(np.mean(np.sum(f64 * f64, axis=1)))
    worst = np.max(np.linalg.norm(f32 - f64, axis=1)) / frms
    print(f"rel L2 {rel:.3e} worst {worst:.3e}", flush=True)
    assert rel < 1e-5
    assert worst < 1e-3
    # third law: the total of all pair forces vanishes (FP64: to round-off)
    assert np.linalg.norm(f64.sum(axis=0)) < 1e-9 * frms * np.sqrt(cfg["N"])


def test_momentum_conserved_fp32(ctx, cfg):
    print("case: C3 momentum over 20 fused FP32 steps", flush=True)
    s = g.system_from_config(ctx, cfg, precision=g.PREC_F32, log_every=10 ** 6)
    p0 = s.download()["vel"].sum(axis=0)  # unit masses
    s.run(20)
    p1 = s.download()["vel"].sum(axis=0)
    s.destroy()
    # each pair adds equal and opposite integers inside a brick; a pair across
    # two bricks is evaluated once per side, in FP32 relative to each brick
    vrms = np.sqrt(np.mean(np.sum(cfg["vel"] ** 2, axis=1)))
    drift = np.linalg.norm(p1 - p0)
    print(f"momentum drift {drift:.3e} (N v_rms = {cfg['N'] * vrms:.3e})", flush=True)
    assert drift < 1e-7 * cfg["N"] * vrms


def test_resort_interval_full_size(ctx, cfg):
    print("case: C3 re-sort interval independence", flush=True)
    out = []
    for every, graph in ((16, 8), (5, 4)):
        s = g.system_from_config(ctx, cfg, precision=g.PREC_F32, log_every=5)
        s.set_tuning(resort_every=every, graph_steps=graph)
        s.run(12)
        d = s.download()
        out.append((d["pos"], d["vel"], s.thermo_rows()))
        s.destroy()
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    assert np.array_equal(out[0][2], out[1][2])


def test_hook_step_equals_fused_full_size(ctx, cfg):
    print("case: C3 hook path vs fused path, 3 steps", flush=True)
    n = 3
    a = g.system_from_config(ctx, cfg, precision=g.PREC_F32, log_every=10 ** 6)
    a.run(n)
    da = a.download()
    a.destroy()
    b = g.system_from_config(ctx, cfg, precision=g.PREC_F32, log_every=10 ** 6)
    b.pair_compute(0, log=False)
    for step in range(1, n + 1):
        b.nve_initial()
        b.pair_compute(step, log=False)
        b.nve_final()
    db = b.download()
    b.destroy()
    assert np.array_equal(da["pos"], db["pos"]) and np.array_equal(da["vel"], db["vel"])
    assert np.array_equal(da["force"], db["force"])
