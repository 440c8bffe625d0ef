#!/usr/bin/env python3
"""Benchmark: bead-steps/s of the fused DPD step on one MI355X.

Workload (default): the 10,077,696-bead DPD fluid (216^3 jittered lattice in a
150^3 box, rho=3, rc=1, a=25, gamma=4.5, sigma=3, dt=0.04), FP32 pair
arithmetic with FP64 accumulators and FP64 state.  A "step" is one whole
velocity-Verlet step: pair force + final half-kick + half-kick + drift + re-binning.

The served-mix half of the metric (512 Poisson replica requests, continuous
batching, p50/p99 step latency, against the plain one-at-a-time path) is the
"serving" object of the same line; --no-serve skips it.

The headline value times the shipped path: blocks of graph-captured steps
replayed on the library's stream, kernel timing off.  A second pass of the same
number of steps, eager, brackets every launch of the dominant kernel (the brick
pair-force kernel k_brick) with HIP events on that stream; the roofline comes
from it.  The CPU baseline is the OpenMP oracle port on bounded samples.

Algorithmic bytes (compulsory HBM traffic, FP64 state, each array touched once):
  force pass (k_brick)   read x 24 + v 24 + id 4 + type 4, write f 24 = 80 B/bead
                         (+ rho 8 read for mDPD)
  integrate (k_integrate) read f, v, x (72), write v, x (48)       = 120 B/bead
  density pass (mDPD)    read x 24 + type 4, write rho 8           =  36 B/bead
  bonds                  12 B/bond
so a DPD step is 200 B/bead and an mDPD step 244 B/bead.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_BEAD_STEP = {"dpd": 200.0, "mdpd": 244.0}
FORCE_PASS_BYTES = {"dpd": 80.0, "mdpd": 88.0}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def time_cpu_port(cfg, seconds, max_steps=200):
    """Steps/s of the OpenMP port (oracle/) on cfg after 2 warm-up steps."""
    import ctypes as C

    from oracle import pyoracle
    s = pyoracle.make_orc_system(cfg, log_every=10 ** 9)
    L = pyoracle.lib()
    nthreads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    h = L.orc_omp_create(C.byref(s), nthreads)
    threads = L.orc_omp_threads(h)
    L.orc_omp_steps(h, 0, 2)  # warm-up: step 0 + one step
    steps, t0, step = 0, time.perf_counter(), 2
    while True:
        L.orc_omp_steps(h, step, 1)
        step += 1
        steps += 1
        el = time.perf_counter() - t0
        if el > seconds or steps >= max_steps:
            break
    L.orc_omp_destroy(h)
    return cfg["N"] * steps / el, steps, threads


def cpu_baseline(seconds=12.0, n_target=1_000_000):
    """OpenMP port of the reference force loop (oracle/): half neighbour list
    with the skin check, the DPD force with per-thread force buffers and nve,
    on the GPU box's host cores.  Headline: a 1M-bead slice of the 10M-bead
    fluid; beside it the 24,000-bead C0 fluid."""
    from gpumdpd_amd import configs
    cfg = configs.scaled_dpd(n_target, seed=2024)
    v, steps, threads = time_cpu_port(cfg, seconds)
    c0 = configs.config_c0()
    v0, steps0, _ = time_cpu_port(c0, min(5.0, seconds), max_steps=2000)
    return {"value": v, "unit": "bead-steps/s", "cores": threads, "kind": "port", "cpu_model": cpu_model(),
            "sample": f"{cfg['N']}-bead DPD fluid (rho=3, same force field as the 10M-bead run), {steps} steps "
                      f"after 2 warm-up, half neighbour list rebuilt on the skin criterion, {threads} OpenMP threads",
            "c0": {"value": v0, "unit": "bead-steps/s", "n_beads": c0["N"], "steps": steps0}}


def serving_leg(ctx, n_req, steps, slots, chunk, plain_every):
    """Served mixed-replica stream (512 Poisson requests at 200/s, 70% DPD at
    rho=3 / 30% mDPD at rho=6, 2,000-200,000 beads log-uniform, 2,000 steps,
    FP32 force) with continuous batching, then the plain path (one replica at a
    time, arrival order) on every plain_every-th request.  Each path runs the
    brick shapes the tuner picks for it: throughput shapes for the batch, and
    for a replica alone the same list plus small bricks."""
    import gpumdpd_amd as g
    from gpumdpd_amd import serve, tune
    t = time.perf_counter()
    tuned = tune.tune_classes(ctx, ("dpd", "mdpd"), g.PREC_F32)
    tuned_alone = tune.tune_classes(ctx, ("dpd", "mdpd"), g.PREC_F32, alone=True)
    log(f"bench: tuned {tuned} (alone: {tuned_alone}) in {time.perf_counter() - t:.1f}s")
    reqs = serve.make_requests(5, n=n_req, steps=steps)
    serve.prepare_inputs(reqs)
    log(f"bench: serving {n_req} requests, {sum(r.n for r in reqs)} beads in total")
    m = serve.serve(ctx, reqs, max_slots=slots, chunk=chunk, tuned=tuned)
    m.pop("kept")
    log(f"bench: served in {m['makespan_s']:.1f}s")
    sample = reqs[::plain_every]
    p = serve.serve_plain(ctx, sample, chunk=chunk, tuned=tuned_alone, realtime=False)
    p.pop("kept")
    log(f"bench: plain path on {len(sample)} requests in {p['makespan_s']:.1f}s")
    return {"metric": "served-mix replica-steps/s", "value": m["replica_steps_per_s"], "unit": "replica-steps/s",
            "step_latency_p50_ms": m["step_latency_p50_ms"], "step_latency_p99_ms": m["step_latency_p99_ms"],
            "step_latency_p999_ms": m["step_latency_p999_ms"],
            "step_latency_samples": m["step_latency_samples"],
            "step_latency_note": "per replica-step, from device step-completion clock readings",
            "request_latency_p50_s": m["request_latency_p50_s"], "request_latency_p99_s": m["request_latency_p99_s"],
            "bead_steps_per_s": m["bead_steps_per_s"], "makespan_s": m["makespan_s"], "requests": n_req,
            "steps_per_request": steps, "slots": slots, "chunk_steps": chunk,
            "tuned": {f"{k[0]}/{'small' if k[1] == 0 else 'large'}": v for k, v in tuned.items()},
            "tuned_alone": {f"{k[0]}/{'small' if k[1] == 0 else 'large'}": v for k, v in tuned_alone.items()},
            "plain": {"replica_steps_per_s": p["replica_steps_per_s"], "bead_steps_per_s": p["bead_steps_per_s"],
                      "step_latency_p50_ms": p["step_latency_p50_ms"], "step_latency_p99_ms": p["step_latency_p99_ms"],
                      "makespan_s": p["makespan_s"],
                      "sample": (f"all {n_req} requests" if plain_every == 1 else
                                 f"every {plain_every}th request ({len(sample)} of {n_req})")
                      + ", one at a time in arrival order, back to back"},
            "speedup_vs_plain": m["replica_steps_per_s"] / p["replica_steps_per_s"]}


def load_traffic(workload):
    p = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(p):
        return None
    try:
        d = json.load(open(p))
        return d.get(workload)
    except Exception:
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--config", default="c3", choices=["c0", "c1", "c2", "c3"])
    ap.add_argument("--precision", default="f32", choices=["f32", "f64"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 1 GiB device to device (reference kernel for scaling rocprof FETCH/WRITE_SIZE)")
    ap.add_argument("--serve-requests", type=int, default=512)
    ap.add_argument("--serve-steps", type=int, default=2000)
    ap.add_argument("--serve-slots", type=int, default=128)
    ap.add_argument("--serve-chunk", type=int, default=64)
    ap.add_argument("--plain-every", type=int, default=1)
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        a = torch.ones(2 ** 28, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": a.numel() * 4}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: --gpus {args.gpus} is not supported "
              "(this build runs a single system or a batch of replicas on one MI355X)", file=sys.stderr)
        sys.exit(2)

    import gpumdpd_amd as g
    from gpumdpd_amd import configs
    t = time.perf_counter()
    cfg = {"c0": configs.config_c0, "c1": configs.config_c1, "c2": configs.config_c2,
           "c3": configs.config_c3}[args.config]()
    log(f"bench: generated {cfg['name']} in {time.perf_counter() - t:.1f}s")
    prec = g.PREC_F32 if args.precision == "f32" else g.PREC_F64
    ctx = g.Context(0)
    s = g.system_from_config(ctx, cfg, precision=prec, log_every=10 ** 6)
    tun = s.tuning()
    log(f"bench: uploaded N={cfg['N']} tuning={tun}")
    s.zero_step()
    s.advance(args.warmup)
    ctx.sync()
    log(f"bench: warm-up {args.warmup} steps done")

    # the shipped path: graph-captured blocks replayed, no per-kernel events
    t0 = time.perf_counter()
    s.advance(args.steps)
    ctx.sync()
    wall = time.perf_counter() - t0
    log(f"bench: timed {args.steps} steps in {wall:.3f}s")
    # kernel pass: the same number of steps, eager, HIP events around each force launch
    ctx.kernel_timing(True)
    s.advance(args.steps)
    ctx.sync()
    k_ms, k_n, _ = ctx.kernel_timing_get()
    phases = ctx.kernel_phases()
    ctx.kernel_timing(False)
    s.finish()

    N = cfg["N"]
    value = N * args.steps / wall
    per = BYTES_PER_BEAD_STEP[cfg["style"]] + (12.0 * len(cfg["bonds"]) / N if "bonds" in cfg else 0.0)
    per_kernel = FORCE_PASS_BYTES[cfg["style"]]
    k_avg = k_ms / max(k_n, 1) / 1e3
    achieved = per_kernel * N / k_avg / 1e9
    step_gbs = per * value / 1e9
    workload = {"c0": "dpd_fluid_24k", "c1": "mdpd_slab_8k", "c2": "diblock_melt_375k",
                "c3": "dpd_fluid_10M"}[args.config]
    out = {
        "metric": ("bead-steps/s at 10M-bead DPD (FP32 force); served-mix replica-steps/s + p99"
                   if args.config == "c3" else f"bead-steps/s ({workload})"),
        "value": value,
        "unit": "bead-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if prec == g.PREC_F32 else "f64",
        "data": "synthetic",
        "config": {"workload": workload, "n_beads": N, "box": list(map(float, cfg["hi"])),
                   "mean_rho": N / float(np.prod(np.asarray(cfg["hi"]) - np.asarray(cfg["lo"]))),
                   "precision": "fp32 pair force, fp64 accumulators and state" if prec == g.PREC_F32 else "fp64",
                   "brick": [tun["bx"], tun["by"], tun["bz"]], "resort_every": tun["resort_every"],
                   "graph_steps": tun["graph_steps"], "timed_path": "graph-replayed fused steps"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": load_traffic(workload),
                     "kernel": "k_brick (pair force pass)", "kernel_ms": k_avg * 1e3,
                     "bytes_per_bead": per_kernel, "launches_timed": k_n,
                     "kernel_timing": "HIP events around every k_brick launch of a second, eager pass of the "
                                      "same number of steps on the library stream",
                     "phases_us_per_wave": phases,
                     "step": {"bytes_per_bead_step": per, "achieved": step_gbs,
                              "frac": step_gbs / HBM_PEAK_GBS}},
    }
    s.destroy()
    if not args.no_serve:
        try:
            out["serving"] = serving_leg(ctx, args.serve_requests, args.serve_steps, args.serve_slots,
                                         args.serve_chunk, args.plain_every)
        except Exception as e:  # reported, not fatal to the headline number
            out["serving"] = {"error": repr(e)}
    ctx.close()
    if not args.no_cpu:
        log("bench: CPU baseline ...")
        out["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
