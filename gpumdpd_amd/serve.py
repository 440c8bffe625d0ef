"""Continuous-batching server for independent replica simulations on one GPU.

A stream of requests (each: pair style, bead count, seed, number of steps)
arrives over time.  The server keeps every admitted replica's state resident
in HBM and advances all of them together: each call of
gdpd_batch_advance covers every active replica's bricks in one launch per
phase per step (graph-captured).  Requests that arrive while others run are
admitted between chunks as soon as a slot and enough of the HBM budget are
free; finished replicas leave the batch and free theirs.

Per-replica results do not depend on what else is in the batch: the kernels'
per-bead sums are fixed-point and each replica keeps its own brick shape, so
a replica gives the same bits alone and in any mix.

Metrics:
  replica-steps/s  total replica-steps / makespan (first arrival to last finish,
                   host clock)
  bead-steps/s     total bead-steps / makespan
  step latency     time between consecutive step completions of a replica, one
                   sample per step: the device writes a clock reading when each
                   step's last kernel runs (gdpd_step_times), so the first step
                   of a chunk includes the host turnaround and any wait to join
                   the batch
  request latency  arrival to finish (host clock)
The plain baseline runs the same requests one at a time in arrival order,
each alone on the GPU, with the same chunking.
"""
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from . import configs


@dataclass
class Request:
    k: int
    style: str
    n: int
    seed: int
    arrival: float
    steps: int
    cfg: dict = None
    bytes: int = 0


@dataclass
class Replica:
    req: Request
    sys: object
    remaining: int
    t_admit: float
    t_last: float
    lat: list = field(default_factory=list)
    t_done: float = 0.0
    seen: int = 0  # completed steps whose latency has been recorded (step 0 = the zero step)


def _collect_latencies(ctx, reps):
    """Append the per-step latencies (ms) of every step each replica completed
    since the last call, from the device's step-completion clock."""
    if not reps:
        return
    times, steps = ctx.step_times([a.sys for a in reps])
    cap = times.shape[1]
    for k, a in enumerate(reps):
        done = int(steps[k]) - 1  # last completed step
        if done > a.seen:
            st = np.arange(a.seen + 1, done + 1)
            a.lat.extend((times[k, st % cap] - times[k, (st - 1) % cap]).tolist())
            a.seen = done


def make_requests(master_seed=5, n=512, rate=200.0, steps=2000, n_max=None, n_scale=1.0):
    """The served stream: Poisson arrivals at `rate`/s, 70% DPD / 30% mDPD,
    log-uniform sizes (2,000-200,000 beads; optionally capped or scaled for
    small runs), seeds from `master_seed`."""
    arr = configs.arrival_times(master_seed, n, rate)
    reqs = []
    for k in range(n):
        style, nb, seed = configs.replica_request(master_seed, k)
        nb = max(200, int(nb * n_scale))
        if n_max:
            nb = min(nb, n_max)
        reqs.append(Request(k, style, nb, seed, float(arr[k]), steps))
    return reqs


def prepare_inputs(reqs):
    """Generate every request's initial state (host arrays) before serving."""
    for r in reqs:
        r.cfg = configs.replica_system(r.style, r.n, r.seed)
        r.bytes = _system_bytes(r.cfg)


def _system_bytes(cfg):
    from .system import make_desc, system_bytes
    desc = make_desc(cfg["N"], cfg["ntypes"], cfg["lo"], cfg["hi"], cfg["dt"], cfg["global_cut"], cfg["skin"],
                     cfg["max_rho"], 1, 100, L.PREC_F32)
    return system_bytes(desc, 0)


def _create(ctx, req, precision, log_every, tuned):
    import gpumdpd_amd as g
    from .tune import size_class
    s = g.system_from_config(ctx, req.cfg, precision=precision, log_every=log_every, upload=False)
    # tuned: brick shape (and re-sort interval, graph length) by (style, size class) or by style
    t = tuned.get((req.style, size_class(req.n)), tuned.get(req.style))
    if isinstance(t, dict):
        s.set_tuning(brick=t["brick"], resort_every=t.get("resort_every", -1), graph_steps=t.get("graph_steps", -1))
    elif t:
        s.set_tuning(brick=t)
    c = req.cfg
    s.upload(c["id"], c["type"], c["pos"], c["vel"], c["masses"])
    return s


def _percentiles(lat):
    if not lat:
        return None, None
    a = np.asarray(lat)
    return float(np.percentile(a, 50)), float(np.percentile(a, 99))


def _p999(lat):
    return float(np.percentile(np.asarray(lat), 99.9)) if lat else None


def serve(ctx, reqs, max_slots=128, hbm_budget=200e9, chunk=64, precision=L.PREC_F32, log_every=100,
          tuned=None, keep=None, realtime=True):
    """Run the stream with continuous batching.  `keep`: request ids whose
    final state (pos, vel, thermo rows) is returned.  Returns a metrics dict."""
    tuned = tuned or {}
    keep = set(keep or ())
    kept = {}
    order = sorted(reqs, key=lambda r: r.arrival)
    nxt = 0
    active = []
    used = 0
    lat_all, req_lat = [], []
    total_steps = 0
    total_bead_steps = 0
    batches = 0
    t0 = time.perf_counter()

    def now():
        return time.perf_counter() - t0

    def admit():
        # new replicas are set up on the context's second stream while the
        # batch's chunk runs on the main one
        nonlocal nxt, used
        t = now()
        opened = False
        try:
            while (nxt < len(order) and (order[nxt].arrival <= t or not realtime) and len(active) + len(joining)
                   < max_slots and used + order[nxt].bytes <= hbm_budget):
                if not opened:
                    ctx.admission(True)
                    opened = True
                r = order[nxt]
                s = _create(ctx, r, precision, log_every, tuned)
                s.zero_step()
                used += r.bytes
                joining.append(Replica(r, s, r.steps - 1, t, t))
                nxt += 1
                t = now()
        finally:
            if opened:
                ctx.admission(False)

    joining = []
    while nxt < len(order) or active or joining:
        for a in joining:
            a.t_last = now()
        active.extend(joining)
        joining = []
        if not active:
            admit()
            if not joining and nxt < len(order):
                time.sleep(max(0.0, order[nxt].arrival - now()))
            continue
        # launch the next chunk for the batch, admit arrivals while it runs
        # (their set-up queues behind it on the same stream), then wait
        k = min(chunk, min(a.remaining for a in active))
        if k > 0:
            ctx.batch_advance([a.sys for a in active], k)
            batches += 1
        admit()
        ctx.sync()
        for a in active:
            a.remaining -= k
        _collect_latencies(ctx, active)
        still, leaving = [], []
        for a in active:
            if a.remaining > 0:
                still.append(a)
            else:
                a.sys.finish()
                leaving.append(a)
        _collect_latencies(ctx, leaving)
        for a in leaving:
            td = now()
            a.t_done = td
            lat_all.extend(a.lat)
            req_lat.append(td - a.req.arrival)
            total_steps += a.req.steps
            total_bead_steps += a.req.steps * a.req.n
            if a.req.k in keep:
                d = a.sys.download()
                kept[a.req.k] = (d["pos"], d["vel"], a.sys.thermo_rows(), d["force"])
            a.sys.destroy()
            used -= a.req.bytes
        active = still
    makespan = now() - (order[0].arrival if (order and realtime) else 0.0)
    p50, p99 = _percentiles(lat_all)
    r50, r99 = _percentiles(req_lat)
    return {"replica_steps_per_s": total_steps / makespan, "bead_steps_per_s": total_bead_steps / makespan,
            "makespan_s": makespan, "step_latency_p50_ms": p50, "step_latency_p99_ms": p99,
            "step_latency_p999_ms": _p999(lat_all),
            "step_latency_samples": len(lat_all), "request_latency_p50_s": r50, "request_latency_p99_s": r99,
            "batches": batches, "requests": len(reqs), "replica_steps": total_steps, "kept": kept}


def serve_plain(ctx, reqs, chunk=64, precision=L.PREC_F32, log_every=100, tuned=None, keep=None,
                realtime=True):
    """Plain baseline: the same requests one at a time in arrival order."""
    tuned = tuned or {}
    keep = set(keep or ())
    kept = {}
    order = sorted(reqs, key=lambda r: r.arrival)
    lat_all, req_lat = [], []
    total_steps = total_bead_steps = 0
    t0 = time.perf_counter()

    def now():
        return time.perf_counter() - t0

    for r in order:
        if realtime and r.arrival > now():
            time.sleep(r.arrival - now())
        s = _create(ctx, r, precision, log_every, tuned)
        s.zero_step()
        a = Replica(r, s, r.steps - 1, now(), now())
        while a.remaining > 0:
            k = min(chunk, a.remaining)
            s.advance(k)
            ctx.sync()
            a.remaining -= k
            _collect_latencies(ctx, [a])
        s.finish()
        _collect_latencies(ctx, [a])
        lat_all.extend(a.lat)
        td = now()
        req_lat.append(td - r.arrival)
        total_steps += r.steps
        total_bead_steps += r.steps * r.n
        if r.k in keep:
            d = s.download()
            kept[r.k] = (d["pos"], d["vel"], s.thermo_rows(), d["force"])
        s.destroy()
    makespan = now() - (order[0].arrival if (order and realtime) else 0.0)
    p50, p99 = _percentiles(lat_all)
    r50, r99 = _percentiles(req_lat)
    return {"replica_steps_per_s": total_steps / makespan, "bead_steps_per_s": total_bead_steps / makespan,
            "makespan_s": makespan, "step_latency_p50_ms": p50, "step_latency_p99_ms": p99,
            "step_latency_p999_ms": _p999(lat_all),
            "step_latency_samples": len(lat_all), "request_latency_p50_s": r50, "request_latency_p99_s": r99,
            "requests": len(reqs), "replica_steps": total_steps, "kept": kept}
