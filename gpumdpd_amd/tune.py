"""Autotuner of the step's launch parameters, per pair style and replica size class.

1. Brick shape.  The force kernels stage one brick of bx*by*bz cells (plus a
   one-cell halo) per workgroup, so the brick shape is the kernel's tile.  The
   best tile depends on the bead density per cell (LDS capacity, occupancy) and
   on the pair style.  The tuner times the force launches of a few candidate
   shapes on a sample system of the class and keeps the fastest.
2. Schedule.  With that brick, it times whole graph-replayed steps for each
   physical re-sort interval (bead arrays re-ordered into cell order every
   `resort_every` steps) and each graph length (`graph_steps` steps captured
   per hipGraph) of a small grid, and keeps the fastest pair.
Results are cached per class for the life of the process.

Two shape lists: the batch list (throughput shapes, for replicas stepped
together) and, for a replica stepped alone, that list plus small bricks.  A
small replica alone is latency-bound, so it runs faster with more, smaller
bricks; in a batch the other replicas already fill the chip, and the same small
bricks only add staging work (the served stream ran 11% slower with them).

The other knobs of the reference's neighbour machinery do not exist in this
design: cells are rebuilt every step (no Verlet skin, no rebuild interval), and
the cell edge is the pair cutoff.
"""
import time

from . import configs
from .system import System  # noqa: F401  (re-exported for callers that build samples)

_CACHE = {}

DPD_SHAPES = [(4, 4, 5), (4, 4, 4), (3, 4, 4), (4, 4, 6), (3, 4, 5)]
MDPD_SHAPES = [(3, 3, 4), (3, 3, 3), (2, 3, 4), (3, 3, 5), (3, 4, 4)]
# added for a replica stepped alone
DPD_SHAPES_ALONE = [(3, 3, 3), (2, 3, 3), (2, 2, 3)]
MDPD_SHAPES_ALONE = [(2, 2, 3), (2, 2, 2), (2, 3, 3)]


def _time_force(ctx, cfg, brick, precision, steps):
    import gpumdpd_amd as g
    s = g.system_from_config(ctx, cfg, precision=precision, log_every=10 ** 6, upload=False)
    try:
        s.set_tuning(brick=brick)
        s.upload(cfg["id"], cfg["type"], cfg["pos"], cfg["vel"], cfg["masses"])
        got = s.tuning()
        s.zero_step()
        s.advance(2)
        ctx.sync()
        ctx.kernel_timing(True)
        t0 = time.perf_counter()
        s.advance(steps)
        ctx.sync()
        wall = time.perf_counter() - t0
        ms, n, _ = ctx.kernel_timing_get()
        ctx.kernel_timing(False)
        s.finish()
        return (got["bx"], got["by"], got["bz"]), ms / max(n, 1), wall / steps
    finally:
        s.destroy()


# replica size classes for serving: tiles are tuned on one sample per class
SIZE_CLASSES = [(0, 12000), (30000, 120000)]  # (lower bound of N, sample size)


def size_class(n):
    k = 0
    for i, (lo, _) in enumerate(SIZE_CLASSES):
        if n >= lo:
            k = i
    return k


RESORT_EVERY = (8, 16, 32)
GRAPH_STEPS = (4, 8, 16)


def shapes_for(style, alone=False):
    base = DPD_SHAPES if style == "dpd" else MDPD_SHAPES
    if alone:
        return base + (DPD_SHAPES_ALONE if style == "dpd" else MDPD_SHAPES_ALONE)
    return base


def tune_classes(ctx, styles, precision, alone=False, verbose=False):
    """{(style, size class): {"brick", "resort_every", "graph_steps"}} for
    serve.serve's `tuned` table (alone=True: for replicas stepped one at a
    time, as serve.serve_plain does)."""
    out = {}
    for st in styles:
        for k, (_, n) in enumerate(SIZE_CLASSES):
            b = tune_brick(ctx, st, precision, n_sample=n, shapes=shapes_for(st, alone), verbose=verbose)
            r, gs = tune_schedule(ctx, st, precision, b, n_sample=n, verbose=verbose)
            out[(st, k)] = {"brick": b, "resort_every": r, "graph_steps": gs}
    return out


def _time_steps(ctx, cfg, brick, resort_every, graph_steps, precision, steps):
    import gpumdpd_amd as g
    s = g.system_from_config(ctx, cfg, precision=precision, log_every=10 ** 6, upload=False)
    try:
        s.set_tuning(brick=brick, resort_every=resort_every, graph_steps=graph_steps)
        s.upload(cfg["id"], cfg["type"], cfg["pos"], cfg["vel"], cfg["masses"])
        s.zero_step()
        s.advance(graph_steps + 1)
        ctx.sync()
        t0 = time.perf_counter()
        s.advance(steps)
        ctx.sync()
        wall = time.perf_counter() - t0
        s.finish()
        return wall / steps
    finally:
        s.destroy()


def tune_schedule(ctx, style, precision, brick, n_sample=60000, steps=64, verbose=False):
    """Fastest (resort_every, graph_steps) by whole-step time of graph-replayed
    steps on an n_sample-bead replica with the given brick."""
    key = ("schedule", style, precision, n_sample, tuple(brick))
    if key in _CACHE:
        return _CACHE[key]
    cfg = configs.replica_system(style, n_sample, 12345)
    best = None
    for r in RESORT_EVERY:
        for gs in GRAPH_STEPS:
            t = _time_steps(ctx, cfg, brick, r, gs, precision, steps)
            if verbose:
                print(f"tune {style}: resort {r} graph {gs} step {t * 1e3:.4f} ms", flush=True)
            if best is None or t < best[1]:
                best = ((r, gs), t)
    _CACHE[key] = best[0]
    return best[0]


def tune_brick(ctx, style, precision, n_sample=60000, steps=6, shapes=None, verbose=False):
    """Fastest brick shape for replicas of `style` ("dpd" at rho=3, "mdpd" at
    rho=6) in `precision`, from force-kernel time on an n_sample-bead replica."""
    shapes = shapes or shapes_for(style)
    key = (style, precision, n_sample, tuple(map(tuple, shapes)))
    if key in _CACHE:
        return _CACHE[key]
    cfg = configs.replica_system(style, n_sample, 12345)
    best = None
    for b in shapes:
        try:
            got, kms, sms = _time_force(ctx, cfg, b, precision, steps)
        except Exception as e:  # a shape that does not fit LDS is skipped
            if verbose:
                print(f"tune {style}: {b} skipped ({e})", flush=True)
            continue
        if verbose:
            print(f"tune {style}: brick {got} force {kms:.3f} ms step {sms * 1e3:.3f} ms", flush=True)
        if best is None or kms < best[1]:
            best = (got, kms)
    if best is None:
        raise RuntimeError(f"no brick shape fits for {style}")
    _CACHE[key] = best[0]
    return best[0]
