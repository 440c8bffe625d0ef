"""The fused, hook, batched and C++-shell paths compute the same step.

Per-bead and per-brick sums are 64-bit fixed point (exact, order-independent)
and a pair's contribution is the same whichever of its beads finds it, so the
order of beads in cells, slots and launches does not matter: every path below
must agree bit for bit.
"""
import os
import subprocess

import numpy as np
import pytest

import gpumdpd_amd as g
from gpumdpd_amd import configs

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def ctx():
    c = g.Context(0)
    yield c
    c.close()


def state(s):
    d = s.download()
    return d["pos"], d["vel"]


@pytest.mark.parametrize("style", ["dpd", "mdpd"])
@pytest.mark.parametrize("prec", [g.PREC_F32, g.PREC_F64])
def test_hook_path_equals_fused_path(ctx, prec, style):
    """For mDPD the hook's pair_compute runs both passes (density, then force)."""
    cfg = configs.scaled_dpd(5000, seed=21) if style == "dpd" else configs.config_c1()
    n = 12
    a = g.system_from_config(ctx, cfg, precision=prec, log_every=4)
    a.run(n)
    b = g.system_from_config(ctx, cfg, precision=prec, log_every=4)
    # the reference's loop: zero_step then initial / pair / final per step
    b.pair_compute(0, log=True)
    for step in range(1, n + 1):
        b.nve_initial()
        b.pair_compute(step, log=(step % 4 == 0))
        b.nve_final()
    pa, va = state(a)
    pb, vb = state(b)
    assert np.array_equal(pa, pb) and np.array_equal(va, vb)
    assert np.array_equal(a.download()["force"], b.download()["force"])
    a.destroy()
    b.destroy()


def test_resort_interval_does_not_change_results(ctx):
    cfg = configs.scaled_dpd(6000, seed=22)
    out = []
    for every, graph in ((1, 0), (16, 8), (7, 3)):
        s = g.system_from_config(ctx, cfg, precision=g.PREC_F32, log_every=10)
        s.set_tuning(resort_every=every, graph_steps=graph)
        s.run(30)
        out.append(state(s) + (s.thermo_rows(),))
        s.destroy()
    for o in out[1:]:
        assert np.array_equal(o[0], out[0][0]) and np.array_equal(o[1], out[0][1])
        assert np.array_equal(o[2], out[0][2])


@pytest.mark.parametrize("prec", [g.PREC_F32, g.PREC_F64])
def test_batch_equals_alone(ctx, prec):
    reqs = [configs.replica_system("dpd", 3000, 5), configs.replica_system("mdpd", 2500, 6),
            configs.replica_system("dpd", 9000, 7)]
    alone = []
    for cfg in reqs:
        s = g.system_from_config(ctx, cfg, precision=prec, log_every=5)
        s.zero_step()
        s.advance(20)
        s.finish()
        alone.append(state(s) + (s.thermo_rows(),))
        s.destroy()
    sys_ = [g.system_from_config(ctx, cfg, precision=prec, log_every=5) for cfg in reqs]
    for s in sys_:
        s.zero_step()
    ctx.batch_advance(sys_, 20)
    for s, ref in zip(sys_, alone):
        s.finish()
        p, v = state(s)
        assert np.array_equal(p, ref[0]) and np.array_equal(v, ref[1])
        assert np.array_equal(s.thermo_rows(), ref[2])
        s.destroy()


def test_batch_rejects_mixed_precision(ctx):
    a = g.system_from_config(ctx, configs.replica_system("dpd", 2000, 9), precision=g.PREC_F32)
    b = g.system_from_config(ctx, configs.replica_system("dpd", 2000, 10), precision=g.PREC_F64)
    a.zero_step()
    b.zero_step()
    from gpumdpd_amd import _lib
    with pytest.raises(_lib.GdpdError) as e:
        ctx.batch_advance([a, b], 4)
    assert "precision" in str(e.value)
    a.destroy()
    b.destroy()


def test_cpp_plugin_shells_hook_equals_fused():
    exe = os.path.join(ROOT, "tests", "cpp", "plugin_driver")
    assert os.path.exists(exe), "build() compiles tests/cpp/plugin_driver"
    r = subprocess.run([exe, "20"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    hook = np.array([list(map(float, l.split()[1:])) for l in lines if l.startswith("hook")])
    fused = np.array([list(map(float, l.split()[1:])) for l in lines if l.startswith("fused")])
    assert hook.shape == fused.shape == (5, 4)
    assert np.allclose(hook, fused, rtol=1e-11, atol=1e-11)
    final = [l for l in lines if l.startswith("final")][0].split()
    assert float(final[1]) == 0.0 and float(final[2]) == 0.0


def test_bonded_hook_path_equals_fused_path(ctx):
    cfg = configs.config_c2(n_chains=800, box=16.0)
    n = 8
    a = g.system_from_config(ctx, cfg, precision=g.PREC_F32, log_every=4)
    a.run(n)
    b = g.system_from_config(ctx, cfg, precision=g.PREC_F32, log_every=4)
    b.pair_compute(0, log=True)
    b.bond_compute(0, log=True)
    for step in range(1, n + 1):
        b.nve_initial()
        b.pair_compute(step, log=False)
        b.bond_compute(step, log=False)
        b.nve_final()
    pa, va = state(a)
    pb, vb = state(b)
    assert np.array_equal(pa, pb) and np.array_equal(va, vb)
    assert np.array_equal(a.download()["force"], b.download()["force"])
    a.destroy()
    b.destroy()


def test_angle_hook_path_equals_fused_path(ctx):
    cfg = configs.config_angles(n_chains=600, box=14.0)
    n = 8
    a = g.system_from_config(ctx, cfg, precision=g.PREC_F64, log_every=4)
    a.run(n)
    b = g.system_from_config(ctx, cfg, precision=g.PREC_F64, log_every=4)
    b.pair_compute(0, log=True)
    b.bond_compute(0, log=True)
    b.angle_compute(0, log=True)
    for step in range(1, n + 1):
        b.nve_initial()
        b.pair_compute(step, log=False)
        b.bond_compute(step, log=False)
        b.angle_compute(step, log=False)
        b.nve_final()
    pa, va = state(a)
    pb, vb = state(b)
    assert np.array_equal(pa, pb) and np.array_equal(va, vb)
    assert np.array_equal(a.download()["force"], b.download()["force"])
    a.destroy()
    b.destroy()


def test_bonds_added_after_graph_steps(ctx):
    """The integrate kernel instance depends on whether bonds exist, so a system
    that gains bonds after graph-captured steps must re-capture its step graph:
    graph-replayed and eager runs agree bit for bit, and the bonds act."""
    cfg = configs.config_c2(n_chains=400, box=13.0)
    plain = {k: v for k, v in cfg.items() if k != "bonds"}
    out = []
    for graph, bonds in ((8, True), (1, True), (8, False)):
        s = g.system_from_config(ctx, plain, precision=g.PREC_F32, log_every=10 ** 6)
        s.set_tuning(graph_steps=graph)
        s.run(20)
        if bonds:
            s.set_bonds(cfg["bonds"], cfg["n_bond_types"], cfg["bond_coeff"])
        s.run(20)
        out.append(state(s))
        s.destroy()
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    assert not np.array_equal(out[0][0], out[2][0])


def test_brick_capacity_overflow_is_reported_and_cleared_by_upload(ctx):
    """LDS capacities follow each uploaded state.  A state that packs more beads
    into one cell than any brick can stage in LDS is refused at upload with
    GDPD_ERR_OVERFLOW; the system keeps its brick and needs a fresh upload
    before it computes again, after which it runs as if nothing had happened."""
    from gpumdpd_amd import _lib
    cfg = configs.scaled_dpd(6000, seed=31)
    s = g.system_from_config(ctx, cfg, precision=g.PREC_F32, log_every=10 ** 6)
    s.run(2)
    L = np.asarray(cfg["hi"]) - np.asarray(cfg["lo"])
    packed = np.asarray(cfg["lo"]) + 0.5 * L + np.random.default_rng(3).random((cfg["N"], 3)) * 1.5
    with pytest.raises(_lib.GdpdError) as e:
        s.upload(cfg["id"], cfg["type"], packed, cfg["vel"], cfg["masses"])
    assert e.value.code == 3 and "LDS staging" in str(e.value)
    with pytest.raises(_lib.GdpdError) as e:
        s.pair_compute(step=0, log=False)
    assert e.value.code == 4
    s.upload(cfg["id"], cfg["type"], cfg["pos"], cfg["vel"], cfg["masses"])
    s.run(2)
    ref = g.system_from_config(ctx, cfg, precision=g.PREC_F32, log_every=10 ** 6)
    ref.run(2)
    assert np.array_equal(state(s)[0], state(ref)[0]) and np.array_equal(state(s)[1], state(ref)[1])
    s.destroy()
    ref.destroy()


def test_batch_with_bonded_and_angle_members_equals_alone(ctx):
    """A batch runs the integrate instance of its most complex member (here one
    with angles) for every member: plain, bonded and angle replicas still give
    the same bits as alone."""
    reqs = [configs.replica_system("dpd", 4000, 8), configs.config_c2(n_chains=300, box=12.0),
            configs.config_angles(n_chains=250, box=11.0)]
    alone = []
    for cfg in reqs:
        s = g.system_from_config(ctx, cfg, precision=g.PREC_F32, log_every=5)
        s.zero_step()
        s.advance(20)
        s.finish()
        alone.append(state(s) + (s.thermo_rows(),))
        s.destroy()
    sys_ = [g.system_from_config(ctx, cfg, precision=g.PREC_F32, log_every=5) for cfg in reqs]
    for s in sys_:
        s.zero_step()
    ctx.batch_advance(sys_, 20)
    for s, ref in zip(sys_, alone):
        s.finish()
        p, v = state(s)
        assert np.array_equal(p, ref[0]) and np.array_equal(v, ref[1])
        assert np.array_equal(s.thermo_rows(), ref[2])
        s.destroy()


def test_tuner_picks_brick_and_schedule(ctx):
    """tune_classes' records carry the brick and the timed schedule, and a replica
    created from such a record runs with them."""
    from gpumdpd_amd import tune
    b = tune.tune_brick(ctx, "dpd", g.PREC_F32, n_sample=4000, steps=3, shapes=[(3, 3, 3), (4, 4, 4)])
    r, gs = tune.tune_schedule(ctx, "dpd", g.PREC_F32, b, n_sample=4000, steps=16)
    assert r in tune.RESORT_EVERY and gs in tune.GRAPH_STEPS
    s = g.system_from_config(ctx, configs.replica_system("dpd", 4000, 12345), precision=g.PREC_F32)
    s.set_tuning(brick=b, resort_every=r, graph_steps=gs)
    t = s.tuning()
    assert (t["bx"], t["by"], t["bz"]) == tuple(b) and t["resort_every"] == r and t["graph_steps"] == gs
    s.destroy()


def test_admission_window(ctx):
    """Systems created and zero-stepped inside an admission window run on the
    context's second stream beside queued main-stream work; they, and the
    systems already running, end with the same bits as without the window.
    Inside the window the running systems are out of reach."""
    from gpumdpd_amd import _lib
    ca, cb = configs.replica_system("dpd", 30000, 13), configs.replica_system("mdpd", 6000, 14)
    ref = []
    for cfg, n in ((ca, 39 + 19), (cb, 19)):
        s = g.system_from_config(ctx, cfg, precision=g.PREC_F32, log_every=5)
        s.zero_step()
        s.advance(n)
        s.finish()
        ref.append(state(s) + (s.thermo_rows(),))
        s.destroy()
    print("case: admission window", flush=True)
    a = g.system_from_config(ctx, ca, precision=g.PREC_F32, log_every=5)
    a.zero_step()
    a.advance(39)  # queued on the main stream
    ctx.admission(True)
    try:
        b = g.system_from_config(ctx, cb, precision=g.PREC_F32, log_every=5)
        b.zero_step()
        with pytest.raises(_lib.GdpdError) as e:
            a.advance(1)
        assert "admission window" in str(e.value)
    finally:
        ctx.admission(False)
    ctx.batch_advance([a, b], 19)
    for s, r in ((a, ref[0]), (b, ref[1])):
        s.finish()
        p, v = state(s)
        assert np.array_equal(p, r[0]) and np.array_equal(v, r[1])
        assert np.array_equal(s.thermo_rows(), r[2])
        s.destroy()


def test_kernel_phases_measure_timed_force_launches(ctx):
    """With kernel timing on, every wave of the force kernel adds its phase
    durations; the phases are positive, the timed results equal the untimed
    ones bit for bit (timing adds clock reads only), and a reset clears them."""
    cfg = configs.scaled_dpd(6000, seed=21)
    ref = g.system_from_config(ctx, cfg, precision=g.PREC_F32)
    ref.run(6)
    want = ref.download()
    ref.destroy()
    s = g.system_from_config(ctx, cfg, precision=g.PREC_F32)
    ctx.kernel_timing(True)
    s.run(6)
    ms, n, _ = ctx.kernel_timing_get()
    ph = ctx.kernel_phases()
    ctx.kernel_timing(False)
    got = s.download()
    assert n == 7 and ms > 0
    assert ph["stage_us"] > 0 and ph["pairs_us"] > 0 and ph["writeback_us"] > 0 and ph["barrier_us"] >= 0
    # a wave's phases fit inside the summed launch time
    assert sum(ph.values()) * 1e-3 < ms
    for k in ("pos", "vel", "force"):
        assert np.array_equal(got[k], want[k])
    ctx.kernel_timing(True)
    assert all(v == 0 for v in ctx.kernel_phases().values())
    ctx.kernel_timing(False)
    s.destroy()
