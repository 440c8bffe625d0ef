"""Runs that outgrow their brick capacities, long-lived image counts, and the
reference neighbour list's skin check after real motion.

Overflow: a brick whose beads exceed its LDS capacity holds its system at the
last completed step; the library sizes the capacities again from that state
and runs the missing steps.  The run must equal one that never overflowed, bit
for bit (the capacities only bound LDS staging, the brick shape is kept).  The
capacities are set low through gdpd_set_capacity, so the overflow is forced
from the host and every kernel exits normally.
"""
import numpy as np
import pytest

import gpumdpd_amd as g
from gpumdpd_amd import configs
from oracle import pyoracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = g.Context(0)
    yield c
    c.close()


def final_state(s):
    d = s.download()
    return d["pos"], d["vel"], d["force"], s.thermo_rows()


def assert_same(a, b):
    for x, y in zip(a, b):
        assert np.array_equal(np.asarray(x).view(np.uint8), np.asarray(y).view(np.uint8))


def auto_capacity(ctx, cfg, prec):
    s = g.system_from_config(ctx, cfg, precision=prec, log_every=10)
    t = s.tuning()
    s.destroy()
    # the automatic sizing is ceil(1.12 max) + 32 staged and ceil(1.15 max) + 16 owned
    return int((t["smax"] - 32) / 1.12), int((t["omax"] - 16) / 1.15), t


def run_fused(ctx, cfg, prec, nsteps, cap=None, cap_at=0):
    """zero_step, advances of 7, 20, 30, ... steps, finish.  cap: LDS capacities
    set before the run (cap_at 0) or after the first advance (cap_at 1)."""
    s = g.system_from_config(ctx, cfg, precision=prec, log_every=10)
    if cap and cap_at == 0:
        s.set_capacity(*cap)
    s.zero_step()
    done = 0
    for i, k in enumerate((7, 20, 30, 10 ** 9)):
        k = min(k, nsteps - 1 - done)
        if k > 0:
            s.advance(k)
            done += k
        if cap and cap_at == 1 and i == 0:
            s.set_capacity(*cap)
    s.finish()
    out = final_state(s)
    t = s.tuning()
    s.destroy()
    return out, t


@pytest.mark.parametrize("style", ["dpd", "mdpd"])
@pytest.mark.parametrize("cap_at", [0, 1])
def test_overflow_recovers_bit_exact(ctx, style, cap_at):
    """Capacities 20% under the starting maximum: set before the run, the zero
    step itself overflows; set after 7 steps, step 8 overflows mid-run."""
    cfg = configs.scaled_dpd(12000, seed=21) if style == "dpd" else configs.replica_system("mdpd", 9000, 31)
    prec = g.PREC_F32
    print(f"case: overflow recovery style={style} cap_at={cap_at}", flush=True)
    smax0, omax0, t0 = auto_capacity(ctx, cfg, prec)
    ref, _ = run_fused(ctx, cfg, prec, 60)
    low = (int(0.8 * smax0), int(0.8 * omax0))
    got, t = run_fused(ctx, cfg, prec, 60, cap=low, cap_at=cap_at)
    assert (t["bx"], t["by"], t["bz"]) == (t0["bx"], t0["by"], t0["bz"])
    assert t["smax"] > low[0] and t["omax"] > low[1]  # grown back from the state
    assert_same(got, ref)


def test_overflow_in_batch_recovers(ctx):
    """One replica of a batch overflows mid-run; it and its batch mates each
    equal their uninterrupted runs."""
    cfgs = [configs.scaled_dpd(6000, seed=41), configs.replica_system("mdpd", 5000, 42),
            configs.scaled_dpd(9000, seed=43)]
    prec = g.PREC_F32
    print("case: overflow inside a batch", flush=True)
    refs = [run_fused(ctx, c, prec, 50)[0] for c in cfgs]
    smax0, omax0, _ = auto_capacity(ctx, cfgs[1], prec)
    systems = [g.system_from_config(ctx, c, precision=prec, log_every=10) for c in cfgs]
    for s in systems:
        s.zero_step()
    for i, k in enumerate((16, 16, 17)):
        ctx.batch_advance(systems, k)
        ctx.sync()
        if i == 0:
            systems[1].set_capacity(int(0.8 * smax0), int(0.8 * omax0))
    for s, ref in zip(systems, refs):
        s.finish()
        assert_same(final_state(s), ref)
        s.destroy()


def test_overflow_hook_path_recovers(ctx):
    cfg = configs.scaled_dpd(6000, seed=51)
    smax0, omax0, _ = auto_capacity(ctx, cfg, g.PREC_F64)
    print("case: overflow on the hook path", flush=True)
    out = []
    for cap in (None, (smax0 - 4, omax0 - 4)):
        s = g.system_from_config(ctx, cfg, precision=g.PREC_F64)
        if cap:
            s.set_capacity(*cap)
        s.pair_compute(step=3, log=True)
        f = s.download()["force"]
        pe, vir, _ = s.download_pe_virial()
        out.append((f, pe, vir))
        s.destroy()
    assert_same(out[0], out[1])


def test_overflow_fix_path_recovers(ctx):
    """A system with group fixes and computes (the hook-ordered device path)."""
    cfg = configs.scaled_dpd(5000, seed=61)
    smax0, omax0, _ = auto_capacity(ctx, cfg, g.PREC_F64)
    print("case: overflow on the fix path", flush=True)
    out = []
    for cap in (None, (smax0 - 3, omax0 - 3)):
        s = g.system_from_config(ctx, cfg, precision=g.PREC_F64, log_every=5)
        s.set_fixes([{"kind": "nve", "group": [1]}])
        s.set_computes([{"kind": "temp", "group": [1], "every": 5}])
        if cap:
            s.set_capacity(*cap)
        s.zero_step()
        s.advance(12)
        s.advance(11)
        s.finish()
        out.append(final_state(s) + (s.records(0),))
        s.destroy()
    assert_same(out[0], out[1])


def test_image_counts_beyond_old_range(ctx):
    """Free beads crossing the periodic box ~750 times keep exact image counts
    and unwrapped positions (the reference's uwpos += dt v, fix/fix_nve.cu:83-85)."""
    L = 3.0
    pos = np.array([[0.5, 0.5, 0.5], [1.5, 1.5, 1.5]])
    vel = np.array([[80.0, -61.0, 7.0], [-75.0, 0.0, 33.0]])
    cfg = {"N": 2, "ntypes": 1, "lo": np.zeros(3), "hi": np.full(3, L), "id": np.array([1, 2], np.int32),
           "type": np.ones(2, np.int32), "pos": pos, "vel": vel, "masses": np.array([1.0]), "global_cut": 1.0,
           "skin": 0.3, "max_rho": 6.0, "temp": 0.0, "dt": 0.04, "saru_seed": 42, "style": "dpd", "rc": 1.0,
           "coeff": np.array([0.0, 0.0, 1.0])}
    nsteps = 700
    print("case: image counts over 700 steps", flush=True)
    s = g.system_from_config(ctx, cfg, precision=g.PREC_F64)
    s.run(nsteps)
    d = s.download()
    s.destroy()
    t = nsteps * cfg["dt"]
    unwrapped = d["pos"] + d["image"] * L
    assert np.allclose(unwrapped, pos + vel * t, rtol=0, atol=1e-9)
    assert np.abs(d["image"]).max() > 600  # beyond the old +-511 field
    assert np.array_equal(d["image"], np.floor((pos + vel * t) / L).astype(np.int64))


def test_neigh_check_after_motion(ctx):
    """Neigh_list::check_update after real motion (force/neigh_list.cu:483-551,
    policy main/utils.cuh:217-242): beads displaced just under and just over
    skin/2, one of them across a periodic face, then a rebuild that equals the
    oracle's list of the moved state."""
    base = configs.scaled_dpd(3000, seed=5)
    N, skin, dt = base["N"], base["skin"], base["dt"]
    lo, hi = np.asarray(base["lo"]), np.asarray(base["hi"])
    half = 0.5 * skin
    rng = np.random.Generator(np.random.PCG64(17))
    idx = rng.choice(N, 40, replace=False)
    dirs = rng.normal(size=(40, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    edge = int(np.argmax(base["pos"][:, 0]))  # nearest the upper x face: it crosses it
    for over in (False, True):
        print(f"case: neigh check over={over}", flush=True)
        vel = np.zeros((N, 3))
        vel[idx] = dirs * (half * (1.0 - 1e-7) / dt)
        vel[edge] = [half * (1.0 - 1e-7) / dt, 0.0, 0.0]
        if over:
            vel[idx[0]] = dirs[0] * (half * (1.0 + 1e-7) / dt)
        cfg = dict(base, vel=vel)
        s = g.system_from_config(ctx, cfg, precision=g.PREC_F64)
        last = s.download()["pos"]
        s.neigh_build()
        assert s.neigh_check() == 0
        s.nve_initial()  # no force has been computed: x += dt v
        moved = s.download()["pos"]
        assert moved[edge, 0] < last[edge, 0]  # wrapped across the face
        flag = s.neigh_check()
        assert flag == pyoracle.neigh_check(moved, last, lo, hi, skin)
        assert flag == int(over)
        s.neigh_build()
        n, nb = s.neigh_download(96)
        on, onb = pyoracle.neigh_build(moved, lo, hi, base["global_cut"] + skin, 96)
        assert np.array_equal(n, on)
        assert np.array_equal(nb, onb)
        assert s.neigh_check() == 0
        s.destroy()


def test_upload_refuses_positions_far_outside_the_box(ctx):
    """Uploaded positions are wrapped into the box one box length at a time, as
    the reference does (system/read_data.cu:191-197); one more than 16 box
    lengths out is refused instead of looping for that many box lengths."""
    from gpumdpd_amd import _lib
    cfg = configs.scaled_dpd(2000, seed=71)
    pos = np.array(cfg["pos"], copy=True)
    pos[5, 1] = 1e11
    s = g.system_from_config(ctx, cfg, precision=g.PREC_F32, upload=False)
    print("case: far-out position at upload", flush=True)
    with pytest.raises(_lib.GdpdError) as e:
        s.upload(cfg["id"], cfg["type"], pos, cfg["vel"], cfg["masses"])
    assert "16 box lengths" in str(e.value)
    # three box lengths out is wrapped back as the reference's loop does
    L = np.asarray(cfg["hi"]) - np.asarray(cfg["lo"])
    pos = np.array(cfg["pos"], copy=True)
    pos[5] += 3.0 * L
    s.upload(cfg["id"], cfg["type"], pos, cfg["vel"], cfg["masses"])
    got = s.download()["pos"][5]
    assert np.allclose(got, cfg["pos"][5], atol=1e-12)
    s.destroy()
