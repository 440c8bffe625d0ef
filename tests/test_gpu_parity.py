"""GPU parity of the HIP path against the FP64 oracle, through the C ABI.

Tolerances (from the build's precision contract):
  FP32 pair arithmetic / FP64 accumulation: global relative L2 force error <= 1e-5
  FP64 mode: <= 1e-11
"""
import numpy as np
import pytest

import gpumdpd_amd as g
from gpumdpd_amd import configs
from oracle import pyoracle

pytestmark = pytest.mark.gpu

F32_TOL = 1e-5
F64_TOL = 1e-11


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.fixture(scope="module")
def ctx():
    c = g.Context(0)
    yield c
    c.close()


def test_gpu_saru_bits_match_golden(ctx):
    import os
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "saru_table.npz"))
    for sd in np.unique(d["seed"]):
        m = d["seed"] == sd
        bits, r32, r64 = ctx.saru_eval(d["tag1"][m], d["tag2"][m], int(sd), d["step"][m])
        assert np.array_equal(bits, d["bits"][m])
        assert np.array_equal(r32.view(np.uint32), d["r32"][m].view(np.uint32))
        assert np.array_equal(r64.view(np.uint64), d["r64"][m].view(np.uint64))


@pytest.mark.parametrize("prec,tol", [(g.PREC_F32, F32_TOL), (g.PREC_F64, F64_TOL)])
@pytest.mark.parametrize("step", [0, 17])
def test_dpd_force_c0(ctx, prec, tol, step):
    cfg = configs.config_c0()
    print(f"case: C0 prec={prec} step={step}", flush=True)
    s = g.system_from_config(ctx, cfg, precision=prec)
    s.pair_compute(step=step, log=True)
    f = s.download()["force"]
    pe, vir, _ = s.download_pe_virial()
    f_ref, pe_ref, _, vir_ref, _ = pyoracle.forces(cfg, step=step)
    assert rel(f, f_ref) < tol
    assert rel(pe, pe_ref) < tol
    assert rel(vir, vir_ref) < tol
    s.destroy()


@pytest.mark.parametrize("prec,tol", [(g.PREC_F32, F32_TOL), (g.PREC_F64, F64_TOL)])
def test_mdpd_density_and_force_c1(ctx, prec, tol):
    cfg = configs.config_c1()
    print(f"case: C1 prec={prec}", flush=True)
    s = g.system_from_config(ctx, cfg, precision=prec)
    s.pair_compute(step=3, log=True)
    f = s.download()["force"]
    pe, vir, rho = s.download_pe_virial()
    f_ref, pe_ref, _, vir_ref, rho_ref = pyoracle.forces(cfg, step=3)
    assert rel(rho, rho_ref) < tol
    assert rel(f, f_ref) < tol
    assert rel(pe, pe_ref) < tol
    s.destroy()


def test_bonded_melt_c2_full(ctx):
    """The whole C2 diblock melt (374,992 beads, 351,555 bonds): pair + bond
    force, per-bead pair PE and virial against the FP64 oracle."""
    cfg = configs.config_c2()
    print(f"case: C2 full, N={cfg['N']}", flush=True)
    f_ref, pe_ref, _, vir_ref, _ = pyoracle.forces(cfg, step=5)
    for prec, tol in [(g.PREC_F32, F32_TOL), (g.PREC_F64, F64_TOL)]:
        s = g.system_from_config(ctx, cfg, precision=prec)
        s.pair_compute(step=5, log=True)
        s.bond_compute(step=5, log=True)
        f = s.download()["force"]
        pe, vir, _ = s.download_pe_virial()
        assert rel(f, f_ref) < tol, prec
        assert rel(pe, pe_ref) < tol, prec
        assert rel(vir, vir_ref) < tol, prec
        s.destroy()


def test_fused_run_matches_oracle_fp64(ctx):
    cfg = configs.scaled_dpd(4000, seed=3)
    print("case: fused 30-step FP64 run", flush=True)
    s = g.system_from_config(ctx, cfg, precision=g.PREC_F64, log_every=10)
    s.run(30)
    st = s.download()
    rows = s.thermo_rows()
    ref_rows, ref_state, _ = pyoracle.run(cfg, 30, log_every=10)
    L = np.asarray(cfg["hi"]) - np.asarray(cfg["lo"])
    d = st["pos"] - ref_state["pos"]
    d -= L * np.round(d / L)
    assert np.max(np.abs(d)) < 1e-9
    assert np.max(np.abs(st["vel"] - ref_state["vel"])) < 1e-8
    assert rel(st["force"], ref_state["force"]) < 1e-9
    assert len(rows) == len(ref_rows) == 4
    assert np.allclose(rows[:, :11], ref_rows, rtol=1e-9, atol=1e-9)
    s.destroy()


def test_reference_neighbour_list(ctx):
    cfg = configs.scaled_dpd(3000, seed=5)
    s = g.system_from_config(ctx, cfg, precision=g.PREC_F64)
    s.neigh_build()
    n_max = 96
    n, nb = s.neigh_download(n_max)
    on, onb = pyoracle.neigh_build(cfg["pos"], cfg["lo"], cfg["hi"], 1.3, n_max)
    assert np.array_equal(n, on)
    assert np.array_equal(np.where(nb >= 0, nb, -1), np.where(onb >= 0, onb, -1))
    assert s.neigh_check() == 0
    s.destroy()


_ORACLE_CACHE = {}


def oracle_forces(name, cfg, step):
    key = (name, step)
    if key not in _ORACLE_CACHE:
        _ORACLE_CACHE[key] = pyoracle.forces(cfg, step=step)
    return _ORACLE_CACHE[key]


@pytest.mark.parametrize("brick", [(4, 4, 1), (1, 1, 1), (3, 3, 3), (2, 3, 7), (6, 1, 2)])
def test_brick_shapes_match_oracle(ctx, brick):
    """Every brick shape, including one-cell-thick bricks whose z-ends are both
    halo, finds each pair exactly once."""
    cfg = configs.config_c0()
    f_ref, pe_ref, _, vir_ref, _ = oracle_forces("c0", cfg, 9)
    for prec, tol in [(g.PREC_F32, F32_TOL), (g.PREC_F64, F64_TOL)]:
        print(f"case: C0 brick={brick} prec={prec}", flush=True)
        s = g.system_from_config(ctx, cfg, precision=prec, upload=False)
        s.set_tuning(brick=brick)
        s.upload(cfg["id"], cfg["type"], cfg["pos"], cfg["vel"], cfg["masses"])
        t = s.tuning()
        assert (t["bx"], t["by"], t["bz"]) == brick
        s.pair_compute(step=9, log=True)
        f = s.download()["force"]
        pe, vir, _ = s.download_pe_virial()
        assert rel(f, f_ref) < tol
        assert rel(pe, pe_ref) < tol
        assert rel(vir, vir_ref) < tol
        s.destroy()


@pytest.mark.parametrize("brick", [(4, 4, 1), (2, 2, 2)])
def test_mdpd_brick_shapes_match_oracle(ctx, brick):
    cfg = configs.config_c1()
    f_ref, pe_ref, _, vir_ref, rho_ref = oracle_forces("c1", cfg, 4)
    print(f"case: C1 brick={brick}", flush=True)
    s = g.system_from_config(ctx, cfg, precision=g.PREC_F64, upload=False)
    s.set_tuning(brick=brick)
    s.upload(cfg["id"], cfg["type"], cfg["pos"], cfg["vel"], cfg["masses"])
    s.pair_compute(step=4, log=True)
    f = s.download()["force"]
    pe, vir, rho = s.download_pe_virial()
    assert rel(rho, rho_ref) < F64_TOL
    assert rel(f, f_ref) < F64_TOL
    assert rel(pe, pe_ref) < F64_TOL
    s.destroy()


@pytest.mark.parametrize("prec,tol", [(g.PREC_F32, F32_TOL), (g.PREC_F64, F64_TOL)])
def test_angles_match_oracle(ctx, prec, tol):
    """Harmonic angles (force/angle_harmonic.cu:168-330) on a semi-flexible melt:
    pair + bond + angle force and virial, and the angle energy of the hook
    thermo and of the fused run's step-0 record."""
    from oracle import pyoracle as O
    cfg = configs.config_angles(n_chains=700, box=15.0)
    print(f"case: angle melt prec={prec}", flush=True)
    s = g.system_from_config(ctx, cfg, precision=prec)
    s.pair_compute(step=2, log=True)
    s.bond_compute(step=2, log=True)
    s.angle_compute(step=2, log=True)
    f = s.download()["force"]
    _, vir, _ = s.download_pe_virial()
    th = s.thermo()
    s.destroy()
    f_ref, _, _, vir_ref, _ = pyoracle.forces(cfg, step=2)
    order = np.argsort(cfg["id"])
    _, ape, _ = O.angle_force(cfg["angles"], cfg["angle_coeff"], np.asarray(cfg["pos"])[order], cfg["lo"], cfg["hi"])
    assert rel(f, f_ref) < tol
    assert rel(vir, vir_ref) < tol
    assert th[12] == pytest.approx(ape.sum(), rel=max(tol, 1e-12))
    r = g.system_from_config(ctx, cfg, precision=prec, log_every=1)
    r.run(0)
    row = r.thermo_rows()[0]
    r.destroy()
    assert row[12] == pytest.approx(ape.sum(), rel=max(tol, 1e-12))


def _tiny(n, box, seed, style="dpd"):
    """A sparse or ragged DPD/mDPD system in a box of 3-12 cells per axis.  Boxes
    under 3 reference bins of rc + skin get skin 0 (the reference then rebuilds
    its list every step), so the oracle's 27-bin stencil never repeats a bin."""
    cfg = configs.dpd_fluid(n, box, seed)
    skin = 0.3 if min(box) >= 3.9 else 0.0
    if style == "dpd":
        cfg.update(style="dpd", temp=1.0, rc=1.0, saru_seed=42, dt=0.04, skin=skin, global_cut=1.0,
                   max_rho=6.0, coeff=configs.dpd_coeff(1, {(1, 1): (25.0, 4.5, 1.0)}), name=f"tiny {n}")
    else:
        cfg.update(style="mdpd", temp=1.0, rc=1.0, rd=0.75, saru_seed=42, dt=0.01, skin=skin, global_cut=1.0,
                   max_rho=8.0, coeff=configs.mdpd_coeff(1, {(1, 1): (-40.0, 25.0, 4.5, 1.0, 0.75)}),
                   name=f"tiny mdpd {n}")
    return cfg


TINY = [
    # (n, box, style): one bead; a pair across all three periodic faces; sparse
    # and ragged boxes whose bricks' halos wrap onto their own cells; a dense
    # 3x3x3-cell box; a ragged box of several bricks; mDPD in a small box
    (1, (3.0, 3.0, 3.0), "dpd"),
    (2, (3.0, 3.0, 3.0), "dpd"),
    (7, (3.2, 3.5, 3.9), "dpd"),
    (60, (3.0, 4.7, 3.3), "dpd"),
    (81, (3.0, 3.0, 3.0), "dpd"),
    (400, (3.7, 11.3, 5.2), "dpd"),
    (150, (3.4, 3.9, 4.6), "mdpd"),
]


@pytest.mark.parametrize("n,box,style", TINY)
def test_tiny_and_ragged_boxes(ctx, n, box, style):
    cfg = _tiny(n, box, seed=100 + n, style=style)
    if n == 2:
        # in range only through the periodic images on every axis (r = 0.6)
        cfg["pos"] = np.array([[0.2, 0.3, 2.9], [2.8, 0.1, 0.3]])
    print(f"case: tiny n={n} box={box} style={style}", flush=True)
    for prec, tol in ((g.PREC_F32, F32_TOL), (g.PREC_F64, F64_TOL)):
        s = g.system_from_config(ctx, cfg, precision=prec)
        s.pair_compute(step=4, log=True)
        f = s.download()["force"]
        pe, vir, _ = s.download_pe_virial()
        f_ref, pe_ref, _, vir_ref, _ = pyoracle.forces(cfg, step=4)
        s.destroy()
        assert rel(f, f_ref) < tol and rel(pe, pe_ref) < tol and rel(vir, vir_ref) < tol
        if n == 1:
            assert np.all(f == 0.0) and np.all(pe == 0.0)
        if n == 2:
            assert np.linalg.norm(f_ref[0]) > 1.0 and np.allclose(f_ref[0], -f_ref[1])
    s = g.system_from_config(ctx, cfg, precision=g.PREC_F64, log_every=5)
    s.run(20)
    st = s.download()
    rows = s.thermo_rows()
    s.destroy()
    ref_rows, ref_state, _ = pyoracle.run(cfg, 20, log_every=5)
    L = np.asarray(cfg["hi"]) - np.asarray(cfg["lo"])
    d = st["pos"] - ref_state["pos"]
    d -= L * np.round(d / L)
    assert np.max(np.abs(d)) < 1e-9
    assert np.max(np.abs(st["vel"] - ref_state["vel"])) < 1e-8
    assert len(rows) == len(ref_rows) == 5
    # one bead has no degrees of freedom left after the momentum term: T is inf in both
    assert np.allclose(rows[:, :11], ref_rows, rtol=1e-9, atol=1e-9, equal_nan=True)


def _mixed(case):
    """Multi-type systems whose pair cutoffs differ per type pair
    (pair_coeff i j A gamma rc, force/pair_dpd.cu:118-160; mDPD adds B and rd,
    pair_mdpd.cu:130-175): the search runs to the largest cutoff and each pair
    is cut at its own rc (and weighted with its own rd)."""
    if case == "dpd3":
        box, rho, nt, style = (9.0, 8.5, 9.5), 3.0, 3, "dpd"
        table = {(1, 1): (25.0, 4.5, 1.0), (1, 2): (40.0, 4.5, 0.8), (1, 3): (30.0, 3.0, 0.6),
                 (2, 2): (25.0, 4.5, 0.9), (2, 3): (35.0, 2.0, 1.0), (3, 3): (20.0, 5.0, 0.7)}
        masses = np.array([1.0, 1.5, 0.8])
    elif case == "dpd_short":
        box, rho, nt, style = (8.0, 8.0, 8.0), 3.0, 2, "dpd"
        table = {(1, 1): (25.0, 4.5, 0.9), (1, 2): (30.0, 4.5, 0.75), (2, 2): (25.0, 4.5, 0.85)}
        masses = np.array([1.0, 1.0])
    else:
        box, rho, nt, style = (7.0, 7.0, 7.0), 6.0, 2, "mdpd"
        table = {(1, 1): (-40.0, 25.0, 4.5, 1.0, 0.75), (1, 2): (-30.0, 20.0, 4.5, 0.85, 0.65),
                 (2, 2): (-35.0, 25.0, 3.0, 0.9, 0.7)}
        masses = np.array([1.0, 1.2])
    n = int(rho * np.prod(box))
    cfg = configs.dpd_fluid(n, box, seed=900 + n)
    typ = 1 + (configs.uniform(configs.substream(901, 2), n) * nt).astype(np.int32)
    cfg.update(ntypes=nt, type=np.minimum(typ, nt).astype(np.int32), masses=masses, style=style, temp=1.0,
               rc=1.0, saru_seed=42, skin=0.3, global_cut=1.0, name=f"mixed cutoffs {case}")
    if style == "dpd":
        cfg.update(dt=0.04, max_rho=6.0, coeff=configs.dpd_coeff(nt, table))
    else:
        cfg.update(dt=0.01, rd=0.75, max_rho=10.0, coeff=configs.mdpd_coeff(nt, table))
    return cfg


@pytest.mark.parametrize("case", ["dpd3", "dpd_short", "mdpd2"])
def test_mixed_cutoffs_match_oracle(ctx, case):
    cfg = _mixed(case)
    print(f"case: mixed cutoffs {case}, N={cfg['N']}", flush=True)
    f_ref, pe_ref, _, vir_ref, rho_ref = pyoracle.forces(cfg, step=4)
    # the per-pair cutoffs matter: with every rc at 1.0 (rd 0.75) the force is far off
    w = 3 if cfg["style"] == "dpd" else 5
    cu = np.asarray(cfg["coeff"]).reshape(-1, w).copy()
    cu[:, w - 1 if w == 3 else 3] = 1.0
    if w == 5:
        cu[:, 4] = 0.75
    assert rel(pyoracle.forces(dict(cfg, coeff=cu.ravel()), step=4)[0], f_ref) > 0.1
    for prec, tol in ((g.PREC_F32, F32_TOL), (g.PREC_F64, F64_TOL)):
        s = g.system_from_config(ctx, cfg, precision=prec)
        s.pair_compute(step=4, log=True)
        f = s.download()["force"]
        pe, vir, rho = s.download_pe_virial()
        s.destroy()
        assert rel(f, f_ref) < tol, prec
        assert rel(pe, pe_ref) < tol and rel(vir, vir_ref) < tol, prec
        if cfg["style"] == "mdpd":
            assert rel(rho, rho_ref) < tol, prec
    s = g.system_from_config(ctx, cfg, precision=g.PREC_F64, log_every=5)
    s.run(20)
    st = s.download()
    rows = s.thermo_rows()
    s.destroy()
    ref_rows, ref_state, _ = pyoracle.run(cfg, 20, log_every=5)
    L = np.asarray(cfg["hi"]) - np.asarray(cfg["lo"])
    d = st["pos"] - ref_state["pos"]
    d -= L * np.round(d / L)
    assert np.max(np.abs(d)) < 1e-9
    assert np.max(np.abs(st["vel"] - ref_state["vel"])) < 1e-8
    assert np.allclose(rows[:, :11], ref_rows, rtol=1e-9, atol=1e-9)


def test_coincident_beads(ctx):
    """Two beads at one point add nothing to each other (force/pair_dpd.cu:
    pairs with r < 1e-10 are skipped); a third bead 0.5 away feels both."""
    cfg = _tiny(3, (4.0, 4.0, 4.0), seed=7)
    cfg["pos"] = np.array([[1.5, 2.0, 2.0], [1.5, 2.0, 2.0], [2.0, 2.0, 2.0]])
    cfg["vel"] = np.array([[0.3, -0.2, 0.1], [-0.1, 0.4, 0.0], [0.0, 0.0, -0.5]])
    print("case: coincident beads", flush=True)
    f_ref, pe_ref, _, _, _ = pyoracle.forces(cfg, step=3)
    assert np.linalg.norm(f_ref[2]) > 1.0
    for prec, tol in ((g.PREC_F32, F32_TOL), (g.PREC_F64, F64_TOL)):
        s = g.system_from_config(ctx, cfg, precision=prec)
        s.pair_compute(step=3, log=True)
        f = s.download()["force"]
        pe, _, _ = s.download_pe_virial()
        s.destroy()
        assert np.all(np.isfinite(f)) and rel(f, f_ref) < tol and rel(pe, pe_ref) < tol
