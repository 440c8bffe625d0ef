"""Pins of the FP64 oracle's force, list, integrator and thermo arithmetic.

Each check is chosen so a plausible mistake (dropped term, sign, wrong power,
wrong tag order, transposed index) fails: forces are checked as gradients of
their potentials by finite differences, the dissipative and random terms by
their physical invariants (perpendicular motion, fluctuation-dissipation
sigma^2 = 2 gamma kT with the independently restated Saru value), the list by
brute force, the mDPD density by the normalisation of its weight, and the
whole step by energy conservation and the Groot-Warren equation of state.
"""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import pyoracle as O
from saru_numpy import saru_bits, uniform_f64

LIB = O.lib()


def two_beads(r, ids=(1, 2), v=((0, 0, 0), (0, 0, 0)), L=10.0):
    pos = np.array([[5.0 + r, 5.0, 5.0], [5.0, 5.0, 5.0]])
    return {"N": 2, "ntypes": 1, "lo": np.zeros(3), "hi": np.full(3, L), "id": np.array(ids, np.int32),
            "type": np.ones(2, np.int32), "pos": pos, "vel": np.array(v, float), "masses": np.array([1.0]),
            "global_cut": 1.0, "skin": 0.3, "max_rho": 6.0, "temp": 0.0, "dt": 0.04, "saru_seed": 42,
            "style": "dpd", "coeff": np.array([25.0, 0.0, 1.0])}


def dpd_pair(cfg, step=0):
    f, pe, _, vir, _ = O.forces(cfg, step=step)
    return f, pe, vir


def total_pe(cfg):
    return dpd_pair(cfg)[1].sum()


def test_conservative_force_is_gradient_of_pair_energy():
    for r in (0.2, 0.55, 0.93):
        cfg = two_beads(r)
        f, pe, _ = dpd_pair(cfg)
        h = 1e-6
        up, dn = two_beads(r + h), two_beads(r - h)
        fd = -(total_pe(up) - total_pe(dn)) / (2 * h)
        # force on bead 1 (at +r along x) is radial
        assert f[0, 0] == pytest.approx(fd, rel=1e-7, abs=1e-9)
        assert np.allclose(f[0], -f[1]) and abs(f[0, 1]) < 1e-14 and abs(f[0, 2]) < 1e-14
        # energy split: half of 0.5*A*rc*wc^2 to each bead
        assert pe[0] == pytest.approx(pe[1])


def test_no_force_at_or_beyond_cutoff():
    for r in (1.0, 1.0 + 1e-9, 1.2):
        cfg = two_beads(r, v=((0, 0, 0), (1.0, 0, 0)))
        cfg["coeff"] = np.array([25.0, 4.5, 1.0])
        cfg["temp"] = 1.0
        f, pe, _ = dpd_pair(cfg)
        assert np.all(f == 0.0) and np.all(pe == 0.0)


def test_dissipative_term_invariants():
    def fd(r, v2):
        cfg = two_beads(r, v=((0, 0, 0), v2))
        cfg["coeff"] = np.array([0.0, 4.5, 1.0])  # A = 0, T = 0 -> sigma = 0
        return dpd_pair(cfg)[0][0]

    # motion perpendicular to the line of centres is not damped
    assert np.allclose(fd(0.5, (0, 0.7, -0.3)), 0.0, atol=1e-15)
    # bead 2 moves towards bead 1 (approach): bead 1 is pushed away (+x)
    f_a = fd(0.5, (1.0, 0, 0))
    assert f_a[0] > 0
    # linear in the relative velocity, weight (1 - r/rc)^2
    assert fd(0.5, (2.0, 0, 0))[0] == pytest.approx(2 * f_a[0], rel=1e-14)
    ratio = fd(0.25, (1.0, 0, 0))[0] / fd(0.5, (1.0, 0, 0))[0]
    assert ratio == pytest.approx((0.75 / 0.5) ** 2, rel=1e-12)
    # magnitude gamma * w^2 * |v_radial|
    assert f_a[0] == pytest.approx(4.5 * 0.25 * 1.0, rel=1e-12)


def test_random_term_fluctuation_dissipation_and_tags():
    for ids, step in (((1, 2), 0), ((7, 3), 5), ((24000, 11), 1000)):
        cfg = two_beads(0.4, ids=ids)
        cfg["coeff"] = np.array([0.0, 4.5, 1.0])
        cfg["temp"] = 1.3
        f = dpd_pair(cfg, step=step)[0][0]
        lo, hi = min(ids), max(ids)
        xi = uniform_f64(saru_bits(np.array([lo]), np.array([hi]), 42, np.array([step])))[0]
        wr = 1.0 - 0.4
        sigma = f[0] / (wr * xi / math.sqrt(0.04))
        # sigma^2 = 2 gamma kT (fluctuation-dissipation)
        assert sigma ** 2 == pytest.approx(2 * 4.5 * 1.3, rel=1e-12)


def test_newton_third_law_random_system():
    from gpumdpd_amd import configs
    cfg = configs.scaled_dpd(800, seed=9)
    f, pe, _, vir, _ = O.forces(cfg, step=3)
    assert np.abs(f.sum(axis=0)).max() < 1e-10 * np.abs(f).max() * 10
    assert pe.min() >= 0.0


def test_neighbour_list_equals_brute_force_with_edges():
    rng = np.random.Generator(np.random.PCG64(3))
    N, L = 2000, 9.1
    pos = rng.random((N, 3)) * L
    pos[:10, 0] = L - 1e-13   # clamp into the last bin
    pos[10:20, 1] = 0.0
    pos[20] = pos[21] + np.array([1.3, 0, 0])  # a pair exactly at the list cut
    pos[20] %= L
    lo, hi = np.zeros(3), np.full(3, L)
    n1, nb1 = O.neigh_build(pos, lo, hi, 1.3, 128)
    n2, nb2 = O.neigh_build(pos, lo, hi, 1.3, 128, brute=True)
    assert np.array_equal(n1, n2)
    for i in range(N):
        assert np.array_equal(np.sort(nb1[i, :n1[i]]), np.sort(nb2[i, :n2[i]]))
    # the half list keeps j > i only
    for i in range(N):
        assert np.all(nb1[i, :n1[i]] > i)


def test_neigh_check_threshold():
    L = 10.0
    pos = np.array([[1.0, 1.0, 1.0], [5.0, 5.0, 5.0]])
    lo, hi = np.zeros(3), np.full(3, L)
    moved = pos.copy()
    moved[0, 0] += 0.25  # exactly skin/2 for skin = 0.5 (representable) -> rebuild (>=)
    assert O.neigh_check(moved, pos, lo, hi, 0.5) == 1
    moved[0, 0] = pos[0, 0] + 0.2499
    assert O.neigh_check(moved, pos, lo, hi, 0.5) == 0
    # a move across the periodic boundary is a small displacement
    a = np.array([[9.95, 1.0, 1.0], [5.0, 5.0, 5.0]])
    b = np.array([[0.05, 1.0, 1.0], [5.0, 5.0, 5.0]])
    assert O.neigh_check(b, a, lo, hi, 0.3) == 0


def mdpd_cfg(pos, A=0.0, B=25.0, gamma=0.0, T=0.0, L=12.0):
    N = len(pos)
    return {"N": N, "ntypes": 1, "lo": np.zeros(3), "hi": np.full(3, L), "id": np.arange(1, N + 1, dtype=np.int32),
            "type": np.ones(N, np.int32), "pos": np.asarray(pos, float), "vel": np.zeros((N, 3)),
            "masses": np.array([1.0]), "global_cut": 1.0, "skin": 0.3, "max_rho": 10.0, "temp": T, "dt": 0.01,
            "saru_seed": 42, "style": "mdpd", "coeff": np.array([A, B, gamma, 1.0, 0.75])}


def test_mdpd_density_weight_is_normalised():
    # 15/(2 pi rd^3) (1 - r/rd)^2 integrates to 1 over the ball: a uniform fluid reads its density
    from gpumdpd_amd import configs
    cfg = configs.scaled_dpd(30000, rho=3.0, seed=4)
    cfg.update(style="mdpd", coeff=np.array([-40.0, 25.0, 4.5, 1.0, 0.75]), max_rho=10.0)
    rho = O.forces(cfg)[4]
    assert rho.mean() == pytest.approx(3.0, rel=0.01)
    # an isolated pair beyond rd contributes nothing
    r2 = O.forces(mdpd_cfg([[5, 5, 5], [5.8, 5, 5]]))[4]
    assert np.all(r2 == 0.0)


def many_body_energy(cfg):
    """Warren's many-body potential sum_i (pi rd^4 B / 30) rho_i^2 plus the pair term."""
    _, _, _, _, rho = O.forces(cfg)
    A, B, rd, rc = cfg["coeff"][0], cfg["coeff"][1], cfg["coeff"][4], cfg["coeff"][3]
    e = (math.pi * rd ** 4 * B / 30.0) * np.sum(rho ** 2)
    pos = cfg["pos"]
    for i in range(len(pos)):
        for j in range(i + 1, len(pos)):
            r = np.linalg.norm(pos[i] - pos[j])
            if r < rc:
                e += 0.5 * A * rc * (1 - r / rc) ** 2
    return e


def test_mdpd_force_is_gradient_of_many_body_potential():
    rng = np.random.Generator(np.random.PCG64(11))
    pos = 5.0 + rng.random((7, 3)) * 1.1
    for A in (0.0, -40.0):
        cfg = mdpd_cfg(pos, A=A)
        f = O.forces(cfg)[0]
        h = 1e-6
        for d in range(3):
            p1, p2 = pos.copy(), pos.copy()
            p1[2, d] += h
            p2[2, d] -= h
            fd = -(many_body_energy(mdpd_cfg(p1, A=A)) - many_body_energy(mdpd_cfg(p2, A=A))) / (2 * h)
            assert f[2, d] == pytest.approx(fd, rel=1e-6, abs=1e-7)


def test_mdpd_with_b_zero_equals_dpd():
    from gpumdpd_amd import configs
    cfg = configs.scaled_dpd(600, seed=2)
    cfg["temp"] = 1.0
    f_dpd = O.forces(cfg, step=4)[0]
    m = dict(cfg)
    m.update(style="mdpd", coeff=np.array([25.0, 0.0, 4.5, 1.0, 0.75]))
    f_m = O.forces(m, step=4)[0]
    assert np.allclose(f_dpd, f_m, rtol=1e-13, atol=1e-12)


def bond_energy_force(r, k=4.0, r0=0.3):
    b = np.zeros(3)
    pos = np.array([[5.0 + r, 5.0, 5.0], [5.0, 5.0, 5.0]])
    f = np.zeros((2, 3))
    bl = np.array([[1, 1, 2]], np.int32)
    bpe = np.zeros(1)
    bx = O.box(b, np.full(3, 10.0))
    LIB.orc_bond_force(1, O.p(bl, C.c_int), O.p(np.array([k, r0])), O.p(pos), C.byref(bx), O.p(f), O.p(bpe), None)
    return bpe[0], f


def test_bond_force_is_gradient_of_bond_energy():
    for r in (0.1, 0.3, 0.9, 1.7):
        e, f = bond_energy_force(r)
        h = 1e-6
        fd = -(bond_energy_force(r + h)[0] - bond_energy_force(r - h)[0]) / (2 * h)
        assert f[0, 0] == pytest.approx(fd, rel=1e-7, abs=1e-9)
        assert e == pytest.approx(4.0 * (r - 0.3) ** 2)
    # bond_coeff 1 4 0 gives F = -8 r (E = k r^2)
    assert bond_energy_force(0.9, 4.0, 0.0)[1][0, 0] == pytest.approx(-8 * 0.9)


def test_nve_free_particle_wraps_and_unwraps():
    N = 3
    pos = np.array([[9.9, 0.05, 5.0], [0.01, 9.99, 5.0], [5.0, 5.0, 5.0]])
    vel = np.array([[2.0, -3.0, 0.0], [-1.0, 1.0, 0.5], [0.0, 0.0, 0.0]])
    typ = np.ones(N, np.int32)
    uw = pos.copy()
    f = np.zeros((N, 3))
    bx = O.box(np.zeros(3), np.full(3, 10.0))
    x0 = pos.copy()
    for _ in range(25):
        LIB.orc_nve_initial(N, 0.04, O.p(typ, C.c_int), O.p(np.array([1.0])), O.p(f), C.byref(bx), O.p(pos),
                            O.p(uw), O.p(vel))
    expect = x0 + vel * 1.0
    assert np.allclose(uw, expect, atol=1e-12)
    assert np.allclose(pos, np.mod(expect, 10.0), atol=1e-12)
    assert np.all((pos >= 0) & (pos < 10))


def test_conservative_run_conserves_energy():
    from gpumdpd_amd import configs
    cfg = configs.scaled_dpd(500, seed=6)
    cfg.update(coeff=np.array([25.0, 0.0, 1.0]), temp=0.0, dt=0.005)
    rows, _, _ = O.run(cfg, 400, log_every=20)
    e = rows[:, 10] + rows[:, 1]
    assert np.max(np.abs(e - e[0])) / abs(e[0]) < 2e-4


def test_thermo_two_bead_closed_form():
    vel = np.array([[0.5, 0.0, 0.0], [-0.5, 0.0, 0.0]])
    out = np.zeros(11)
    bx = O.box(np.zeros(3), np.array([10.0, 10.0, 10.0]))
    typ = np.ones(2, np.int32)
    vir = np.zeros((2, 6))
    vir[:, 0] = 0.3
    pe = np.array([0.1, 0.2])
    LIB.orc_thermo(2, O.p(typ, C.c_int), O.p(np.array([2.0])), O.p(vel), O.p(pe), O.p(vir), C.byref(bx), 0,
                   None, O.p(out))
    ke = 0.5 * 2.0 * 0.25 * 2
    assert out[10] == pytest.approx(ke)
    assert out[0] == pytest.approx(2 * ke / 3.0)   # 3N - 3 = 3 degrees of freedom
    assert out[1] == pytest.approx(0.3)
    assert out[4] == pytest.approx((2 * 2.0 * 0.25 + 0.6) / 1000.0)
    assert out[3] == pytest.approx(out[4] / 3.0)


def test_run_order_matches_hand_sequenced_steps():
    from gpumdpd_amd import configs
    cfg = configs.scaled_dpd(300, seed=8)
    rows, st, _ = O.run(cfg, 3, log_every=1)
    # hand sequence: force(0); then for step 1..3: initial, rebuild, force(step), final
    N = cfg["N"]
    pos, vel = cfg["pos"].copy(), cfg["vel"].copy()
    typ = cfg["type"].astype(np.int32)
    bx = O.box(cfg["lo"], cfg["hi"])
    f = O.forces(cfg, pos=pos, vel=vel, step=0)[0]
    uw = pos.copy()
    for step in (1, 2, 3):
        LIB.orc_nve_initial(N, cfg["dt"], O.p(typ, C.c_int), O.p(cfg["masses"]), O.p(f), C.byref(bx), O.p(pos),
                            O.p(uw), O.p(vel))
        f = O.forces(cfg, pos=pos, vel=vel, step=step)[0]
        LIB.orc_nve_final(N, cfg["dt"], O.p(typ, C.c_int), O.p(cfg["masses"]), O.p(f), O.p(vel))
    assert np.array_equal(pos, st["pos"]) and np.array_equal(vel, st["vel"])
    assert len(rows) == 4


def test_pressure_mean_field_at_uniform_start_and_thermostat():
    from gpumdpd_amd import configs
    cfg = configs.scaled_dpd(3000, seed=12)
    rows, _, builds = O.run(cfg, 300, log_every=5)
    # uniformly random positions have g(r) = 1, so the virial pressure is the
    # mean-field value rho T + (2 pi / 3) rho^2 * int_0^rc r^3 a (1 - r/rc) dr
    #                     = rho T + (2 pi / 3) rho^2 a / 20  (= 26.56 at rho=3, a=25, T=1)
    T0 = rows[0, 0]
    mf = 3.0 * T0 + (2 * math.pi / 3) * 9.0 * 25.0 / 20.0
    assert rows[0, 3] == pytest.approx(mf, abs=0.4)
    # correlations then lower it; the DPD thermostat holds T near kT = 1
    assert rows[20:, 3].mean() < rows[0, 3] - 1.0
    assert rows[20:, 0].mean() == pytest.approx(1.0, abs=0.06)
    # the list rebuilds almost every step at dt=0.04, skin=0.3
    assert builds > 250


# ---- harmonic angles (force/angle_harmonic.cu:168-330) ---------------------------

def _angle_energy(al, coeff, pos, lo, hi):
    return float(O.angle_force(al, coeff, pos, lo, hi)[1].sum())


def test_angle_closed_form_right_angle():
    # i on x, vertex at the origin, k on y: theta = 90 deg; theta0 = 60 deg
    lo, hi = [-10.0] * 3, [10.0] * 3
    pos = np.array([[1.0, 0.0, 0.0], [0.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
    k = 3.0
    f, ape, vir = O.angle_force([[1, 1, 2, 3]], [k, 60.0], pos, lo, hi)
    dth = np.pi / 2 - np.pi / 3
    assert ape[0] == pytest.approx(k * dth * dth, rel=1e-14)
    # -dE/dr pulls i toward k and k toward i
    assert np.allclose(f[0], [0.0, 2 * k * dth, 0.0], atol=1e-13)
    assert np.allclose(f[2], [2 * k * dth, 0.0, 0.0], atol=1e-13)
    assert np.allclose(f[1], -(f[0] + f[2]), atol=1e-13)


def test_angle_force_is_gradient_of_angle_energy():
    rng = np.random.default_rng(11)
    lo, hi = [0.0] * 3, [8.0] * 3
    n = 12
    pos = np.cumsum(rng.normal(0.0, 0.6, (n, 3)), axis=0) % 8.0
    al = np.array([[1 + (a % 2), a + 1, a + 2, a + 3] for a in range(n - 2)], dtype=np.int32)
    coeff = [4.0, 120.0, 2.5, 75.0]
    f, _, _ = O.angle_force(al, coeff, pos, lo, hi)
    h = 1e-6
    for i in range(n):
        for d in range(3):
            pp, pm = pos.copy(), pos.copy()
            pp[i, d] += h
            pm[i, d] -= h
            g = (_angle_energy(al, coeff, pp, lo, hi) - _angle_energy(al, coeff, pm, lo, hi)) / (2 * h)
            assert f[i, d] == pytest.approx(-g, rel=1e-5, abs=1e-6)


def test_angle_forces_have_no_net_force_or_torque_and_vanish_at_theta0():
    rng = np.random.default_rng(12)
    lo, hi = [-50.0] * 3, [50.0] * 3
    pos = rng.normal(0.0, 1.0, (3, 3))
    f, _, vir = O.angle_force([[1, 1, 2, 3]], [5.0, 100.0], pos, lo, hi)
    assert np.allclose(f.sum(axis=0), 0.0, atol=1e-12)
    assert np.allclose(np.cross(pos, f).sum(axis=0), 0.0, atol=1e-11)
    # the three thirds of the virial are equal
    assert np.allclose(vir[0], vir[1]) and np.allclose(vir[1], vir[2])
    d1, d2 = pos[0] - pos[1], pos[2] - pos[1]
    th = np.degrees(np.arccos(d1 @ d2 / np.linalg.norm(d1) / np.linalg.norm(d2)))
    f0, ape0, _ = O.angle_force([[1, 1, 2, 3]], [5.0, th], pos, lo, hi)
    assert np.allclose(f0, 0.0, atol=1e-9) and ape0[0] < 1e-18
