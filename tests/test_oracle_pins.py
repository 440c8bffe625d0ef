"""Independent pins of two oracle functions that a restatement alone cannot fix.

1. The mDPD per-bead energy split (force/pair_mdpd.cu:479-485):

       pe_i = sum_j [ 1/2 u_c(r_ij) + 1/4 B rd rho_i wd(r_ij)^2 ],   u_c = 1/2 A rc wc^2.

   Pass 1 gives rho_i = sum_j 15/(2 pi rd^3) wd(r_ij)^2 over the same pairs
   (pair_mdpd.cu:266-289; every pair inside rd is inside rc when rd <= rc), so
   the many-body share of bead i is exactly (pi rd^4 / 30) B rho_i^2: Warren's
   density energy psi(rho_i) of that bead.  Physics therefore fixes the split:
   each bead carries half of each pair energy plus its own psi(rho_i), and the
   sum over beads is the potential whose negative gradient is the conservative
   force.  The expected values below are computed from geometry (hand-worked
   distances on a line, brute-force pair distances, finite differences), never
   through the oracle's force routine.  A rho_i <-> rho_j swap, a 1/4 -> 1/2
   slip or a dropped pair half each fails one of them.

2. Saru's TEA stage (force/pair_dpd.cu:227-232) is the block cipher of Wheeler
   and Needham (TEA, Fast Software Encryption 1994) run for 8 cycles with the
   key (k0, k1, k2, k3) of pair_dpd.cu:205-208.  `tea_encrypt` below is written
   from the paper's description (v0 += ((v1 << 4) + k0) ^ (v1 + sum) ^
   ((v1 >> 5) + k1); v1 += ((v0 << 4) + k2) ^ (v0 + sum) ^ ((v0 >> 5) + k3);
   sum += delta before each cycle) and reproduces the published 32-cycle known
   answer for an all-zero key and block.  The oracle's TEA stage must equal it
   on the oracle's pre-mixed blocks.  The pre-mix (pair_dpd.cu:213-226) and the
   output avalanche (:233-234) have no second published source; they stay
   pinned by the two transcriptions of tests/test_oracle_saru.py.
"""
import math

import numpy as np
import pytest

from oracle import pyoracle as O

RD, RC = 0.75, 1.0


def mdpd_cfg(pos, A=-40.0, B=25.0, L=12.0):
    N = len(pos)
    return {"N": N, "ntypes": 1, "lo": np.zeros(3), "hi": np.full(3, L), "id": np.arange(1, N + 1, dtype=np.int32),
            "type": np.ones(N, np.int32), "pos": np.asarray(pos, float), "vel": np.zeros((N, 3)),
            "masses": np.array([1.0]), "global_cut": RC, "skin": 0.3, "max_rho": 30.0, "temp": 0.0, "dt": 0.01,
            "saru_seed": 42, "style": "mdpd", "coeff": np.array([A, B, 0.0, RC, RD])}


def test_mdpd_pe_split_hand_worked_line():
    # Four beads on a line: 1 at 0, 2 at 0.25, 3 at 0.5, 4 at 1.4.
    # Pairs inside rc = 1: (1,2) r=0.25, (1,3) r=0.5, (2,3) r=0.25, (3,4) r=0.9;
    # (3,4) lies beyond rd = 0.75, so it adds pair energy but no density.
    A, B = -40.0, 25.0
    x0 = np.array([5.0, 5.0, 5.0])
    pos = [x0, x0 + [0.25, 0, 0], x0 + [0.5, 0, 0], x0 + [1.4, 0, 0]]
    _, pe, _, _, rho = O.forces(mdpd_cfg(pos, A=A, B=B))
    # weights: wd(0.25) = 2/3, wd(0.5) = 1/3, wd(0.9) = 0; wc(0.25) = 3/4, wc(0.5) = 1/2, wc(0.9) = 1/10
    c = 15.0 / (2.0 * math.pi * RD ** 3)
    rho_ref = c * np.array([4 / 9 + 1 / 9, 4 / 9 + 4 / 9, 1 / 9 + 4 / 9, 0.0])
    assert np.allclose(rho, rho_ref, rtol=1e-14, atol=0)
    assert rho[1] != rho[0]  # the densities of a pair's two beads differ
    uc = {0.25: 0.5 * A * RC * (3 / 4) ** 2, 0.5: 0.5 * A * RC * (1 / 2) ** 2, 0.9: 0.5 * A * RC * (1 / 10) ** 2}
    k = 0.25 * B * RD
    pe_ref = np.array([
        0.5 * (uc[0.25] + uc[0.5]) + k * rho_ref[0] * (4 / 9 + 1 / 9),
        0.5 * (uc[0.25] + uc[0.25]) + k * rho_ref[1] * (4 / 9 + 4 / 9),
        0.5 * (uc[0.5] + uc[0.25] + uc[0.9]) + k * rho_ref[2] * (1 / 9 + 4 / 9),
        0.5 * uc[0.9],
    ])
    assert np.allclose(pe, pe_ref, rtol=1e-13, atol=1e-13)
    # the numbers that a rho_i <-> rho_j swap or a 1/4 -> 1/2 slip would give differ by far more
    swapped0 = 0.5 * (uc[0.25] + uc[0.5]) + k * (rho_ref[1] * 4 / 9 + rho_ref[2] * 1 / 9)
    assert abs(swapped0 - pe[0]) > 1.0
    doubled0 = 0.5 * (uc[0.25] + uc[0.5]) + 2 * k * rho_ref[0] * (5 / 9)
    assert abs(doubled0 - pe[0]) > 1.0
    assert pe.sum() == pytest.approx(sum(uc[r] for r in (0.25, 0.5, 0.25, 0.9))
                                     + math.pi * RD ** 4 / 30 * B * np.sum(rho_ref ** 2), rel=1e-13)


def _pair_table(pos, L):
    d = pos[:, None, :] - pos[None, :, :]
    d -= L * np.round(d / L)
    r = np.sqrt((d ** 2).sum(-1))
    np.fill_diagonal(r, np.inf)
    return r


def test_mdpd_pe_is_half_pairs_plus_own_density_energy():
    # a dense, uneven cluster (a droplet-like lump plus a sparse halo), so
    # neighbouring beads carry very different densities
    rng = np.random.Generator(np.random.PCG64(31))
    L = 8.0
    core = 4.0 + rng.normal(0.0, 0.45, (160, 3))
    halo = rng.random((60, 3)) * L
    pos = np.concatenate([core, halo]) % L
    A, B = -40.0, 25.0
    _, pe, _, _, rho = O.forces(mdpd_cfg(pos, A=A, B=B, L=L))
    r = _pair_table(pos, L)
    wd = np.clip(1.0 - r / RD, 0.0, None)
    rho_ref = (15.0 / (2.0 * math.pi * RD ** 3) * wd ** 2).sum(1)
    assert np.allclose(rho, rho_ref, rtol=1e-12, atol=1e-12)
    assert rho.std() > 0.5 * rho.mean() and rho.max() > 50 * rho[rho > 0].min()  # genuinely uneven
    wc = np.where(r <= RC, 1.0 - r / RC, 0.0)
    half_pairs = 0.5 * (0.5 * A * RC * wc ** 2).sum(1)
    psi = math.pi * RD ** 4 / 30.0 * B * rho_ref ** 2
    assert np.allclose(pe, half_pairs + psi, rtol=1e-11, atol=1e-10)


def test_mdpd_logged_pe_is_the_potential_of_the_force():
    # gamma = 0 and T = 0 leave the conservative force; -d(sum pe)/dx must equal it
    rng = np.random.Generator(np.random.PCG64(5))
    pos0 = 5.0 + rng.random((9, 3)) * 1.2
    cfg = mdpd_cfg(pos0)
    f = O.forces(cfg)[0]
    h = 1e-6
    for bead in (0, 4, 7):
        for d in range(3):
            p1, p2 = pos0.copy(), pos0.copy()
            p1[bead, d] += h
            p2[bead, d] -= h
            e1 = O.forces(mdpd_cfg(p1))[1].sum()
            e2 = O.forces(mdpd_cfg(p2))[1].sum()
            assert f[bead, d] == pytest.approx(-(e1 - e2) / (2 * h), rel=1e-6, abs=1e-7)


# --------------------------------------------------------------------------
# Saru: TEA stage against the textbook cipher
# --------------------------------------------------------------------------

M32 = 0xFFFFFFFF


def tea_encrypt(v0, v1, key, cycles=32, delta=0x9E3779B9):
    """TEA encryption of the 64-bit block (v0, v1), Wheeler and Needham 1994."""
    s = 0
    for _ in range(cycles):
        s = (s + delta) & M32
        v0 = (v0 + ((((v1 << 4) & M32) + key[0]) & M32 ^ (v1 + s) & M32 ^ ((v1 >> 5) + key[1]) & M32)) & M32
        v1 = (v1 + ((((v0 << 4) & M32) + key[2]) & M32 ^ (v0 + s) & M32 ^ ((v0 >> 5) + key[3]) & M32)) & M32
    return v0, v1


SARU_KEY = (0xA341316C, 0xC8013EA4, 0xAD90777D, 0x7E95761E)


def test_textbook_tea_known_answer():
    # the published test vector: zero key, zero block, 32 cycles
    assert tea_encrypt(0, 0, (0, 0, 0, 0)) == (0x41EA3A0A, 0x94BAA940)


def test_saru_tea_stage_is_textbook_tea():
    rng = np.random.Generator(np.random.PCG64(2024))
    n = 3000
    t1 = rng.integers(0, 2 ** 32, n, dtype=np.uint64)
    t2 = rng.integers(0, 2 ** 32, n, dtype=np.uint64)
    st = rng.integers(-(2 ** 31), 2 ** 31, n, dtype=np.int64)
    for a, b, step in zip(t1, t2, st):
        st0, ws0, st8, ws8, out = O.saru_stages(int(a), int(b), 42, int(step))
        assert (st8, ws8) == tea_encrypt(st0, ws0, SARU_KEY, cycles=8)
        assert out == int(O.lib().orc_saru_u32(int(a), int(b), 42, int(step)))
    # call-site keys: tags sorted, seed + step (pair_dpd.cu:213, 342-344)
    st0, ws0, st8, ws8, out = O.saru_stages(3, 17, 42, 5)
    assert (st8, ws8) == tea_encrypt(st0, ws0, SARU_KEY, cycles=8)
