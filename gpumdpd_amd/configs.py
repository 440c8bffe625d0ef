"""Deterministic synthetic inputs for the benchmark systems.

The reference reads LAMMPS data files and has no generator, so inputs are made
here with a counter-based splitmix64 stream (uniforms in [0,1) from the top 53
bits) and Box-Muller normals: the same seed gives the same system on any host.

    C0  DPD fluid, rho=3, 20^3 box, 24,000 beads, seed 1234
    C1  mDPD liquid slab, rho=6 in the centre third of 40x10x10, 8,000 beads, seed 7
    C2  AB diblock melt, 23,437 chains x 16 beads (8 A + 8 B), 50^3, seed 99
    C3  DPD fluid on a 216^3 jittered lattice in a 150^3 box, seed 2024
    C4  served stream: 512 requests, Poisson arrivals at 200/s, master seed 5
"""
import math

import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed, n, offset=0):
    """n consecutive outputs of the splitmix64 stream started at `seed`."""
    with np.errstate(over="ignore"):
        x = np.uint64(seed) + GOLDEN * (np.arange(offset, offset + n, dtype=np.uint64) + np.uint64(1))
        z = (x ^ (x >> np.uint64(30))) * M1
        z = (z ^ (z >> np.uint64(27))) * M2
        return z ^ (z >> np.uint64(31))


def substream(seed, tag):
    """A child seed for purpose `tag` (positions, velocities, ...)."""
    return int(splitmix64(int(seed) ^ (0x5851F42D4C957F2D * (tag + 1) & 0xFFFFFFFFFFFFFFFF), 1)[0])


def uniform(seed, n):
    return (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed, n):
    m = (n + 1) // 2
    u = uniform(seed, 2 * m)
    u1 = 1.0 - u[:m]  # (0, 1]
    u2 = u[m:]
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.concatenate([r * np.cos(2.0 * math.pi * u2), r * np.sin(2.0 * math.pi * u2)])
    return z[:n]


def maxwell(seed, n, kT=1.0, mass=1.0):
    return normal(seed, 3 * n).reshape(n, 3) * math.sqrt(kT / mass)


def dpd_coeff(ntypes, table):
    """Packed {A, gamma, rc} per type pair from {(ti, tj): (A, gamma, rc)}."""
    out = []
    for ti in range(1, ntypes + 1):
        for tj in range(ti, ntypes + 1):
            out.extend(table[(ti, tj)])
    return np.array(out, dtype=np.float64)


def mdpd_coeff(ntypes, table):
    out = []
    for ti in range(1, ntypes + 1):
        for tj in range(ti, ntypes + 1):
            out.extend(table[(ti, tj)])
    return np.array(out, dtype=np.float64)


def dpd_fluid(n, box, seed, rho=None):
    """Uniform random positions in a cubic/rectangular box and MB velocities."""
    box = np.asarray(box, dtype=np.float64)
    pos = uniform(substream(seed, 0), 3 * n).reshape(n, 3) * box
    vel = maxwell(substream(seed, 1), n)
    return {
        "N": n, "ntypes": 1, "lo": np.zeros(3), "hi": box.copy(),
        "id": np.arange(1, n + 1, dtype=np.int32), "type": np.ones(n, dtype=np.int32),
        "pos": pos, "vel": vel, "masses": np.array([1.0]),
    }


def config_c0(seed=1234):
    s = dpd_fluid(24000, (20.0, 20.0, 20.0), seed)
    s.update(style="dpd", temp=1.0, rc=1.0, saru_seed=42, dt=0.04, skin=0.3, global_cut=1.0,
             max_rho=6.0, coeff=dpd_coeff(1, {(1, 1): (25.0, 4.5, 1.0)}), name="C0 DPD fluid 24k")
    return s


def config_c1(seed=7):
    n, L = 8000, np.array([40.0, 10.0, 10.0])
    u = uniform(substream(seed, 0), 3 * n).reshape(n, 3)
    pos = np.empty((n, 3))
    pos[:, 0] = L[0] / 3.0 + u[:, 0] * (L[0] / 3.0)
    pos[:, 1:] = u[:, 1:] * L[1:]
    s = {"N": n, "ntypes": 1, "lo": np.zeros(3), "hi": L, "id": np.arange(1, n + 1, dtype=np.int32),
         "type": np.ones(n, dtype=np.int32), "pos": pos, "vel": maxwell(substream(seed, 1), n),
         "masses": np.array([1.0])}
    s.update(style="mdpd", temp=1.0, rc=1.0, rd=0.75, saru_seed=42, dt=0.01, skin=0.3, global_cut=1.0,
             max_rho=8.0, coeff=mdpd_coeff(1, {(1, 1): (-40.0, 25.0, 4.5, 1.0, 0.75)}),
             name="C1 mDPD slab 8k")
    return s


def config_c2(seed=99, n_chains=23437, chain=16, box=50.0, step=0.9):
    n = n_chains * chain
    L = np.array([box, box, box])
    start = uniform(substream(seed, 0), 3 * n_chains).reshape(n_chains, 3) * L
    dirs = normal(substream(seed, 1), 3 * n_chains * (chain - 1)).reshape(n_chains, chain - 1, 3)
    dirs /= np.linalg.norm(dirs, axis=2, keepdims=True)
    walk = np.concatenate([start[:, None, :], start[:, None, :] + np.cumsum(dirs * step, axis=1)], axis=1)
    pos = np.mod(walk.reshape(n, 3), L)
    types = np.tile(np.array([1] * (chain // 2) + [2] * (chain - chain // 2), dtype=np.int32), n_chains)
    ids = np.arange(1, n + 1, dtype=np.int32)
    first = (np.arange(n_chains) * chain)[:, None] + np.arange(chain - 1)[None, :]
    bonds = np.stack([np.ones(first.size, dtype=np.int32), first.ravel() + 1, first.ravel() + 2], axis=1)
    s = {"N": n, "ntypes": 2, "lo": np.zeros(3), "hi": L, "id": ids, "type": types, "pos": pos,
         "vel": maxwell(substream(seed, 2), n), "masses": np.array([1.0, 1.0]),
         "bonds": bonds.astype(np.int32), "n_bond_types": 1, "bond_coeff": np.array([4.0, 0.0])}
    s.update(style="dpd", temp=1.0, rc=1.0, saru_seed=42, dt=0.04, skin=0.3, global_cut=1.0, max_rho=6.0,
             coeff=dpd_coeff(2, {(1, 1): (25.0, 4.5, 1.0), (1, 2): (40.0, 4.5, 1.0), (2, 2): (25.0, 4.5, 1.0)}),
             name="C2 diblock melt 375k")
    return s


def config_c3(seed=2024, m=216, box=150.0, jitter=0.1):
    n = m ** 3
    a = box / m
    g = np.arange(m, dtype=np.float64) * a + 0.5 * a
    gx, gy, gz = np.meshgrid(g, g, g, indexing="ij")
    pos = np.stack([gx.ravel(), gy.ravel(), gz.ravel()], axis=1)
    pos += (uniform(substream(seed, 0), 3 * n).reshape(n, 3) * 2.0 - 1.0) * jitter
    L = np.array([box, box, box])
    pos = np.mod(pos, L)
    s = {"N": n, "ntypes": 1, "lo": np.zeros(3), "hi": L, "id": np.arange(1, n + 1, dtype=np.int32),
         "type": np.ones(n, dtype=np.int32), "pos": pos, "vel": maxwell(substream(seed, 1), n),
         "masses": np.array([1.0])}
    s.update(style="dpd", temp=1.0, rc=1.0, saru_seed=42, dt=0.04, skin=0.3, global_cut=1.0, max_rho=6.0,
             coeff=dpd_coeff(1, {(1, 1): (25.0, 4.5, 1.0)}), name=f"C3 DPD fluid {n}")
    return s


def scaled_dpd(n_target, rho=3.0, seed=2024):
    """A cubic DPD fluid of about n_target beads at density rho (C0 physics)."""
    L = (n_target / rho) ** (1.0 / 3.0)
    s = dpd_fluid(int(n_target), (L, L, L), seed)
    s.update(style="dpd", temp=1.0, rc=1.0, saru_seed=42, dt=0.04, skin=0.3, global_cut=1.0, max_rho=6.0,
             coeff=dpd_coeff(1, {(1, 1): (25.0, 4.5, 1.0)}), name=f"DPD fluid {int(n_target)}")
    return s


def replica_request(master_seed, k):
    """Request k of the served stream: style, size, seed (C4)."""
    u = uniform(substream(master_seed, 100 + k), 3)
    style = "dpd" if u[0] < 0.7 else "mdpd"
    n = int(round(math.exp(math.log(2000.0) + u[1] * (math.log(200000.0) - math.log(2000.0)))))
    return style, n, int(splitmix64(master_seed, 1, offset=1000 + k)[0] & 0x7FFFFFFF)


def arrival_times(master_seed, n, rate=200.0):
    u = uniform(substream(master_seed, 7), n)
    return np.cumsum(-np.log(1.0 - u) / rate)


def replica_system(style, n, seed):
    rho = 3.0 if style == "dpd" else 6.0
    L = (n / rho) ** (1.0 / 3.0)
    s = dpd_fluid(n, (L, L, L), seed)
    if style == "dpd":
        s.update(style="dpd", temp=1.0, rc=1.0, saru_seed=seed & 0xFFFF, dt=0.04, skin=0.3, global_cut=1.0,
                 max_rho=6.0, coeff=dpd_coeff(1, {(1, 1): (25.0, 4.5, 1.0)}))
    else:
        s.update(style="mdpd", temp=1.0, rc=1.0, rd=0.75, saru_seed=seed & 0xFFFF, dt=0.01, skin=0.3,
                 global_cut=1.0, max_rho=8.0, coeff=mdpd_coeff(1, {(1, 1): (-40.0, 25.0, 4.5, 1.0, 0.75)}))
    s["name"] = f"{style} replica {n}"
    return s


def config_angles(n_chains=1500, box=20.0, seed=31):
    """Semi-flexible diblock chains: the C2 melt at a given size plus harmonic
    angles over every consecutive triple of a chain (type 1: k=3, theta0=180
    degrees for the A half; type 2: k=2, theta0=120 for triples touching B)."""
    cfg = config_c2(seed=seed, n_chains=n_chains, box=box)
    chain = 16
    first = np.arange(n_chains)[:, None] * chain + np.arange(chain - 2)[None, :]
    ids = first.ravel() + 1
    t = np.where((first.ravel() % chain) + 2 < chain // 2, 1, 2)
    cfg["angles"] = np.stack([t, ids, ids + 1, ids + 2], axis=1).astype(np.int32)
    cfg["n_angle_types"] = 2
    cfg["angle_coeff"] = np.array([3.0, 180.0, 2.0, 120.0])
    cfg["name"] = f"angle melt {cfg['N']}"
    return cfg


def config_wetting(seed=31, n_fluid=3000, box=(10.0, 10.0, 14.0)):
    """Small wetting-study geometry for the group fixes and computes: a DPD
    fluid (type 1) between two frozen wall slabs one cell thick at the bottom
    (type 2) and top (type 3) of z, plus a 20-bead probe cluster (type 4) in
    the fluid.  rho = 3 everywhere; the fluid starts inside z in [1.2, L-1.2).
    """
    box = np.asarray(box, dtype=np.float64)
    area = box[0] * box[1]
    n_wall = int(round(3.0 * area * 1.0))
    u = uniform(substream(seed, 0), 3 * (n_fluid + 2 * n_wall + 20)).reshape(-1, 3)
    fl = u[:n_fluid] * np.array([box[0], box[1], box[2] - 2.4]) + np.array([0.0, 0.0, 1.2])
    wlo = u[n_fluid:n_fluid + n_wall] * np.array([box[0], box[1], 1.0])
    whi = u[n_fluid + n_wall:n_fluid + 2 * n_wall] * np.array([box[0], box[1], 1.0]) + np.array([0, 0, box[2] - 1.0])
    probe = u[n_fluid + 2 * n_wall:] * 1.5 + box / 2.0 - 0.75
    pos = np.concatenate([fl, wlo, whi, probe])
    typ = np.concatenate([np.full(n_fluid, 1), np.full(n_wall, 2), np.full(n_wall, 3), np.full(20, 4)]).astype(np.int32)
    n = len(pos)
    vel = maxwell(substream(seed, 1), n)
    vel[typ == 2] = 0.0
    vel[typ == 3] = 0.0
    table = {}
    for ti in range(1, 5):
        for tj in range(ti, 5):
            table[(ti, tj)] = (25.0 if ti == tj else 40.0, 4.5, 1.0)
    return {"N": n, "ntypes": 4, "lo": np.zeros(3), "hi": box.copy(), "id": np.arange(1, n + 1, dtype=np.int32),
            "type": typ, "pos": pos, "vel": vel, "masses": np.array([1.0, 1.0, 1.0, 2.0]),
            "style": "dpd", "temp": 1.0, "rc": 1.0, "saru_seed": 77, "dt": 0.02, "skin": 0.3,
            "global_cut": 1.0, "max_rho": 8.0, "coeff": dpd_coeff(4, table), "name": "wetting walls + probe"}
