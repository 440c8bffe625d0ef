"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or the reported CPU baseline.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


def build():
    r = subprocess.run(["make", "-C", HERE, "-s"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)
    return LIB_PATH


class Box(C.Structure):
    _fields_ = [("lo", C.c_double * 3), ("hi", C.c_double * 3)]


class OrcSystem(C.Structure):
    _fields_ = [
        ("N", C.c_int), ("ntypes", C.c_int), ("style", C.c_int), ("box", Box),
        ("id", C.POINTER(C.c_int)), ("type", C.POINTER(C.c_int)),
        ("pos", C.POINTER(C.c_double)), ("vel", C.POINTER(C.c_double)), ("uwpos", C.POINTER(C.c_double)),
        ("masses", C.c_double * 16), ("coeff", C.c_double * (5 * 136)),
        ("temp", C.c_double), ("dt", C.c_double), ("seed", C.c_uint32),
        ("global_cut", C.c_double), ("skin", C.c_double), ("max_rho", C.c_double),
        ("check_every", C.c_int), ("log_every", C.c_int), ("n_bonds", C.c_int),
        ("bondlist", C.POINTER(C.c_int)), ("bond_coeff", C.c_double * 32),
    ]


_lib = None
_PD, _PI, _PU = C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_uint32)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_saru_u32.restype = C.c_uint32
        L.orc_saru_u32.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_int32]
        L.orc_saru_rand_f32.restype = C.c_float
        L.orc_saru_rand_f32.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_int32]
        L.orc_saru_rand_f64.restype = C.c_double
        L.orc_saru_rand_f64.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_int32]
        L.orc_saru_batch.argtypes = [C.c_int, _PU, _PU, C.c_uint32, C.POINTER(C.c_int32), _PU,
                                     C.POINTER(C.c_float), _PD]
        L.orc_pair_index.argtypes = [C.c_int, C.c_int, C.c_int]
        L.orc_neigh_build.argtypes = [C.c_int, _PD, C.POINTER(Box), C.c_double, C.c_int, _PI, _PI]
        L.orc_neigh_brute.argtypes = [C.c_int, _PD, C.POINTER(Box), C.c_double, C.c_int, _PI, _PI]
        L.orc_neigh_check.argtypes = [C.c_int, _PD, _PD, C.POINTER(Box), C.c_double]
        L.orc_n_max_neigh.argtypes = [C.c_double, C.c_double, C.c_double]
        L.orc_dpd_force.argtypes = [C.c_int, C.c_int, _PI, _PI, _PD, _PD, C.POINTER(Box), _PI, _PI, C.c_int,
                                    _PD, C.c_double, C.c_double, C.c_uint32, C.c_uint32, _PD, _PD, _PD]
        L.orc_mdpd_density.argtypes = [C.c_int, C.c_int, _PI, _PD, C.POINTER(Box), _PI, _PI, C.c_int, _PD, _PD]
        L.orc_mdpd_force.argtypes = [C.c_int, C.c_int, _PI, _PI, _PD, _PD, _PD, C.POINTER(Box), _PI, _PI,
                                     C.c_int, _PD, C.c_double, C.c_double, C.c_uint32, C.c_uint32, _PD, _PD,
                                     _PD]
        L.orc_bond_force.argtypes = [C.c_int, _PI, _PD, _PD, C.POINTER(Box), _PD, _PD, _PD]
        L.orc_angle_force.argtypes = [C.c_int, _PI, _PD, _PD, C.POINTER(Box), _PD, _PD, _PD]
        L.orc_nve_initial.argtypes = [C.c_int, C.c_double, _PI, _PD, _PD, C.POINTER(Box), _PD, _PD, _PD]
        L.orc_nve_final.argtypes = [C.c_int, C.c_double, _PI, _PD, _PD, _PD]
        L.orc_thermo.argtypes = [C.c_int, _PI, _PD, _PD, _PD, _PD, C.POINTER(Box), C.c_int, _PD, _PD]
        L.orc_run.argtypes = [C.POINTER(OrcSystem), C.c_int, _PD, C.c_int, _PD, _PI]
        L.orc_omp_create.restype = C.c_void_p
        L.orc_omp_create.argtypes = [C.POINTER(OrcSystem), C.c_int]
        L.orc_omp_steps.argtypes = [C.c_void_p, C.c_uint32, C.c_int]
        L.orc_omp_threads.argtypes = [C.c_void_p]
        L.orc_omp_destroy.argtypes = [C.c_void_p]
        _lib = L
    return _lib


def p(a, t=C.c_double):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


def box(lo, hi):
    b = Box()
    for k in range(3):
        b.lo[k], b.hi[k] = float(lo[k]), float(hi[k])
    return b


def saru_batch(t1, t2, seed, steps):
    t1 = np.ascontiguousarray(t1, dtype=np.uint32)
    t2 = np.ascontiguousarray(t2, dtype=np.uint32)
    steps = np.ascontiguousarray(steps, dtype=np.int32)
    n = len(t1)
    s, r32, r64 = np.empty(n, np.uint32), np.empty(n, np.float32), np.empty(n, np.float64)
    lib().orc_saru_batch(n, p(t1, C.c_uint32), p(t2, C.c_uint32), int(seed) & 0xFFFFFFFF, p(steps, C.c_int32),
                         p(s, C.c_uint32), p(r32, C.c_float), p(r64))
    return s, r32, r64


def n_max_neigh(global_cut, skin, max_rho):
    return lib().orc_n_max_neigh(global_cut, skin, max_rho)


def neigh_build(pos, lo, hi, bin_size, n_max, brute=False):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    N = len(pos)
    n = np.zeros(N, np.int32)
    nb = np.full((N, n_max), -1, np.int32)
    b = box(lo, hi)
    fn = lib().orc_neigh_brute if brute else lib().orc_neigh_build
    dropped = fn(N, p(pos), C.byref(b), float(bin_size), int(n_max), p(n, C.c_int), p(nb, C.c_int))
    if dropped != 0:
        raise RuntimeError(f"oracle neighbour list: code {dropped}")
    return n, nb


def neigh_check(pos, last, lo, hi, skin):
    b = box(lo, hi)
    return lib().orc_neigh_check(len(pos), p(np.ascontiguousarray(pos)), p(np.ascontiguousarray(last)),
                                 C.byref(b), float(skin))


def forces(cfg, pos=None, vel=None, step=0, with_list=None):
    """Pair (+bond) force, per-bead PE, virial and density at a given state, FP64.

    Follows Pair::compute_force then Bond::compute_force with a log step.
    Returns (force (N,3), pe (N,), bond_pe_total, virial (N,6), rho (N,) or None).
    """
    L = lib()
    N, nt = cfg["N"], cfg["ntypes"]
    pos = np.ascontiguousarray(cfg["pos"] if pos is None else pos, dtype=np.float64)
    vel = np.ascontiguousarray(cfg["vel"] if vel is None else vel, dtype=np.float64)
    ids = np.ascontiguousarray(cfg["id"], dtype=np.int32)
    order = np.argsort(ids)
    # oracle arrays are indexed by id - 1
    idx = np.ascontiguousarray(ids[order])
    typ = np.ascontiguousarray(np.asarray(cfg["type"], dtype=np.int32)[order])
    pos, vel = np.ascontiguousarray(pos[order]), np.ascontiguousarray(vel[order])
    b = box(cfg["lo"], cfg["hi"])
    bin_size = cfg["global_cut"] + cfg["skin"]
    n_max = max(64, n_max_neigh(cfg["global_cut"], cfg["skin"], max(cfg["max_rho"], 10.0)))
    n, nb = with_list if with_list is not None else neigh_build(pos, cfg["lo"], cfg["hi"], bin_size, n_max)
    n_max = nb.shape[1]
    coeff = np.ascontiguousarray(cfg["coeff"], dtype=np.float64)
    f, pe, vir = np.zeros((N, 3)), np.zeros(N), np.zeros((N, 6))
    rho = None
    if cfg["style"] == "mdpd":
        rho = np.zeros(N)
        L.orc_mdpd_density(N, nt, p(typ, C.c_int), p(pos), C.byref(b), p(n, C.c_int), p(nb, C.c_int), n_max,
                           p(coeff), p(rho))
        L.orc_mdpd_force(N, nt, p(idx, C.c_int), p(typ, C.c_int), p(pos), p(vel), p(rho), C.byref(b),
                         p(n, C.c_int), p(nb, C.c_int), n_max, p(coeff), cfg["temp"], cfg["dt"],
                         cfg["saru_seed"] & 0xFFFFFFFF, int(step), p(f), p(pe), p(vir))
    else:
        L.orc_dpd_force(N, nt, p(idx, C.c_int), p(typ, C.c_int), p(pos), p(vel), C.byref(b), p(n, C.c_int),
                        p(nb, C.c_int), n_max, p(coeff), cfg["temp"], cfg["dt"], cfg["saru_seed"] & 0xFFFFFFFF,
                        int(step), p(f), p(pe), p(vir))
    bpe_total = 0.0
    if "bonds" in cfg:
        bl = np.ascontiguousarray(cfg["bonds"], dtype=np.int32)
        bpe = np.zeros(len(bl))
        L.orc_bond_force(len(bl), p(bl, C.c_int), p(np.ascontiguousarray(cfg["bond_coeff"], dtype=np.float64)),
                         p(pos), C.byref(b), p(f), p(bpe), p(vir))
        bpe_total = float(bpe.sum())
    if "angles" in cfg:
        fa, _, va = angle_force(cfg["angles"], cfg["angle_coeff"], pos, cfg["lo"], cfg["hi"])
        f += fa
        vir += va
    return f, pe, bpe_total, vir, rho


def angle_force(anglelist, angle_coeff, pos, lo, hi):
    """Harmonic angle force, per-angle PE and virial at positions pos (N,3),
    rows indexed by id - 1.  Returns (force (N,3), angle_pe (n_angles,), virial (N,6))."""
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    N = len(pos)
    al = np.ascontiguousarray(anglelist, dtype=np.int32)
    f, vir = np.zeros((N, 3)), np.zeros((N, 6))
    ape = np.zeros(len(al))
    b = box(lo, hi)
    lib().orc_angle_force(len(al), p(al, C.c_int), p(np.ascontiguousarray(angle_coeff, dtype=np.float64)),
                          p(pos), C.byref(b), p(f), p(ape), p(vir))
    return f, ape, vir


def make_orc_system(cfg, log_every=100, check_every=1):
    """An OrcSystem over copies of cfg's arrays (kept alive on the returned object)."""
    s = OrcSystem()
    N = cfg["N"]
    ids = np.ascontiguousarray(cfg["id"], dtype=np.int32)
    order = np.argsort(ids)
    keep = {
        "id": np.ascontiguousarray(ids[order]),
        "type": np.ascontiguousarray(np.asarray(cfg["type"], dtype=np.int32)[order]),
        "pos": np.ascontiguousarray(np.asarray(cfg["pos"], dtype=np.float64)[order]),
        "vel": np.ascontiguousarray(np.asarray(cfg["vel"], dtype=np.float64)[order]),
    }
    keep["uwpos"] = keep["pos"].copy()
    s.N, s.ntypes, s.style = N, cfg["ntypes"], 1 if cfg["style"] == "mdpd" else 0
    s.box = box(cfg["lo"], cfg["hi"])
    s.id, s.type = p(keep["id"], C.c_int), p(keep["type"], C.c_int)
    s.pos, s.vel, s.uwpos = p(keep["pos"]), p(keep["vel"]), p(keep["uwpos"])
    for k, m in enumerate(cfg["masses"]):
        s.masses[k] = m
    for k, c in enumerate(cfg["coeff"]):
        s.coeff[k] = c
    s.temp, s.dt, s.seed = cfg["temp"], cfg["dt"], cfg["saru_seed"] & 0xFFFFFFFF
    s.global_cut, s.skin, s.max_rho = cfg["global_cut"], cfg["skin"], max(cfg["max_rho"], 10.0)
    s.check_every, s.log_every = check_every, log_every
    if "bonds" in cfg:
        keep["bonds"] = np.ascontiguousarray(cfg["bonds"], dtype=np.int32)
        s.n_bonds = len(keep["bonds"])
        s.bondlist = p(keep["bonds"], C.c_int)
        for k, c in enumerate(cfg["bond_coeff"]):
            s.bond_coeff[k] = c
    s._keep = keep
    return s


def run(cfg, nsteps, log_every=100, check_every=1):
    """Whole reference-order run; returns (thermo rows (n,11), final state dict, rebuilds)."""
    s = make_orc_system(cfg, log_every, check_every)
    rows = np.zeros((nsteps // log_every + 2, 11))
    force = np.zeros((cfg["N"], 3))
    builds = C.c_int()
    nr = lib().orc_run(C.byref(s), int(nsteps), p(rows), len(rows), p(force), C.byref(builds))
    if nr < 0:
        raise RuntimeError("oracle run: neighbour list overflow")
    k = s._keep
    return rows[:nr], {"pos": k["pos"], "vel": k["vel"], "uwpos": k["uwpos"], "force": force}, builds.value
