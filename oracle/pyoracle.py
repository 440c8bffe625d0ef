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
OBSERVER = C.CFUNCTYPE(None, C.c_int, C.POINTER(C.c_double), C.c_void_p)
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
        L.orc_saru_stages.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_int32, _PU]
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
        L.orc_run_observe.argtypes = [C.POINTER(OrcSystem), C.c_int, _PD, C.c_int, _PD, _PI, C.c_int,
                                      OBSERVER, C.c_void_p]
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


def saru_stages(t1, t2, seed, step):
    """(premixed st, ws, st and ws after the 8 TEA cycles, finished value) of one call."""
    out = (C.c_uint32 * 5)()
    lib().orc_saru_stages(int(t1) & 0xFFFFFFFF, int(t2) & 0xFFFFFFFF, int(seed) & 0xFFFFFFFF, int(step),
                          C.cast(out, _PU))
    return tuple(int(v) for v in out)


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


def run_observed(cfg, nsteps, every, fn, log_every=100, check_every=1):
    """orc_run that calls fn(step, pos (N, 3) rows by id - 1) after every every-th step."""
    s = make_orc_system(cfg, log_every, check_every)
    N = cfg["N"]

    def cb(step, pos, user):
        fn(int(step), np.ctypeslib.as_array(pos, shape=(N * 3,)).reshape(N, 3).copy())

    cfn = OBSERVER(cb)
    rows = np.zeros((nsteps // log_every + 2, 11))
    nr = lib().orc_run_observe(C.byref(s), int(nsteps), p(rows), len(rows), None, None, int(every), cfn, None)
    if nr < 0:
        raise RuntimeError("oracle run: neighbour list overflow")
    return rows[:nr]


# ---------------------------------------------------------------------------
# Group fixes and computes: a restatement of fix/fix_*.cu and compute/compute_*.cu
# in the reference's step order (main/utils.cuh:166-314), on id-ordered arrays.
# Groups are sets of types (system/system.cu:235-301).  Hook order inside one
# hook point: the order of `fixes`, with post_integrate's wall_reflect fixes
# before its two_wall fixes (the reference walks an unordered_map there).
# ---------------------------------------------------------------------------

def _wrap(x, lo, hi):
    """fix/fix_nve.cu:68-76 (and fix_move.cu, fix_recenter.cu): shift by whole
    box lengths until lo <= x < hi, one coordinate array at a time."""
    x = np.array(x, dtype=np.float64)
    L = hi - lo
    while True:
        up, dn = x >= hi, x < lo
        if not (up.any() or dn.any()):
            return x
        x = x - L * (up.astype(np.float64) - dn.astype(np.float64))


def _in_group(types, group):
    return np.isin(types, sorted(group))


def _reflect(x, v, lo, hi):
    """fix_wall_reflect.cu:96-118 / fix_two_wall.cu:118-140 on one axis."""
    x, v = x.copy(), v.copy()
    below, above = x < lo, (x > hi) & ~(x < lo)
    disp_b, disp_a = 2.0 * (lo - x[below]), 2.0 * (hi - x[above])
    x[below] += disp_b
    v[below] = -v[below]
    x[above] += disp_a
    v[above] = -v[above]
    d = np.zeros_like(x)
    d[below], d[above] = disp_b, disp_a
    return x, v, d


def _spring_terms(fx, com):
    """fix_spring.cu:300-350: spring on the group's centre of mass."""
    k, c, r0 = fx["k"], np.asarray(fx["xyz"], dtype=np.float64), fx["r0"]
    d = com - c
    r = float(np.sqrt(d @ d))
    dr = r - r0
    if r > 1.0e-10:
        f = -k * dr * d / r
        fm = float(np.sqrt(f @ f))
        return f, (-fm if dr < 0.0 else fm), 0.5 * k * dr * dr
    return np.zeros(3), 0.0, 0.0


def run_fixes(cfg, nsteps, fixes, computes=(), log_every=1):
    """Reference-order run with group fixes and computes, FP64.

    fixes: dicts with "kind" in nve, move, wall_reflect, two_wall, spring,
    recenter, rvcm; "group" (set of types) and the kind's parameters
    (move: v; wall_reflect: lo, hi, dir; two_wall: lo_group, hi_group, dir;
    spring: k, xyz, r0, every (records; 0 none); recenter: every, xyz;
    rvcm: every, flags (x, y, z)).
    computes: dicts with "kind" in temp, msd, rho, tf, wall_contact, "group",
    "every" and (rho) region (xlo, xhi, ylo, yhi, zlo, zhi) or (wall_contact)
    wall_group, side (+1 lo, -1 hi), dir, cutoff.

    Returns (thermo rows (n, 13), records {("compute", k) | ("fix", k): list of
    rows [step, v...]}, final state dict by id-1).
    """
    N = cfg["N"]
    ids = np.asarray(cfg["id"], dtype=np.int64)
    order = np.argsort(ids)
    types = np.asarray(cfg["type"], dtype=np.int64)[order]
    pos = np.asarray(cfg["pos"], dtype=np.float64)[order].copy()
    vel = np.asarray(cfg["vel"], dtype=np.float64)[order].copy()
    lo, hi = np.asarray(cfg["lo"], dtype=np.float64), np.asarray(cfg["hi"], dtype=np.float64)
    L = hi - lo
    vol = float(np.prod(L))
    masses = np.asarray(cfg["masses"], dtype=np.float64)
    m = masses[types - 1]
    dt = cfg["dt"]
    pos = np.stack([_wrap(pos[:, d], lo[d], hi[d]) for d in range(3)], axis=1)
    uw = pos.copy()
    x0 = uw.copy()
    sorted_cfg = dict(cfg, id=np.arange(1, N + 1, dtype=np.int32), type=types.astype(np.int32))
    rec = {}
    rows = []
    gmask = [_in_group(types, f["group"]) for f in fixes]
    cmask = [_in_group(types, c["group"]) for c in computes]

    def force(step):
        f, pe, bpe, vir, _ = forces(sorted_cfg, pos=pos, vel=vel, step=step)
        return f, pe, bpe, vir

    def com(g):
        return (m[g, None] * uw[g]).sum(axis=0) / m[g].sum()

    def post_force(step, f):
        for k, fx in enumerate(fixes):
            if fx["kind"] != "spring":
                continue
            g = gmask[k]
            F, fs, e = _spring_terms(fx, com(g))
            f[g] += (F / m[g].sum())[None, :] * m[g, None]
            if fx.get("every", 0) > 0 and step % fx["every"] == 0:
                rec.setdefault(("fix", k), []).append([step, F[0], F[1], F[2], fs, e])

    def thermo(step, pe, bpe, vir):
        ke = 0.5 * float((m[:, None] * vel * vel).sum())
        mvv = [(m * vel[:, a] * vel[:, b]).sum() for a, b in ((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2))]
        p6 = [(mvv[c] + vir[:, c].sum()) / vol for c in range(6)]
        ape = float(angle_force(cfg["angles"], cfg["angle_coeff"], pos, lo, hi)[1].sum()) if "angles" in cfg else 0.0
        rows.append([2.0 * ke / (3.0 * N - 3.0), pe.sum(), bpe, (p6[0] + p6[3] + p6[5]) / 3.0] + p6
                    + [ke, float(step), ape])

    def do_computes(step, f):
        for k, c in enumerate(computes):
            if step % c["every"] != 0:
                continue
            g = cmask[k]
            kind = c["kind"]
            if kind == "temp":
                ke = 0.5 * float((m[g, None] * vel[g] ** 2).sum())
                val = [2.0 * ke / (3.0 * g.sum() - 3.0)]
            elif kind == "msd":
                val = [float(((uw[g] - x0[g]) ** 2).sum()) / g.sum()]
            elif kind == "rho":
                r = c["region"]
                inside = g.copy()
                for d in range(3):
                    inside &= (pos[:, d] >= r[2 * d]) & (pos[:, d] < r[2 * d + 1])
                val = [inside.sum() / ((r[1] - r[0]) * (r[3] - r[2]) * (r[5] - r[4]))]
            elif kind == "tf":
                val = list(f[g].sum(axis=0))
            elif kind == "wall_contact":
                w = _in_group(types, c["wall_group"])
                xd = pos[:, c["dir"]]
                if c["side"] > 0:
                    b = xd[w].max()
                    inside = g & (xd >= b) & (xd < b + c["cutoff"])
                else:
                    b = xd[w].min()
                    inside = g & (xd >= b - c["cutoff"]) & (xd < b)
                val = [float(inside.sum())]
            rec.setdefault(("compute", k), []).append([step] + val + [0.0] * (5 - len(val)))

    # zero_step (main/utils.cuh:166-205)
    f, pe, bpe, vir = force(0)
    post_force(0, f)
    if log_every > 0:
        thermo(0, pe, bpe, vir)
    do_computes(0, f)
    for step in range(1, nsteps + 1):
        # initial_integrate: nve (fix_nve.cu:30-77), move (fix_move.cu:45-90)
        for k, fx in enumerate(fixes):
            g = gmask[k]
            if fx["kind"] == "nve":
                vel[g] = vel[g] + ((0.5 * dt) / m[g])[:, None] * f[g]
                for d in range(3):
                    pos[g, d] = _wrap(pos[g, d] + dt * vel[g, d], lo[d], hi[d])
                uw[g] += dt * vel[g]
            elif fx["kind"] == "move":
                vm = np.asarray(fx["v"], dtype=np.float64)
                vel[g] = vm
                for d in range(3):
                    pos[g, d] = _wrap(pos[g, d] + vm[d] * dt, lo[d], hi[d])
                uw[g] += vm * dt
        # post_integrate: wall_reflect, then two_wall
        for k, fx in enumerate(fixes):
            if fx["kind"] != "wall_reflect":
                continue
            g, d = gmask[k], fx["dir"]
            x, v, disp = _reflect(pos[g, d], vel[g, d], fx["lo"], fx["hi"])
            pos[g, d], vel[g, d] = x, v
            uw[g, d] += disp
        for k, fx in enumerate(fixes):
            if fx["kind"] != "two_wall":
                continue
            g, d = gmask[k], fx["dir"]
            wlo = pos[_in_group(types, fx["lo_group"]), d].max()
            whi = pos[_in_group(types, fx["hi_group"]), d].min()
            x, v, disp = _reflect(pos[g, d], vel[g, d], wlo, whi)
            pos[g, d], vel[g, d] = x, v
            uw[g, d] += disp
        f, pe, bpe, vir = force(step)
        post_force(step, f)
        # final_integrate: nve (fix_nve.cu:120-152)
        for k, fx in enumerate(fixes):
            if fx["kind"] == "nve":
                g = gmask[k]
                vel[g] = vel[g] + ((0.5 * dt) / m[g])[:, None] * f[g]
        # end_of_step: rvcm (fix_rvcm.cu:220-292), recenter (fix_recenter.cu:260-340)
        for k, fx in enumerate(fixes):
            if fx["kind"] not in ("rvcm", "recenter") or step % fx["every"] != 0:
                continue
            g = gmask[k]
            if fx["kind"] == "rvcm":
                vcm = (m[g, None] * vel[g]).sum(axis=0) / m[g].sum()
                vel[g] -= np.asarray(fx["flags"], dtype=np.float64) * vcm
            else:
                shift = np.asarray(fx["xyz"], dtype=np.float64) - com(g)
                for d in range(3):
                    pos[g, d] = _wrap(pos[g, d] + shift[d], lo[d], hi[d])
                uw[g] += shift
        if log_every > 0 and step % log_every == 0:
            thermo(step, pe, bpe, vir)
        do_computes(step, f)
    return np.array(rows), rec, {"pos": pos, "vel": vel, "uwpos": uw, "force": f}
