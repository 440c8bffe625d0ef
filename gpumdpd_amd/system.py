"""Python front of libgdpd: a device context and the systems ("replicas") on it.

Method names follow the reference's plugin hooks so tests read like the
reference's step loop (main/utils.cuh:166-314):
    Pair::compute_force        -> System.pair_compute(step, log)
    Bond::compute_force        -> System.bond_compute(step, log)
    nve::initial_integrate     -> System.nve_initial()
    nve::final_integrate       -> System.nve_final()
    Neigh_list::build          -> System.neigh_build()
    Neigh_list::check_update   -> System.neigh_check()
    Thermo::process            -> System.thermo()
and the fused path runs a whole `run N` (System.run) in one launch per step.

Host arrays are indexed like the reference's host vectors: row = id - 1,
positions/velocities/forces (N, 3), virial (N, 6).
"""
import ctypes as C

import numpy as np

from . import _lib as L


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype)) if a is not None else None


def type_mask(types):
    """Group of bead types as the ABI's bitmask (bit t-1 for type t)."""
    m = 0
    for t in types:
        m |= 1 << (int(t) - 1)
    return m


class Context:
    """One device + one HIP stream (libgdpd gdpd_ctx)."""

    def __init__(self, device=0):
        self.lib = L.load()
        h = C.c_void_p()
        rc = self.lib.gdpd_create(C.byref(h), device)
        if rc != L.GDPD_OK:
            raise L.GdpdError(rc, f"gdpd_create(device={device}) failed")
        self.h = h

    def check(self, rc):
        if rc != L.GDPD_OK:
            raise L.GdpdError(rc, self.lib.gdpd_last_error(self.h).decode())

    def close(self):
        if self.h:
            self.lib.gdpd_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def sync(self):
        self.check(self.lib.gdpd_sync(self.h))

    def admission(self, on):
        """Open (True) or close (False) an admission window: systems created and
        zero-stepped inside it are set up on a second stream, beside the work
        already queued for the others."""
        self.check(self.lib.gdpd_admission(self.h, 1 if on else 0))

    def batch_advance(self, systems, nsteps):
        ids = np.array([s.sid for s in systems], dtype=np.int32)
        self.check(self.lib.gdpd_batch_advance(self.h, _ptr(ids, C.c_int), len(ids), int(nsteps)))

    def saru_eval(self, tag1, tag2, seed, steps):
        t1 = np.ascontiguousarray(tag1, dtype=np.uint32)
        t2 = np.ascontiguousarray(tag2, dtype=np.uint32)
        st = np.ascontiguousarray(steps, dtype=np.int32)
        n = len(t1)
        bits, r32, r64 = np.empty(n, np.uint32), np.empty(n, np.float32), np.empty(n, np.float64)
        self.check(self.lib.gdpd_saru_eval(self.h, n, _ptr(t1, C.c_uint32), _ptr(t2, C.c_uint32),
                                           int(seed) & 0xFFFFFFFF, _ptr(st, C.c_int32), _ptr(bits, C.c_uint32),
                                           _ptr(r32, C.c_float), _ptr(r64, C.c_double)))
        return bits, r32, r64

    def step_times(self, systems):
        """(times (n, STEP_TS_CAP) in ms of the device clock, step counters (n,)):
        the completion time of step s of system k sits at [k, s % STEP_TS_CAP]."""
        ids = np.array([s.sid for s in systems], dtype=np.int32)
        n = len(ids)
        t = np.zeros((max(n, 1), L.STEP_TS_CAP), dtype=np.uint64)
        st = np.zeros(max(n, 1), dtype=np.int64)
        tick = C.c_double()
        self.check(self.lib.gdpd_step_times(self.h, _ptr(ids, C.c_int), n, _ptr(t, C.c_uint64),
                                            _ptr(st, C.c_int64), C.byref(tick)))
        return t[:n].astype(np.float64) * (tick.value * 1e-6), st[:n]

    def kernel_timing(self, on):
        self.check(self.lib.gdpd_kernel_timing(self.h, 1 if on else 0))

    def kernel_timing_get(self):
        ms, n, tot = C.c_double(), C.c_int64(), C.c_double()
        self.check(self.lib.gdpd_kernel_timing_get(self.h, C.byref(ms), C.byref(n), C.byref(tot)))
        return ms.value, n.value, tot.value

    def kernel_phases(self):
        """Mean microseconds per wave and timed force launch: staging, pair
        work, closing-barrier wait, write-back."""
        out = np.zeros(4)
        self.check(self.lib.gdpd_kernel_phases(self.h, _ptr(out, C.c_double)))
        return dict(zip(["stage_us", "pairs_us", "barrier_us", "writeback_us"], out.tolist()))


def make_desc(n_beads, n_types, lo, hi, dt, global_cut=1.0, skin=0.3, max_rho=6.0,
              check_every=1, log_every=100, precision=L.PREC_F32):
    d = L.SystemDesc()
    d.n_beads, d.n_types = int(n_beads), int(n_types)
    for k in range(3):
        d.lo[k], d.hi[k] = float(lo[k]), float(hi[k])
    d.dt, d.global_cut, d.skin, d.max_rho = float(dt), float(global_cut), float(skin), float(max_rho)
    d.check_every, d.log_every, d.precision = int(check_every), int(log_every), int(precision)
    return d


def system_bytes(desc, n_bonds=0):
    return int(L.load().gdpd_system_bytes(C.byref(desc), int(n_bonds)))


class System:
    """One simulated system on a Context (libgdpd system handle)."""

    def __init__(self, ctx, desc):
        self.ctx, self.lib, self.desc = ctx, ctx.lib, desc
        sid = C.c_int()
        ctx.check(self.lib.gdpd_sys_create(ctx.h, C.byref(desc), C.byref(sid)))
        self.sid = sid.value
        self.N = desc.n_beads

    def destroy(self):
        if self.sid is not None and self.ctx.h:
            self.ctx.check(self.lib.gdpd_sys_destroy(self.ctx.h, self.sid))
        self.sid = None

    # ---- state ----
    def upload(self, ids, types, pos, vel, masses):
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        types = np.ascontiguousarray(types, dtype=np.int32)
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        vel = None if vel is None else np.ascontiguousarray(vel, dtype=np.float64).reshape(-1, 3)
        masses = np.ascontiguousarray(masses, dtype=np.float64)
        self.ctx.check(self.lib.gdpd_upload(self.ctx.h, self.sid, _ptr(ids, C.c_int), _ptr(types, C.c_int),
                                            _ptr(pos, C.c_double), _ptr(vel, C.c_double),
                                            _ptr(masses, C.c_double)))

    def download(self):
        N = self.N
        pos, vel, force = (np.empty((N, 3)) for _ in range(3))
        img = np.empty((N, 3), dtype=np.int32)
        self.ctx.check(self.lib.gdpd_download(self.ctx.h, self.sid, _ptr(pos, C.c_double),
                                              _ptr(vel, C.c_double), _ptr(img, C.c_int),
                                              _ptr(force, C.c_double)))
        return {"pos": pos, "vel": vel, "image": img, "force": force}

    def download_pe_virial(self):
        N = self.N
        pe, vir, rho = np.empty(N), np.empty((N, 6)), np.empty(N)
        self.ctx.check(self.lib.gdpd_download_pe_virial(self.ctx.h, self.sid, _ptr(pe, C.c_double),
                                                        _ptr(vir, C.c_double), _ptr(rho, C.c_double)))
        return pe, vir, rho

    # ---- force fields ----
    def set_pair_dpd(self, temp, rc, seed, coeff):
        c = np.ascontiguousarray(coeff, dtype=np.float64)
        self.ctx.check(self.lib.gdpd_set_pair_dpd(self.ctx.h, self.sid, float(temp), float(rc),
                                                  int(seed) & 0xffffffff, _ptr(c, C.c_double)))

    def set_pair_mdpd(self, temp, rc, rd, seed, coeff):
        c = np.ascontiguousarray(coeff, dtype=np.float64)
        self.ctx.check(self.lib.gdpd_set_pair_mdpd(self.ctx.h, self.sid, float(temp), float(rc), float(rd),
                                                   int(seed) & 0xffffffff, _ptr(c, C.c_double)))

    def set_bonds(self, bondlist, n_bond_types, bond_coeff):
        b = np.ascontiguousarray(bondlist, dtype=np.int32).reshape(-1, 3)
        c = np.ascontiguousarray(bond_coeff, dtype=np.float64)
        self.ctx.check(self.lib.gdpd_set_bonds(self.ctx.h, self.sid, len(b), int(n_bond_types),
                                               _ptr(b, C.c_int), _ptr(c, C.c_double)))

    def set_angles(self, anglelist, n_angle_types, angle_coeff):
        a = np.ascontiguousarray(anglelist, dtype=np.int32).reshape(-1, 4)
        c = np.ascontiguousarray(angle_coeff, dtype=np.float64)
        self.ctx.check(self.lib.gdpd_set_angles(self.ctx.h, self.sid, len(a), int(n_angle_types),
                                                _ptr(a, C.c_int), _ptr(c, C.c_double)))

    # ---- hooks ----
    def neigh_build(self):
        self.ctx.check(self.lib.gdpd_neigh_build(self.ctx.h, self.sid))

    def neigh_check(self):
        f = C.c_int()
        self.ctx.check(self.lib.gdpd_neigh_check(self.ctx.h, self.sid, C.byref(f)))
        return f.value

    def neigh_download(self, n_max):
        n = np.empty(self.N, dtype=np.int32)
        nb = np.full((self.N, n_max), -1, dtype=np.int32)
        self.ctx.check(self.lib.gdpd_neigh_download(self.ctx.h, self.sid, _ptr(n, C.c_int),
                                                    _ptr(nb, C.c_int), int(n_max)))
        return n, nb

    def pair_compute(self, step, log=True):
        self.ctx.check(self.lib.gdpd_pair_compute(self.ctx.h, self.sid, int(step), 1 if log else 0))

    def bond_compute(self, step, log=True):
        self.ctx.check(self.lib.gdpd_bond_compute(self.ctx.h, self.sid, int(step), 1 if log else 0))

    def angle_compute(self, step, log=True):
        self.ctx.check(self.lib.gdpd_angle_compute(self.ctx.h, self.sid, int(step), 1 if log else 0))

    def nve_initial(self):
        self.ctx.check(self.lib.gdpd_nve_initial(self.ctx.h, self.sid))

    def nve_final(self):
        self.ctx.check(self.lib.gdpd_nve_final(self.ctx.h, self.sid))

    def thermo(self):
        out = np.empty(L.THERMO_LEN)
        self.ctx.check(self.lib.gdpd_thermo(self.ctx.h, self.sid, _ptr(out, C.c_double)))
        return out

    # ---- fused path ----
    def run(self, nsteps):
        self.ctx.check(self.lib.gdpd_run(self.ctx.h, self.sid, int(nsteps)))

    def zero_step(self):
        self.ctx.check(self.lib.gdpd_zero_step(self.ctx.h, self.sid))

    def advance(self, nsteps):
        self.ctx.check(self.lib.gdpd_advance(self.ctx.h, self.sid, int(nsteps)))

    def finish(self):
        self.ctx.check(self.lib.gdpd_finish(self.ctx.h, self.sid))

    def thermo_rows(self, max_rows=4096):
        out = np.empty((max_rows, L.THERMO_LEN))
        n = C.c_int()
        self.ctx.check(self.lib.gdpd_thermo_rows(self.ctx.h, self.sid, _ptr(out, C.c_double), max_rows,
                                                 C.byref(n)))
        return out[: n.value].copy()

    # ---- group fixes and computes ----
    def set_fixes(self, fixes):
        """fixes: dicts as in oracle-style specs: kind, group (types), and
        move v; wall_reflect lo, hi, dir; two_wall lo_group, hi_group, dir;
        spring k, xyz, r0, every; recenter every, xyz; rvcm every, flags."""
        arr = (L.Fix * max(1, len(fixes)))()
        for k, f in enumerate(fixes):
            a = arr[k]
            a.kind = L.FIX_KINDS[f["kind"]]
            a.group = type_mask(f["group"])
            a.dir = int(f.get("dir", 0))
            a.every = int(f.get("every", 0))
            kind = f["kind"]
            if kind == "move":
                a.p[0], a.p[1], a.p[2] = f["v"]
            elif kind == "wall_reflect":
                a.p[0], a.p[1] = f["lo"], f["hi"]
            elif kind == "two_wall":
                a.group2, a.group3 = type_mask(f["lo_group"]), type_mask(f["hi_group"])
            elif kind == "spring":
                a.p[0] = f["k"]
                a.p[1], a.p[2], a.p[3] = f["xyz"]
                a.p[4] = f["r0"]
            elif kind == "recenter":
                a.p[0], a.p[1], a.p[2] = f["xyz"]
            elif kind == "rvcm":
                a.flags = sum(1 << d for d in range(3) if f["flags"][d])
        self.ctx.check(self.lib.gdpd_set_fixes(self.ctx.h, self.sid, arr, len(fixes)))

    def set_computes(self, computes):
        arr = (L.Compute * max(1, len(computes)))()
        for k, c in enumerate(computes):
            a = arr[k]
            a.kind = L.COMPUTE_KINDS[c["kind"]]
            a.group = type_mask(c["group"])
            a.every = int(c["every"])
            if c["kind"] == "rho":
                for q in range(6):
                    a.p[q] = c["region"][q]
            elif c["kind"] == "wall_contact":
                a.group2 = type_mask(c["wall_group"])
                a.side, a.dir, a.p[0] = int(c["side"]), int(c["dir"]), c["cutoff"]
        self.ctx.check(self.lib.gdpd_set_computes(self.ctx.h, self.sid, arr, len(computes)))

    def fix_hook(self, index, hook, step):
        out = np.zeros(5)
        self.ctx.check(self.lib.gdpd_fix_hook(self.ctx.h, self.sid, int(index), L.HOOKS[hook], int(step),
                                              _ptr(out, C.c_double)))
        return out

    def compute_eval(self, index):
        out = np.zeros(3)
        self.ctx.check(self.lib.gdpd_compute_eval(self.ctx.h, self.sid, int(index), _ptr(out, C.c_double)))
        return out

    def records(self, index, fix=False, max_rows=L.RECORD_CAP):
        out = np.empty((max_rows, L.RECORD_LEN))
        n = C.c_int()
        slot = (L.MAX_COMPUTES + index) if fix else index
        self.ctx.check(self.lib.gdpd_records(self.ctx.h, self.sid, slot, _ptr(out, C.c_double), max_rows,
                                             C.byref(n)))
        return out[: n.value].copy()

    def step(self):
        s = C.c_int64()
        self.ctx.check(self.lib.gdpd_get_step(self.ctx.h, self.sid, C.byref(s)))
        return s.value

    def set_tuning(self, brick=(0, 0, 0), resort_every=-1, graph_steps=-1):
        self.ctx.check(self.lib.gdpd_set_tuning(self.ctx.h, self.sid, int(brick[0]), int(brick[1]),
                                                int(brick[2]), int(resort_every), int(graph_steps)))

    def set_capacity(self, smax=0, omax=0):
        """LDS capacities of the bricks (staged, owned beads; 0 = from the state)."""
        self.ctx.check(self.lib.gdpd_set_capacity(self.ctx.h, self.sid, int(smax), int(omax)))

    def tuning(self):
        out = np.zeros(8, dtype=np.int32)
        self.ctx.check(self.lib.gdpd_get_tuning(self.ctx.h, self.sid, _ptr(out, C.c_int)))
        keys = ["bx", "by", "bz", "smax", "omax", "resort_every", "graph_steps", "nbrick"]
        return dict(zip(keys, out.tolist()))
