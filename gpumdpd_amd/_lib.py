"""ctypes binding of libgdpd.so (the C ABI in include/gdpd.h).

Loads the in-tree library and nothing else: there is no CPU fallback.  If the
library is missing or cannot be loaded the import raises.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgdpd.so")

GDPD_OK = 0
ERR_NAMES = {1: "GDPD_ERR_ARG", 2: "GDPD_ERR_HIP", 3: "GDPD_ERR_OVERFLOW", 4: "GDPD_ERR_STATE",
             5: "GDPD_ERR_NOMEM", 6: "GDPD_ERR_NAN"}
PREC_F32, PREC_F64 = 0, 1
STYLE_DPD, STYLE_MDPD = 1, 2
THERMO_LEN = 13
FIX_KINDS = {"nve": 1, "move": 2, "wall_reflect": 3, "two_wall": 4, "spring": 5, "recenter": 6, "rvcm": 7}
COMPUTE_KINDS = {"temp": 1, "msd": 2, "rho": 3, "tf": 4, "wall_contact": 5}
HOOKS = {"initial_integrate": 0, "post_integrate": 1, "post_force": 2, "final_integrate": 3, "end_of_step": 4}
MAX_FIXES = MAX_COMPUTES = 8
RECORD_LEN, RECORD_CAP = 6, 1024
STEP_TS_CAP = 256


class SystemDesc(C.Structure):
    _fields_ = [
        ("n_beads", C.c_int), ("n_types", C.c_int),
        ("lo", C.c_double * 3), ("hi", C.c_double * 3),
        ("dt", C.c_double), ("global_cut", C.c_double), ("skin", C.c_double),
        ("max_rho", C.c_double), ("check_every", C.c_int), ("log_every", C.c_int),
        ("precision", C.c_int),
    ]


class Fix(C.Structure):
    _fields_ = [
        ("kind", C.c_int), ("group", C.c_uint32), ("group2", C.c_uint32), ("group3", C.c_uint32),
        ("every", C.c_int), ("dir", C.c_int), ("flags", C.c_int), ("p", C.c_double * 5),
    ]


class Compute(C.Structure):
    _fields_ = [
        ("kind", C.c_int), ("group", C.c_uint32), ("group2", C.c_uint32), ("every", C.c_int),
        ("dir", C.c_int), ("side", C.c_int), ("p", C.c_double * 6),
    ]


_P = C.c_void_p
_I = C.c_int
_PI = C.POINTER(C.c_int)
_PD = C.POINTER(C.c_double)

# name: (restype, argtypes) -- one line per declaration of include/gdpd.h
PROTOTYPES = {
    "gdpd_create": (_I, [C.POINTER(_P), _I]),
    "gdpd_destroy": (_I, [_P]),
    "gdpd_last_error": (C.c_char_p, [_P]),
    "gdpd_system_bytes": (C.c_int64, [C.POINTER(SystemDesc), _I]),
    "gdpd_sys_create": (_I, [_P, C.POINTER(SystemDesc), _PI]),
    "gdpd_sys_destroy": (_I, [_P, _I]),
    "gdpd_upload": (_I, [_P, _I, _PI, _PI, _PD, _PD, _PD]),
    "gdpd_download": (_I, [_P, _I, _PD, _PD, _PI, _PD]),
    "gdpd_download_pe_virial": (_I, [_P, _I, _PD, _PD, _PD]),
    "gdpd_set_pair_dpd": (_I, [_P, _I, C.c_double, C.c_double, C.c_uint32, _PD]),
    "gdpd_set_pair_mdpd": (_I, [_P, _I, C.c_double, C.c_double, C.c_double, C.c_uint32, _PD]),
    "gdpd_set_bonds": (_I, [_P, _I, _I, _I, _PI, _PD]),
    "gdpd_set_angles": (_I, [_P, _I, _I, _I, _PI, _PD]),
    "gdpd_neigh_build": (_I, [_P, _I]),
    "gdpd_neigh_check": (_I, [_P, _I, _PI]),
    "gdpd_neigh_download": (_I, [_P, _I, _PI, _PI, _I]),
    "gdpd_pair_compute": (_I, [_P, _I, C.c_uint32, _I]),
    "gdpd_bond_compute": (_I, [_P, _I, C.c_uint32, _I]),
    "gdpd_angle_compute": (_I, [_P, _I, C.c_uint32, _I]),
    "gdpd_nve_initial": (_I, [_P, _I]),
    "gdpd_nve_final": (_I, [_P, _I]),
    "gdpd_thermo": (_I, [_P, _I, _PD]),
    "gdpd_run": (_I, [_P, _I, _I]),
    "gdpd_zero_step": (_I, [_P, _I]),
    "gdpd_advance": (_I, [_P, _I, _I]),
    "gdpd_finish": (_I, [_P, _I]),
    "gdpd_thermo_rows": (_I, [_P, _I, _PD, _I, _PI]),
    "gdpd_get_step": (_I, [_P, _I, C.POINTER(C.c_int64)]),
    "gdpd_set_fixes": (_I, [_P, _I, C.POINTER(Fix), _I]),
    "gdpd_fix_hook": (_I, [_P, _I, _I, _I, C.c_uint32, _PD]),
    "gdpd_set_computes": (_I, [_P, _I, C.POINTER(Compute), _I]),
    "gdpd_compute_eval": (_I, [_P, _I, _I, _PD]),
    "gdpd_records": (_I, [_P, _I, _I, _PD, _I, _PI]),
    "gdpd_batch_advance": (_I, [_P, _PI, _I, _I]),
    "gdpd_set_tuning": (_I, [_P, _I, _I, _I, _I, _I, _I]),
    "gdpd_get_tuning": (_I, [_P, _I, _PI]),
    "gdpd_kernel_timing": (_I, [_P, _I]),
    "gdpd_kernel_timing_get": (_I, [_P, _PD, C.POINTER(C.c_int64), _PD]),
    "gdpd_kernel_phases": (_I, [_P, _PD]),
    "gdpd_sync": (_I, [_P]),
    "gdpd_admission": (_I, [_P, _I]),
    "gdpd_thermo_noalloc": (_I, [_P, _I, _PD]),
    "gdpd_set_capacity": (_I, [_P, _I, _I, _I]),
    "gdpd_step_times": (_I, [_P, _PI, _I, C.POINTER(C.c_uint64), C.POINTER(C.c_int64), _PD]),
    "gdpd_saru_eval": (_I, [_P, _I, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.c_uint32,
                            C.POINTER(C.c_int32), C.POINTER(C.c_uint32), C.POINTER(C.c_float), _PD]),
}

_lib = None


def load():
    """Load libgdpd.so from the package directory (raises if it is absent)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class GdpdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERR_NAMES.get(code, code)}: {msg}")
        self.code = code
