"""gpumdpd_amd: MI355X-native (gfx950 HIP) DPD/mDPD particle step behind a C ABI.

The compute lives in libgdpd.so (gpumdpd_amd/csrc, include/gdpd.h).  This
package is the Python front used by tests and the benchmark: it loads that
library through ctypes and never falls back to CPU code.
"""
from ._lib import PREC_F32, PREC_F64, GdpdError  # noqa: F401
from .system import Context, System, make_desc, system_bytes  # noqa: F401


def system_from_config(ctx, cfg, precision=PREC_F32, log_every=100, check_every=1, upload=True):
    """Create, parametrise and upload a System from a configs.* dictionary."""
    desc = make_desc(cfg["N"], cfg["ntypes"], cfg["lo"], cfg["hi"], cfg["dt"], cfg["global_cut"],
                     cfg["skin"], cfg["max_rho"], check_every, log_every, precision)
    s = System(ctx, desc)
    if cfg["style"] == "mdpd":
        s.set_pair_mdpd(cfg["temp"], cfg["rc"], cfg["rd"], cfg["saru_seed"], cfg["coeff"])
    else:
        s.set_pair_dpd(cfg["temp"], cfg["rc"], cfg["saru_seed"], cfg["coeff"])
    if "bonds" in cfg:
        s.set_bonds(cfg["bonds"], cfg["n_bond_types"], cfg["bond_coeff"])
    if "angles" in cfg:
        s.set_angles(cfg["angles"], cfg["n_angle_types"], cfg["angle_coeff"])
    if upload:
        s.upload(cfg["id"], cfg["type"], cfg["pos"], cfg["vel"], cfg["masses"])
    return s
