"""gdpd_thermo_noalloc (the allocation-free thermo, partial sums in a static
device array) against gdpd_thermo (per-call device allocation), through the C
ABI: equal bit for bit on two calls in a row, the second on a smaller system,
so a leftover of the first call's partials would show."""
import ctypes as C

import numpy as np
import pytest

import gpumdpd_amd as g
from gpumdpd_amd import configs

pytestmark = pytest.mark.gpu


def thermo_both(ctx, s):
    a = np.empty(13)
    b = np.empty(13)
    ctx.check(ctx.lib.gdpd_thermo(ctx.h, s.sid, a.ctypes.data_as(C.POINTER(C.c_double))))
    ctx.check(ctx.lib.gdpd_thermo_noalloc(ctx.h, s.sid, b.ctypes.data_as(C.POINTER(C.c_double))))
    return a, b


def test_thermo_noalloc_equals_thermo():
    ctx = g.Context(0)
    try:
        big = configs.scaled_dpd(60000, seed=71)
        small = configs.replica_system("mdpd", 7000, 72)
        out = []
        for cfg, n in ((big, 9), (small, 5)):
            print(f"case: thermo_noalloc {cfg['style']} N={cfg['N']}", flush=True)
            s = g.system_from_config(ctx, cfg, precision=g.PREC_F32, log_every=10)
            s.run(n)
            a, b = thermo_both(ctx, s)
            assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), (a, b)
            out.append(b)
            s.destroy()
        assert not np.array_equal(out[0], out[1])
    finally:
        ctx.close()


def test_thermo_noalloc_rejects_unprepared_system():
    from gpumdpd_amd import _lib
    ctx = g.Context(0)
    try:
        s = g.system_from_config(ctx, configs.scaled_dpd(3000, seed=73), precision=g.PREC_F32)
        b = np.empty(13)
        rc = ctx.lib.gdpd_thermo_noalloc(ctx.h, s.sid, b.ctypes.data_as(C.POINTER(C.c_double)))
        assert rc == 4, _lib.ERR_NAMES.get(rc, rc)  # GDPD_ERR_STATE before the first computation
        s.run(2)
        a, b = thermo_both(ctx, s)
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
        s.destroy()
    finally:
        ctx.close()
