"""The C1 mDPD liquid-vapour slab's density profile on the GPU (FP32 pair
force) against the FP64 oracle's (tests/golden/c1_profile.npz, written by
tests/golden/make_c1_profile.py): the same 5,000-step run from the same state,
sampled every 10 steps over steps 3010-5000 and reduced by tests/c1_profile.py.

The two runs are independent chaotic trajectories, so they agree
statistically: the liquid plateau density and the slab width (between the
half-plateau crossings) must match within 4 combined block standard errors,
and the vapour must stay empty."""
import os

import numpy as np
import pytest

import c1_profile as P
import gpumdpd_amd as g
from gpumdpd_amd import configs

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c1_profile.npz")


def test_c1_profile_matches_oracle():
    ref = np.load(GOLD)
    cfg = configs.config_c1()
    lo, hi = np.asarray(cfg["lo"], float), np.asarray(cfg["hi"], float)
    print("case: C1 slab 5,000 steps, FP32 force", flush=True)
    ctx = g.Context(0)
    try:
        s = g.system_from_config(ctx, cfg, precision=g.PREC_F32, log_every=1000)
        s.zero_step()  # leaves x(1)
        s.advance(P.FIRST - 1)
        frames = [P.frame_profile(s.download()["pos"], lo, hi)]
        for _ in range((P.LAST - P.FIRST) // P.EVERY):
            s.advance(P.EVERY)
            frames.append(P.frame_profile(s.download()["pos"], lo, hi))
        s.finish()
        s.destroy()
    finally:
        ctx.close()
    assert len(frames) == int(ref["nframes"])
    got = P.summarize(frames, lo, hi)
    print(f"GPU plateau {got['plateau']:.4f} +- {got['plateau_se']:.4f}, width {got['width']:.3f} +- "
          f"{got['width_se']:.3f}; oracle {float(ref['plateau']):.4f} +- {float(ref['plateau_se']):.4f}, "
          f"{float(ref['width']):.3f} +- {float(ref['width_se']):.3f}", flush=True)
    tol_p = 4.0 * np.hypot(got["plateau_se"], float(ref["plateau_se"]))
    tol_w = 4.0 * np.hypot(got["width_se"], float(ref["width_se"]))
    assert abs(got["plateau"] - float(ref["plateau"])) < tol_p
    assert abs(got["width"] - float(ref["width"])) < tol_w
    assert got["vapour"] < 0.01 and float(ref["vapour"]) < 0.01
