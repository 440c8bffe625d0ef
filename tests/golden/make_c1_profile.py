"""Writes tests/golden/c1_profile.npz: the x density profile of the C1 mDPD
liquid-vapour slab (8,000 beads at rho = 6 in the centre third of a 40x10x10
box, A = -40, B = 25, rd = 0.75, gamma = 4.5, kT = 1, dt = 0.01, seed 7, Saru
seed 42) from a 5,000-step FP64 run of the CPU oracle, sampled every 10 steps
over steps 3010-5000, with block standard errors (tests/c1_profile.py).
About three minutes on one core.  Run from the repository root:

    python tests/golden/make_c1_profile.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import c1_profile as P  # noqa: E402
from gpumdpd_amd import configs  # noqa: E402
from oracle import pyoracle  # noqa: E402


def main():
    cfg = configs.config_c1()
    lo, hi = np.asarray(cfg["lo"], float), np.asarray(cfg["hi"], float)
    frames = []

    def obs(step, pos):
        if P.FIRST <= step <= P.LAST:
            frames.append(P.frame_profile(pos, lo, hi))

    rows = pyoracle.run_observed(cfg, P.LAST, P.EVERY, obs, log_every=100)
    s = P.summarize(frames, lo, hi)
    out = os.path.join(HERE, "c1_profile.npz")
    np.savez(out, profile=s["profile"], centres=s["centres"], nframes=len(frames),
             plateau=s["plateau"], plateau_se=s["plateau_se"], width=s["width"], width_se=s["width_se"],
             vapour=s["vapour"], vapour_se=s["vapour_se"], thermo=rows)
    print(f"{out}: {len(frames)} frames, plateau {s['plateau']:.4f} +- {s['plateau_se']:.4f}, "
          f"width {s['width']:.3f} +- {s['width_se']:.3f}, vapour {s['vapour']:.4f}")


if __name__ == "__main__":
    main()
