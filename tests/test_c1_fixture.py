"""The committed C1 profile fixture is the oracle's and is self-consistent:
its stored summary equals a reduction of its stored mean profile's shape, the
plateau is Warren's liquid density for A = -40, B = 25 (about 6.05-6.1) and the
slab holds all 8,000 beads."""
import os

import numpy as np

import c1_profile as P

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c1_profile.npz")


def test_fixture_is_a_slab_of_all_beads():
    d = np.load(GOLD)
    prof = d["profile"]
    assert len(prof) == 80 and int(d["nframes"]) == (P.LAST - P.FIRST) // P.EVERY + 1
    # integral of the density over the box = bead count
    assert abs(prof.sum() * P.BIN * 10.0 * 10.0 - 8000.0) < 1e-6
    assert 5.9 < float(d["plateau"]) < 6.25 and float(d["plateau_se"]) < 0.02
    # slab width: 8,000 beads at the plateau density in a 10 x 10 cross-section
    assert abs(float(d["width"]) - 8000.0 / (100.0 * float(d["plateau"]))) < 0.5
    xl, xr = P.half_crossings(prof, d["centres"])
    assert abs((xr - xl) - float(d["width"])) < 0.05
