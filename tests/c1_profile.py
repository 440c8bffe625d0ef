"""Density profile across the C1 liquid-vapour slab, in the manner of the
reference's `compute rho` (compute/compute_rho.cu:111-214: beads of a group
inside an axis-aligned box, divided by its volume), taken for 80 slabs of
width 0.5 along x.  Each sampled frame is first centred on the slab's
periodic centre of mass along x (circular mean), so the slab's slow drift does
not smear the profile.  Shared by the fixture script (oracle run) and the GPU
test, so both reduce their frames identically.
"""
import numpy as np

BIN = 0.5
FIRST, LAST, EVERY = 3010, 5000, 10   # sampled steps: 3010, 3020, ..., 5000
NBLOCK = 10                           # block averages for standard errors
CORE = 3.0                            # plateau: bins within +-CORE of the centre


def frame_profile(pos, lo, hi):
    """Counts per 0.5-wide x slab of one frame, centred on the slab."""
    L = hi[0] - lo[0]
    x = pos[:, 0] - lo[0]
    th = 2.0 * np.pi * x / L
    xc = (np.arctan2(np.sin(th).sum(), np.cos(th).sum()) * L / (2.0 * np.pi)) % L
    xs = (x - xc + 0.5 * L) % L
    nb = int(round(L / BIN))
    counts, _ = np.histogram(xs, bins=nb, range=(0.0, L))
    area = (hi[1] - lo[1]) * (hi[2] - lo[2])
    return counts / (BIN * area)


def half_crossings(prof, centres):
    """x where the profile falls to half its plateau, left and right of the centre."""
    mid = 0.5 * (centres[0] + centres[-1] + BIN)
    core = np.abs(centres - mid) < CORE
    half = 0.5 * prof[core].mean()
    i_mid = int(np.argmin(np.abs(centres - mid)))
    # walk outwards from the centre to the first bin below half
    r = i_mid
    while r + 1 < len(prof) and prof[r + 1] >= half:
        r += 1
    xr = centres[r] + (prof[r] - half) / (prof[r] - prof[r + 1]) * BIN
    l = i_mid
    while l - 1 >= 0 and prof[l - 1] >= half:
        l -= 1
    xl = centres[l] - (prof[l] - half) / (prof[l] - prof[l - 1]) * BIN
    return xl, xr


def summarize(frames, lo, hi):
    """frames: (nsample, nbins) densities.  Returns the mean profile and the
    plateau density, slab width and vapour density with block standard errors."""
    frames = np.asarray(frames)
    nb = frames.shape[1]
    centres = (np.arange(nb) + 0.5) * BIN
    blocks = np.array_split(frames, NBLOCK)
    stats = []
    for b in blocks:
        p = b.mean(0)
        mid = 0.5 * (hi[0] - lo[0])
        core = np.abs(centres - mid) < CORE
        vap = np.abs(centres - mid) > 0.5 * (hi[0] - lo[0]) - 5.0
        xl, xr = half_crossings(p, centres)
        stats.append((p[core].mean(), xr - xl, p[vap].mean()))
    stats = np.array(stats)
    mean = stats.mean(0)
    se = stats.std(0, ddof=1) / np.sqrt(len(stats))
    return {"profile": frames.mean(0), "centres": centres, "plateau": mean[0], "plateau_se": se[0],
            "width": mean[1], "width_se": se[1], "vapour": mean[2], "vapour_se": se[2]}
