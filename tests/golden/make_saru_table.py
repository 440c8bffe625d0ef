"""Writes tests/golden/saru_table.npz: Saru bits and both uniform conversions for
a table of (tag1, tag2, seed, step) tuples, including edge values.

The table is produced by the numpy restatement (tests/saru_numpy.py) and
checked bit for bit against the C oracle before it is written.  The reference
ships no fixtures; this table pins later changes of either restatement and is
what the GPU Saru test compares against.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from saru_numpy import saru_bits, uniform_f32, uniform_f64  # noqa: E402

from oracle import pyoracle  # noqa: E402


def table():
    rng = np.random.Generator(np.random.PCG64(20241203))
    n = 4096
    t1 = rng.integers(1, 2 ** 31, n, dtype=np.int64).astype(np.uint32)
    t2 = rng.integers(1, 2 ** 31, n, dtype=np.int64).astype(np.uint32)
    lo, hi = np.minimum(t1, t2), np.maximum(t1, t2)
    steps = rng.integers(0, 2 ** 31, n, dtype=np.int64).astype(np.int32)
    seeds = np.full(n, 42, dtype=np.uint32)
    edge = [(1, 2, 0, 0), (1, 1, 0, 0), (0, 0, 0, 0), (0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF, -1),
            (1, 24000, 42, 0), (23999, 24000, 42, 1000), (5, 7, 123456, 2 ** 31 - 1), (2 ** 31 - 1, 2 ** 31, 7, 17)]
    for a, b, sd, st in edge:
        lo = np.append(lo, np.uint32(a))
        hi = np.append(hi, np.uint32(b))
        seeds = np.append(seeds, np.uint32(sd))
        steps = np.append(steps, np.int32(st))
    return lo, hi, seeds, steps


def main():
    t1, t2, seeds, steps = table()
    s = np.array([saru_bits(t1[k:k + 1], t2[k:k + 1], seeds[k], steps[k:k + 1])[0] for k in range(len(t1))],
                 dtype=np.uint32)
    r32, r64 = uniform_f32(s), uniform_f64(s)
    for sd in np.unique(seeds):
        m = seeds == sd
        so, ro32, ro64 = pyoracle.saru_batch(t1[m], t2[m], int(sd), steps[m])
        assert np.array_equal(so, s[m]), "numpy and C oracle Saru bits differ"
        assert np.array_equal(ro32.view(np.uint32), r32[m].view(np.uint32))
        assert np.array_equal(ro64.view(np.uint64), r64[m].view(np.uint64))
    np.savez_compressed(os.path.join(HERE, "saru_table.npz"), tag1=t1, tag2=t2, seed=seeds, step=steps, bits=s,
                        r32=r32, r64=r64)
    print(f"wrote {len(s)} rows")


if __name__ == "__main__":
    main()
