"""Independent numpy-uint32 restatement of the Saru pair hash.

Written from force/pair_dpd.cu:212-236 separately from the C oracle, so the two
must agree bit for bit.  Vectorised over arrays of tags and steps.
"""
import numpy as np


def _u(x):
    return np.asarray(x, dtype=np.uint32)


def _asr(v, n):
    return (v.view(np.int32) >> np.int32(n)).view(np.uint32)


def saru_bits(tag1, tag2, seed, step):
    with np.errstate(over="ignore"):
        s1 = _u(tag1).copy()
        s2 = _u(tag2).copy()
        s3 = (np.uint32(seed) + np.asarray(step, dtype=np.int64).astype(np.uint32)).astype(np.uint32)
        s3 ^= (s1 << np.uint32(7)) ^ (s2 >> np.uint32(6))
        s2 += (s1 >> np.uint32(4)) ^ (s3 >> np.uint32(15))
        s1 ^= (s2 << np.uint32(9)) + (s3 << np.uint32(8))
        s3 ^= np.uint32(0xA5366B4D) * ((s2 >> np.uint32(11)) ^ (s1 << np.uint32(1)))
        s2 += np.uint32(0x72BE1579) * ((s1 << np.uint32(4)) ^ (s3 >> np.uint32(16)))
        s1 ^= np.uint32(0x3F38A6ED) * ((s3 >> np.uint32(5)) ^ _asr(s2, 22))
        s2 += s1 * s3
        s1 += s3 ^ (s2 >> np.uint32(2))
        s2 ^= _asr(s2, 17)
        state = np.uint32(0x79DEDEA3) * (s1 ^ _asr(s1, 14))
        wstate = (state + s2) ^ _asr(state, 8)
        state = state + (wstate * (wstate ^ np.uint32(0xDDDF97F5)))
        wstate = np.uint32(0xABCB96F7) + (wstate >> np.uint32(1))
        total = np.uint32(0)
        for _ in range(8):
            total = np.uint32(total + np.uint32(0x9E3779B9))
            state += ((wstate << np.uint32(4)) + np.uint32(0xA341316C)) ^ (wstate + total) ^ (
                (wstate >> np.uint32(5)) + np.uint32(0xC8013EA4))
            wstate += ((state << np.uint32(4)) + np.uint32(0xAD90777D)) ^ (state + total) ^ (
                (state >> np.uint32(5)) + np.uint32(0x7E95761E))
        v = (state ^ (state >> np.uint32(26))) + wstate
        return ((v ^ (v >> np.uint32(20))) * np.uint32(0x6957F5A7)).astype(np.uint32)


def uniform_f32(s):
    u = s.astype(np.float32) * np.float32(2.0 ** -32)
    return np.float32(np.float32(1.7320508075688772935274463) * (u * np.float32(2.0) - np.float32(1.0)))


def uniform_f64(s):
    u = (s.astype(np.float32) * np.float32(2.0 ** -32)).astype(np.float64)
    return 1.7320508075688772935274463 * (u * 2.0 - 1.0)
