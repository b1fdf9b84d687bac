"""Counter-based normal variates for the stochastic GPE (test infrastructure only).

MATLAB's ``randn`` stream (SGPEtask/solve_split.m:59) cannot be reproduced, so the
stochastic solver draws its noise from Philox4x32-10 (Salmon, Moraes, Dror, Shaw,
"Parallel random numbers: as easy as 1, 2, 3", SC'11) keyed by a seed, one counter
per (step, grid point): counter = (point, step_lo, step_hi, 0), key = (seed_lo,
seed_hi).  The four 32-bit outputs make two uniforms in (0, 1) with 53 bits each,
u = ((w0 << 32 | w1) >> 11 + 1/2) 2^-53, and the Box-Muller transform turns them into
the real and imaginary parts of one complex normal variate:
    re = sqrt(-2 ln u1) cos(2 pi u2),  im = sqrt(-2 ln u1) sin(2 pi u2).
The device generator follows the same definition.
"""
import numpy as np

M0, M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
W0, W1 = np.uint32(0x9E3779B9), np.uint32(0xBB67AE85)
MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10(c0, c1, c2, c3, k0, k1):
    """Ten Philox rounds on uint32 arrays (broadcast); returns four uint32 arrays."""
    c0, c1, c2, c3 = (np.asarray(x, dtype=np.uint32) for x in (c0, c1, c2, c3))
    k0 = np.asarray(k0, dtype=np.uint32)
    k1 = np.asarray(k1, dtype=np.uint32)
    with np.errstate(over="ignore"):
        for r in range(10):
            p0 = M0 * c0.astype(np.uint64)
            p1 = M1 * c2.astype(np.uint64)
            hi0, lo0 = (p0 >> np.uint64(32)).astype(np.uint32), (p0 & MASK).astype(np.uint32)
            hi1, lo1 = (p1 >> np.uint64(32)).astype(np.uint32), (p1 & MASK).astype(np.uint32)
            c0, c1, c2, c3 = hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0
            if r < 9:
                k0 = (k0 + W0).astype(np.uint32)
                k1 = (k1 + W1).astype(np.uint32)
    return c0, c1, c2, c3


def _u53(a, b):
    w = (a.astype(np.uint64) << np.uint64(32)) | b.astype(np.uint64)
    return ((w >> np.uint64(11)).astype(np.float64) + 0.5) * (2.0 ** -53)


def complex_normals(n, step, seed):
    """n complex standard normal pairs (re, im each N(0, 1)) for one step."""
    idx = np.arange(n, dtype=np.uint64)
    w0, w1, w2, w3 = philox4x32_10(idx.astype(np.uint32), np.uint32(step & 0xFFFFFFFF),
                                   np.uint32((step >> 32) & 0xFFFFFFFF), np.uint32(0),
                                   np.uint32(seed & 0xFFFFFFFF), np.uint32((seed >> 32) & 0xFFFFFFFF))
    u1, u2 = _u53(w0, w1), _u53(w2, w3)
    r = np.sqrt(-2.0 * np.log(u1))
    th = 2.0 * np.pi * u2
    return r * np.cos(th), r * np.sin(th)
