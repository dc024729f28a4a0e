"""Counter-based RNG and libm-free elementary functions of the oracle (TEST INFRASTRUCTURE).

The reference seeds every random stream from the clock (curand_rand_mat.cu:42 time(NULL),
stpm.c:1123 srand(time(NULL)), CPU Spectral.cpp:120 srand(clock())), so none of its draws can be
reproduced. The build replaces them with Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel
random numbers: as easy as 1, 2, 3", SC'11), keyed by (seed, stream) and indexed by counters, so a
draw is a pure function of its indices and is identical in the oracle, the C++ host library and the
HIP kernels regardless of thread count.

det_log/det_exp use only IEEE-754 +, -, *, / and exact exponent manipulation (frexp/ldexp), with a
fixed operation order, so the oracle (numpy float64) and the library (C++ compiled with
-ffp-contract=off) produce the same bits. They are written from the series
  log(m) = 2*atanh(s) = 2*(s + s^3/3 + s^5/5 + ...), s = (m-1)/(m+1), m in [sqrt(1/2), sqrt(2))
  exp(r) = sum r^n/n!, |r| <= ln2/2, after x = k*ln2 + r.
They are accurate to a few ulp, which is all the generator needs; bit-agreement is what matters.
"""
from __future__ import annotations

import numpy as np

M0 = np.uint64(0xD2511F53)
M1 = np.uint64(0xCD9E8D57)
W0 = np.uint32(0x9E3779B9)
W1 = np.uint32(0xBB67AE85)
MASK32 = np.uint64(0xFFFFFFFF)

# stream identifiers (third counter word) -- shared with csrc/philox.h
STREAM_THETA = 1
STREAM_POW = 2
STREAM_GAMMA = 3
STREAM_SPLIT = 4
STREAM_EDGE = 5
STREAM_BALL = 6
STREAM_OMEGA = 16  # + index of the Gaussian test matrix


def philox4x32(c0, c1, c2, c3, k0, k1, rounds: int = 10):
    """Vectorised Philox4x32-R. Counters/keys are uint32 arrays (broadcastable)."""
    c0 = np.asarray(c0, np.uint32).astype(np.uint64)
    c1 = np.asarray(c1, np.uint32).astype(np.uint64)
    c2 = np.asarray(c2, np.uint32).astype(np.uint64)
    c3 = np.asarray(c3, np.uint32).astype(np.uint64)
    k0 = np.asarray(k0, np.uint32).astype(np.uint64)
    k1 = np.asarray(k1, np.uint32).astype(np.uint64)
    for rnd in range(rounds):
        if rnd > 0:
            k0 = (k0 + np.uint64(W0)) & MASK32
            k1 = (k1 + np.uint64(W1)) & MASK32
        p0 = M0 * c0
        p1 = M1 * c2
        hi0, lo0 = p0 >> np.uint64(32), p0 & MASK32
        hi1, lo1 = p1 >> np.uint64(32), p1 & MASK32
        c0, c1, c2, c3 = (hi1 ^ c1 ^ k0), lo1, (hi0 ^ c3 ^ k1), lo0
    return tuple(x.astype(np.uint32) for x in (c0, c1, c2, c3))


def seed_key(seed: int):
    s = int(seed) & 0xFFFFFFFFFFFFFFFF
    return np.uint32(s & 0xFFFFFFFF), np.uint32(s >> 32)


def u53(w0, w1):
    """Uniform double in [0, 1) from two 32-bit words (53 significant bits)."""
    a = (np.asarray(w0, np.uint64) >> np.uint64(5)).astype(np.float64)
    b = (np.asarray(w1, np.uint64) >> np.uint64(6)).astype(np.float64)
    return (a * 67108864.0 + b) * (1.0 / 9007199254740992.0)


def u53_open0(w0, w1):
    """Uniform double in (0, 1]."""
    a = (np.asarray(w0, np.uint64) >> np.uint64(5)).astype(np.float64)
    b = (np.asarray(w1, np.uint64) >> np.uint64(6)).astype(np.float64)
    return (a * 67108864.0 + b + 1.0) * (1.0 / 9007199254740992.0)


def draw(seed: int, stream: int, i0, i1, sub=0):
    """Four 32-bit words for counter (i0, i1, stream, sub) under key seed."""
    k0, k1 = seed_key(seed)
    return philox4x32(i0, i1, np.uint32(stream), sub, k0, k1)


# ---- libm-free log / exp -------------------------------------------------------------------
LN2_HI = 6.93147180369123816490e-01  # high part of ln 2 (exact in 32 bits)
LN2_LO = 1.90821492927058770002e-10
INV_LN2 = 1.44269504088896338700e+00
SQRT_HALF = 0.70710678118654752440

_LOG_TERMS = 12   # s^(2n+1)/(2n+1), n = 0..11 : |s| <= 0.1716 -> error < 1e-19
_EXP_TERMS = 14   # r^n/n!, n = 0..13 : |r| <= 0.3466 -> error < 1e-17


def det_log(x):
    """log(x) for x > 0 (float64 array) with a fixed IEEE operation sequence."""
    x = np.asarray(x, np.float64)
    m, e = np.frexp(x)                  # x = m * 2^e, m in [0.5, 1)
    small = m < SQRT_HALF
    m = np.where(small, m * 2.0, m)     # exact
    e = np.where(small, e - 1, e).astype(np.float64)
    s = (m - 1.0) / (m + 1.0)
    s2 = s * s
    # Horner on p(t) = sum_{n} t^n/(2n+1), highest term first
    p = np.full_like(s, 1.0 / (2 * _LOG_TERMS - 1))
    for n in range(_LOG_TERMS - 2, -1, -1):
        p = p * s2 + 1.0 / (2 * n + 1)
    logm = 2.0 * s * p
    return e * LN2_HI + (e * LN2_LO + logm)


def det_exp(x):
    """exp(x) for finite x (float64 array) with a fixed IEEE operation sequence."""
    x = np.asarray(x, np.float64)
    k = np.floor(x * INV_LN2 + 0.5)
    r = (x - k * LN2_HI) - k * LN2_LO
    p = np.full_like(r, 1.0 / _fact(_EXP_TERMS - 1))
    for n in range(_EXP_TERMS - 2, -1, -1):
        p = p * r + 1.0 / _fact(n)
    ki = np.clip(k, -1100, 1100).astype(np.int64)
    return np.ldexp(p, ki)


def _fact(n: int) -> float:
    f = 1.0
    for i in range(2, n + 1):
        f *= i
    return f


def det_sqrt(x):
    """IEEE sqrt is correctly rounded everywhere; kept as a named step for symmetry with C."""
    return np.sqrt(np.asarray(x, np.float64))
