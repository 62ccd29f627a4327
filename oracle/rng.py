"""numpy restatement of the seeded transmission-data sampler (kv_poisson_data,
kevernals_amd/csrc/kv_rng.hip).

TEST INFRASTRUCTURE ONLY: the checker of the HIP kernel; the product path
never imports it.  The reference simulates no data, so there is nothing of
its to follow; the restatement is pinned in tests/test_rng.py by:

* Philox4x64-10 against numpy's own Philox bit generator (an independent
  implementation of the published algorithm of Salmon et al., SC 2011);
* exp / log / log k! against the math library (within a few ulps);
* the inversion branch against the Poisson CDF of scipy on the same
  uniforms, and both branches against the Poisson distribution (moments and
  chi-square goodness of fit at several means).

Vectorised over elements: every element's draw runs the same IEEE
operations in the same order as the kernel, so counts and line integrals are
bit-identical to the device's.
"""
from __future__ import annotations

import numpy as np

M0, M1 = np.uint64(0xD2E7470EE14C6C93), np.uint64(0xCA5A826395121157)
W0, W1 = np.uint64(0x9E3779B97F4A7C15), np.uint64(0xBB67AE8584CAA73B)
KEY1 = 0x6B76
LN2 = 0.6931471805599453094
LN2_HI, LN2_LO = float.fromhex("0x1.62e42fp-1"), float.fromhex("0x1.df473de6af279p-26")
HALF_LOG_2PI = 0.91893853320467274178
_M32 = np.uint64(0xFFFFFFFF)


def _mul128(a, b):
    """(hi, lo) 64-bit halves of the 128-bit products a*b (uint64 arrays)."""
    a_lo, a_hi = a & _M32, a >> np.uint64(32)
    b_lo, b_hi = b & _M32, b >> np.uint64(32)
    ll = a_lo * b_lo
    lh = a_lo * b_hi
    hl = a_hi * b_lo
    hh = a_hi * b_hi
    mid = (ll >> np.uint64(32)) + (lh & _M32) + (hl & _M32)
    hi = hh + (lh >> np.uint64(32)) + (hl >> np.uint64(32)) + (mid >> np.uint64(32))
    lo = a * b
    return hi, lo


def philox4x64_10(c0, c1, c2, c3, k0, k1):
    """Philox4x64 with 10 rounds on arrays of counters; returns 4 arrays."""
    with np.errstate(over="ignore"):
        c = [np.asarray(x, dtype=np.uint64).copy() for x in (c0, c1, c2, c3)]
        k0 = np.asarray(k0, dtype=np.uint64).copy()
        k1 = np.asarray(k1, dtype=np.uint64).copy()
        for _ in range(10):
            hi0, lo0 = _mul128(np.broadcast_to(M0, c[0].shape), c[0])
            hi1, lo1 = _mul128(np.broadcast_to(M1, c[2].shape), c[2])
            c = [hi1 ^ c[1] ^ k0, lo1, hi0 ^ c[3] ^ k1, lo0]
            k0 = k0 + W0
            k1 = k1 + W1
    return c


class Streams:
    """Uniform streams of elements idx under a seed: block j of element i is
    Philox(counter (i, j, 0, 0), key (seed, 0x6b76)), four uniforms each."""

    def __init__(self, idx, seed):
        self.idx = np.asarray(idx, dtype=np.uint64)
        self.seed = np.uint64(seed)
        self.pos = np.zeros(self.idx.shape, dtype=np.int64)  # uniforms used

    def next(self, mask):
        """One uniform for each element where mask holds (others: NaN)."""
        out = np.full(self.idx.shape, np.nan)
        sel = np.nonzero(mask)[0]
        if sel.size:
            blk = (self.pos[sel] // 4).astype(np.uint64)
            w = philox4x64_10(self.idx[sel], blk, np.zeros_like(blk), np.zeros_like(blk),
                              np.full(sel.shape, self.seed, np.uint64), np.full(sel.shape, KEY1, np.uint64))
            lane = self.pos[sel] % 4
            word = np.choose(lane, w)
            out[sel] = (word >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
            self.pos[sel] += 1
        return out


def kv_log(x):
    x = np.asarray(x, dtype=np.float64)
    m, e = np.frexp(x)
    low = m < 0.70710678118654752440
    m = np.where(low, m * 2.0, m)
    e = np.where(low, e - 1, e).astype(np.float64)
    s = (m - 1.0) / (m + 1.0)
    s2 = s * s
    t = np.full(x.shape, 1.0 / 23.0)
    for k in range(21, 0, -2):
        t = 1.0 / float(k) + s2 * t
    return e * LN2 + 2.0 * s * t


def kv_exp(x):
    x = np.asarray(x, dtype=np.float64)
    xc = np.clip(x, -745.0, 709.0)
    n = np.rint(xc / LN2)
    r = (xc - n * LN2_HI) - n * LN2_LO
    t = np.ones(x.shape)
    for k in range(18, 0, -1):
        t = 1.0 + r * t / float(k)
    out = np.ldexp(t, n.astype(np.int64))
    out = np.where(x < -745.0, 0.0, out)
    return np.where(x > 709.0, np.inf, out)


def kv_log_factorial(k):
    k = np.asarray(k, dtype=np.float64)
    out = np.zeros(k.shape)
    small = k < 32.0
    if small.any():
        acc = np.zeros(k.shape)
        for i in range(2, 32):
            acc = np.where(small & (float(i) <= k), acc + kv_log(np.full(k.shape, float(i))), acc)
        out = np.where(small, acc, out)
    n = k + 1.0
    n2 = n * n
    series = (1.0 / 12.0 - (1.0 / 360.0 - (1.0 / 1260.0 - (1.0 / 1680.0) / n2) / n2) / n2) / n
    big = (n - 0.5) * kv_log(np.where(small, 2.0, n)) - n + HALF_LOG_2PI + series
    return np.where(small, out, big)


def poisson(lam, idx, seed):
    """Counts for means lam (float64 array) of elements idx under a seed."""
    lam = np.asarray(lam, dtype=np.float64)
    st = Streams(idx, seed)
    k = np.zeros(lam.shape)
    pos = lam > 0.0
    # inversion below 10
    inv = pos & (lam < 10.0)
    if inv.any():
        u = st.next(inv)
        p = kv_exp(-np.where(inv, lam, 0.0))
        f = p.copy()
        live = inv & (u > f)
        kk = np.zeros(lam.shape)
        while live.any():
            kk = np.where(live, kk + 1.0, kk)
            p = np.where(live, p * lam / np.where(live, kk, 1.0), p)
            f = np.where(live, f + p, f)
            live = live & (u > f) & (kk < 1000.0)
        k = np.where(inv, kk, k)
    # PTRS from 10
    pt = pos & (lam >= 10.0)
    if pt.any():
        L = np.where(pt, lam, 10.0)
        loglam = kv_log(L)
        slam = kv_exp(0.5 * loglam)
        b = 0.931 + 2.53 * slam
        a = -0.059 + 0.02483 * b
        invalpha = 1.1239 + 1.1328 / (b - 3.4)
        vr = 0.9277 - 3.6224 / (b - 2.0)
        live = pt.copy()
        res = np.zeros(lam.shape)
        for _ in range(1000):
            if not live.any():
                break
            U = st.next(live) - 0.5
            V = st.next(live)
            us = 0.5 - np.abs(U)
            with np.errstate(invalid="ignore", divide="ignore"):
                kk = np.floor((2.0 * a / us + b) * U + L + 0.43)
            fast = live & (us >= 0.07) & (V <= vr)
            res = np.where(fast, kk, res)
            live = live & ~fast
            reject = live & ((kk < 0.0) | ((us < 0.013) & (V > us)))
            test = live & ~reject
            if test.any():
                with np.errstate(invalid="ignore", divide="ignore"):
                    lhs = kv_log(np.where(test, V, 1.0)) + kv_log(invalpha) - kv_log(
                        np.where(test, a / (us * us) + b, 1.0))
                    rhs = -L + kk * loglam - kv_log_factorial(np.where(test, kk, 0.0))
                acc = test & (lhs <= rhs)
                res = np.where(acc, kk, res)
                live = live & ~acc
        res = np.where(live, np.floor(L), res)
        k = np.where(pt, res, k)
    return k


def poisson_data(ell, i0, seed):
    """kv_poisson_data: -ln(clamp(I, 1, i0) / i0), I ~ Poisson(i0 exp(-ell))
    per element (index = position in the flattened array), float32."""
    ell = np.asarray(ell, dtype=np.float32)
    flat = ell.ravel().astype(np.float64)
    i0 = float(np.float32(i0))
    lam = i0 * kv_exp(-flat)
    cnt = poisson(lam, np.arange(flat.size, dtype=np.uint64), seed)
    c = np.minimum(np.maximum(cnt, 1.0), i0)
    return (-kv_log(c / i0)).astype(np.float32).reshape(ell.shape)
