"""FP64 CPU oracle of the trajectory-extrapolation passes (kv_traj_gram,
kv_traj_combine).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the HIP
kernels; the product path never uses it.  The passes have no counterpart in
the reference (they serve the accelerated reconstruction built on the mode-A
MLEM update, Reconstructors.h:182-242 in its mode-A form), so this
restatement is pinned by closed forms in tests/test_traj.py: a single-mode
history whose Gram matrix is lambda^(i+j) times one weighted sum, hand-summed
two-voxel cases that fix the weight (sens * newest iterate, not the oldest)
and the orientation of the ratios (newer over older), and linear recurrences
whose extrapolation is exact.

Definitions (iterates x_0..x_h, oldest first, float32 volumes):
    d_i(v) = log(x_{i+1}(v) / x_i(v))      (0 unless both are positive)
    w(v)   = sens(v) * x_h(v)              (0 unless both are positive)
    G_ij   = sum_v w(v) d_i(v) d_j(v)
    y(v)   = x_h(v) * exp(sum_i c_i d_i(v))   (x_h(v) where it is not positive)
"""
from __future__ import annotations

import numpy as np


def log_ratios(iterates):
    xs = [np.asarray(x, dtype=np.float64).ravel() for x in iterates]
    out = []
    for a, b in zip(xs[:-1], xs[1:]):
        ok = (a > 0) & (b > 0)
        d = np.zeros_like(a)
        d[ok] = np.log(b[ok] / a[ok])
        out.append(d)
    return out


def gram(iterates, sens):
    d = log_ratios(iterates)
    xh = np.asarray(iterates[-1], dtype=np.float64).ravel()
    s = np.asarray(sens, dtype=np.float64).ravel()
    w = np.where((s > 0) & (xh > 0), s * xh, 0.0)
    D = np.stack(d)
    return (D * w) @ D.T


def combine(iterates, coef):
    d = log_ratios(iterates)
    xh = np.asarray(iterates[-1], dtype=np.float64)
    e = np.zeros(xh.size)
    for c, di in zip(coef, d):
        e += c * di
    y = np.where(xh.ravel() > 0, xh.ravel() * np.exp(e), xh.ravel())
    return y.reshape(xh.shape)


def ritz_coefficients(gram, pairs, target, horizon, rel_tol=1e-11, ritz_max=1.0 - 1e-5):
    """The Rayleigh-Ritz coefficient solve restated with numpy's symmetric
    eigensolver (LAPACK), the checker of the library's own solve
    (kv_ritz_coefficients, cyclic Jacobi): coefficients c over the h
    log-ratios with sum_i c_i d_i ~ p(J) J d_target, p(t) = sum_{k=1..N} t^k,
    and the Ritz values."""
    G = np.asarray(gram, dtype=np.float64)
    h = G.shape[0]
    bi = [int(i) for i, _ in pairs]
    bj = [int(j) for _, j in pairs]
    Gb = G[np.ix_(bi, bi)]
    H = G[np.ix_(bi, bj)]
    ev, U = np.linalg.eigh(Gb)
    coef = np.zeros(h)
    if not ev.size or ev[-1] <= 0.0 or horizon == 0:
        return coef, np.zeros(0)
    keep = ev > ev[-1] * rel_tol
    Q = U[:, keep] / np.sqrt(ev[keep])
    T = Q.T @ H @ Q
    T = 0.5 * (T + T.T)
    th, Y = np.linalg.eigh(T)
    th = np.clip(th, 0.0, ritz_max)
    p = th * (1.0 - th ** horizon) / (1.0 - th)
    c = Q.T @ Gb[:, bi.index(target)]
    cb = Q @ (Y @ (p * th * (Y.T @ c)))
    for k, i in enumerate(bi):
        coef[i] += cb[k]
    return coef, th
