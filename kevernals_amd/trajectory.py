"""Trajectory extrapolation of mode-A MLEM: a jump from iteration k to an
estimate of iteration k + N.

Late MLEM is close to a linear iteration of the log-ratios,
d_{i+1} = J d_i with d_i = log(x_{i+1}/x_i), and its Jacobian at the fixed
point, J = I - diag(x/sens) A^T diag(data/(Ax)^2) A, is self-adjoint in the
inner product weighted by sens*x (log domain) with eigenvalues in [0, 1).
The last h+1 iterates therefore carry a Krylov space of J for free: with
B = [d_0 .. d_{h-2}] and JB = [d_1 .. d_{h-1}], the Gram matrices
B^T W B and B^T W JB (both blocks of the one h x h matrix that kv_traj_gram
returns) give the Rayleigh-Ritz projection T of J, and the remaining path
    log x_{h+N} - log x_h = sum_{i=1..N} J^i d_{h-1} = p(J) J d_{h-2},
    p(t) = sum_{i=1..N} t^i,
is approximated by B Q p(T) T (coordinates of d_{h-2}) in the W-orthonormal
basis B Q.  The volume passes (Gram matrix, combination) are the HIP kernels
behind kv_traj_gram / kv_traj_combine; the (h-1) x (h-1) eigenproblems here
are host arithmetic in FP64.

This has no counterpart in the reference; it accelerates recons::MLEM in its
mode-A form (Reconstructors.h:182-242).
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from . import TRAJ_MAX, Context


def extrapolation_coefficients(gram: np.ndarray, horizon: int, rel_tol: float = 1e-11,
                               ritz_max: float = 1.0 - 1e-5):
    """Coefficients c (length h) with log x_{h+horizon} ~ log x_h + sum_i c_i d_i,
    and the Ritz values used.  Directions of the history whose Gram
    eigenvalue is below rel_tol times the largest are dropped (they are
    rounding noise of the FP32 iterates); Ritz values are clamped to
    [0, ritz_max] (EM's Jacobian has its spectrum in [0, 1))."""
    G = np.asarray(gram, dtype=np.float64)
    h = G.shape[0]
    if G.shape != (h, h) or h < 2:
        raise ValueError(f"need an h x h Gram matrix with h >= 2, got {G.shape}")
    if horizon < 0:
        raise ValueError("horizon must be >= 0")
    n = h - 1
    Gb, H = G[:n, :n], G[:n, 1:]
    ev, U = np.linalg.eigh(Gb)
    if not ev.size or ev[-1] <= 0.0 or horizon == 0:
        return np.zeros(h), np.zeros(0)
    keep = ev > ev[-1] * rel_tol
    Q = U[:, keep] / np.sqrt(ev[keep])
    T = Q.T @ H @ Q
    T = 0.5 * (T + T.T)
    th, Y = np.linalg.eigh(T)
    th = np.clip(th, 0.0, ritz_max)
    p = th * (1.0 - th ** horizon) / (1.0 - th)
    c = Q.T @ Gb[:, n - 1]
    coef = Q @ (Y @ (p * th * (Y.T @ c)))
    return np.append(coef, 0.0), th


@dataclass
class Extrapolation:
    coef: np.ndarray
    ritz: np.ndarray
    horizon: int
    history: int


class Trajectory:
    """The last `history`+1 iterates of one reconstruction, as device copies
    in a ring (history <= TRAJ_MAX)."""

    def __init__(self, ctx: Context, shape, history: int):
        import torch

        if not 2 <= history <= TRAJ_MAX:
            raise ValueError(f"history must be in 2..{TRAJ_MAX}")
        self.ctx = ctx
        self.history = history
        self.slots = [torch.empty(shape, device="cuda") for _ in range(history + 1)]
        self.count = 0

    def push(self, x):
        """Record an iterate (copied on the context's stream)."""
        dst = self.slots[self.count % len(self.slots)]
        self.ctx.copy(dst, x, x.numel() * 4)
        self.count += 1

    def iterates(self) -> List:
        n = min(self.count, len(self.slots))
        start = self.count - n
        return [self.slots[i % len(self.slots)] for i in range(start, self.count)]

    def reset(self):
        self.count = 0

    def extrapolate(self, horizon: int, out, rel_tol: float = 1e-11) -> Optional[Extrapolation]:
        """out = estimate of the iterate `horizon` iterations after the
        newest one; None (out untouched) with fewer than 3 iterates."""
        xs = self.iterates()
        if len(xs) < 3:
            return None
        G = self.ctx.traj_gram(xs)
        coef, ritz = extrapolation_coefficients(G, horizon, rel_tol)
        self.ctx.traj_combine(xs, coef, out)
        return Extrapolation(coef, ritz, horizon, len(xs) - 1)
