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
behind kv_traj_gram / kv_traj_combine; the (h-1) x (h-1) eigenproblems are
the library's host arithmetic in FP64 (kv_ritz_coefficients).

This has no counterpart in the reference; it accelerates recons::MLEM in its
mode-A form (Reconstructors.h:182-242).
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from . import TRAJ_MAX, Context, ritz_coefficients_native


def ritz_coefficients(gram: np.ndarray, pairs, target: int, horizon: int, rel_tol: float = 1e-11,
                      ritz_max: float = 1.0 - 1e-5):
    """Coefficients c over all h log-ratios with
        sum_i c_i d_i ~ p(J) J d_target,   p(t) = sum_{i=1..horizon} t^i,
    from a Gram matrix of h log-ratios and the list of pairs (i, j) with
    d_j = J d_i (consecutive log-ratios of one stretch of MLEM iterations;
    the pairs may come from several stretches).  d_target must be one of the
    pairs' first members.  Returns (c, Ritz values).

    Rayleigh-Ritz on the span of the pairs' first members B: Gram matrices
    B^T W B and B^T W JB, a W-orthonormal basis B Q from the eigenvectors of
    B^T W B whose eigenvalues exceed rel_tol times the largest (the rest is
    rounding noise of FP32 iterates), T = Q^T (B^T W JB) Q symmetrised, Ritz
    values clamped to [0, ritz_max] (EM's Jacobian has its spectrum in
    [0, 1)).  The eigenproblems are solved by the library's host code
    (kv_ritz_coefficients, cyclic Jacobi in FP64)."""
    G = np.asarray(gram, dtype=np.float64)
    h = G.shape[0]
    if G.shape != (h, h) or h < 2:
        raise ValueError(f"need an h x h Gram matrix with h >= 2, got {G.shape}")
    if horizon < 0:
        raise ValueError("horizon must be >= 0")
    pairs = [(int(i), int(j)) for i, j in pairs]
    if not pairs or any(not (0 <= i < h and 0 <= j < h) for i, j in pairs):
        raise ValueError("pairs must index the Gram matrix")
    if target not in [i for i, _ in pairs]:
        raise ValueError("the target must be the first member of a pair")
    if h > TRAJ_MAX:
        raise ValueError(f"at most {TRAJ_MAX} log-ratios")
    # the host solve of the library (kv_ritz_coefficients), which the C++
    # accelerated reconstruction uses too
    return ritz_coefficients_native(G, pairs, target, horizon, rel_tol, ritz_max)


def extrapolation_coefficients(gram: np.ndarray, horizon: int, rel_tol: float = 1e-11,
                               ritz_max: float = 1.0 - 1e-5):
    """One stretch of h+1 consecutive iterates: coefficients c (length h) with
    log x_{h+horizon} ~ log x_h + sum_i c_i d_i (d_{h-1} = J d_{h-2}, so the
    remaining path is p(J) J d_{h-2}), and the Ritz values used."""
    G = np.asarray(gram, dtype=np.float64)
    h = G.shape[0]
    if G.shape != (h, h) or h < 2:
        raise ValueError(f"need an h x h Gram matrix with h >= 2, got {G.shape}")
    return ritz_coefficients(G, [(i, i + 1) for i in range(h - 1)], h - 2, horizon, rel_tol, ritz_max)


def two_stretch_coefficients(gram: np.ndarray, na: int, horizon: int, rel_tol: float = 1e-11):
    """Iterates x_0..x_{na-1} of stretch A, then a jump, then stretch B
    (the jump's result and the MLEM iterates after it), as one list of h+1
    iterates: coefficients c over its h log-ratios such that
    x_newest * exp(sum_i c_i d_i) estimates the iterate `horizon` iterations
    after x_{na-1}, and the Ritz values used.  Log-ratios 0..na-2 belong to
    stretch A, na-1 is the jump (no J-pair), na..h-1 to stretch B."""
    G = np.asarray(gram, dtype=np.float64)
    h = G.shape[0]
    if na < 4 or h - na < 2:
        raise ValueError("need at least 4 iterates in stretch A and 3 in stretch B")
    pairs = [(i, i + 1) for i in range(na - 2)] + [(i, i + 1) for i in range(na, h - 1)]
    coef, ritz = ritz_coefficients(G, pairs, na - 3, horizon, rel_tol)
    # combination base: log x_{na-1} = log x_newest - (jump + stretch B log-ratios)
    coef[na - 1:] -= 1.0
    return coef, ritz


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


@dataclass
class Schedule:
    """Two-stretch jump schedule.  MLEM-A runs `first` iterations, recording
    the iterates from `start` on (stretch A); a probe then jumps to an
    estimate of iteration `mid` and runs `second` MLEM iterations from there
    (stretch B: their log-ratios are filtered by about mid powers of J, so
    they span the slow modes that stretch A cannot resolve); the final jump
    starts again from iteration `first`, with the Ritz basis of both
    stretches, and lands on an estimate of iteration `horizon`.  Cost: first
    + second MLEM iterations, two Gram passes and two combinations.  The
    probe is side work: the estimate stays at iteration `first` until the
    final jump replaces it."""
    start: int = 2
    first: int = 10
    mid: int = 40
    second: int = 8
    horizon: int = 100

    def __post_init__(self):
        na = self.first - self.start + 1
        if self.start < 0 or na < 4:
            raise ValueError("stretch A needs at least 4 iterates (first - start >= 3)")
        if self.second < 2:
            raise ValueError("stretch B needs at least 2 iterations")
        if na + self.second > TRAJ_MAX + 1:
            raise ValueError(f"the two stretches hold at most {TRAJ_MAX + 1} iterates")
        if not self.first < self.mid < self.horizon:
            raise ValueError("need first < mid < horizon")

    @property
    def cost(self) -> int:
        return self.first + self.second


class Jumper:
    """The recorded stretches of one reconstruction and its two jumps
    (device copies; one context per reconstruction)."""

    def __init__(self, ctx: Context, shape, schedule: Schedule):
        import torch

        self.ctx = ctx
        self.s = schedule
        self.na = schedule.first - schedule.start + 1
        self.slots = [torch.empty(shape, device="cuda") for _ in range(self.na + schedule.second + 1)]
        self.count = 0

    def record(self, x):
        """Copy the next iterate (a device tensor or address of a volume of
        this geometry) into its slot, on the context's stream."""
        if self.count >= len(self.slots):
            raise RuntimeError("both stretches are recorded already")
        dst = self.slots[self.count]
        self.ctx.copy(dst, x, dst.numel() * 4)
        self.count += 1

    @property
    def estimate(self):
        """Stretch A's last iterate (iteration `first`): the estimate while
        stretch B probes ahead."""
        return self.slots[self.na - 1]

    def first_jump(self, out) -> Extrapolation:
        """out = estimate of iteration `mid` from stretch A (complete)."""
        if self.count != self.na:
            raise RuntimeError("stretch A is not complete")
        xs = self.slots[:self.na]
        coef, ritz = extrapolation_coefficients(self.ctx.traj_gram(xs), self.s.mid - self.s.first)
        self.ctx.traj_combine(xs, coef, out)
        return Extrapolation(coef, ritz, self.s.mid - self.s.first, self.na - 1)

    def final_jump(self, out) -> Extrapolation:
        """out = estimate of iteration `horizon`, from iteration `first`, with
        the Ritz basis of both stretches (complete: the first jump's result
        and the `second` iterations after it recorded)."""
        if self.count != len(self.slots):
            raise RuntimeError("stretch B is not complete")
        horizon = self.s.horizon - self.s.first
        coef, ritz = two_stretch_coefficients(self.ctx.traj_gram(self.slots), self.na, horizon)
        self.ctx.traj_combine(self.slots, coef, out)
        return Extrapolation(coef, ritz, horizon, len(self.slots) - 1)
