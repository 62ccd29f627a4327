"""Trajectory extrapolation on the CPU: the oracle of the two volume passes
pinned by closed forms, and the host coefficient solve
(kevernals_amd.trajectory.extrapolation_coefficients) pinned by linear
recurrences whose extrapolation is exact."""
import numpy as np
import pytest

from conftest import rng
from kevernals_amd.trajectory import extrapolation_coefficients
from oracle import traj


def recurrence(nvox, lams, h, seed):
    """log x_k(v) = L(v) + a(v) * lam_{g(v)}^k: a linear iteration of the
    log-ratios with a diagonal Jacobian (one eigenvalue per voxel group)."""
    r = rng(seed)
    L = np.log(0.02 + 0.05 * r.random(nvox))
    a = 0.5 * r.standard_normal(nvox)
    g = r.integers(0, len(lams), nvox)
    lam = np.asarray(lams)[g]

    def x(k):
        return np.exp(L + a * lam ** k)

    sens = 0.5 + r.random(nvox)
    return x, sens


def test_gram_single_mode_closed_form():
    """d_i = lam^i d_0 (every voxel the same eigenvalue): G_ij =
    lam^(i+j) * sum_v w d_0^2 with w = sens * x_h; a weight taken from the
    oldest iterate, or ratios taken the other way round, fail this."""
    x, sens = recurrence(500, [0.8], 6, 1)
    its = [x(k) for k in range(7)]
    G = traj.gram(its, sens)
    d0 = np.log(its[1] / its[0])
    base = np.sum(sens * its[-1] * d0 * d0)
    i, j = np.meshgrid(np.arange(6), np.arange(6), indexing="ij")
    np.testing.assert_allclose(G, base * 0.8 ** (i + j), rtol=1e-10)


def test_gram_and_combine_by_hand():
    its = [np.array([1.0, 2.0, -1.0]), np.array([2.0, 2.0, 1.0]), np.array([4.0, 1.0, 1.0])]
    sens = np.array([1.0, 3.0, 5.0])
    # d_0 = [log 2, 0, 0 (x_0 < 0)], d_1 = [log 2, log 1/2, 0]; w = [4, 3, 5]
    l2 = np.log(2.0)
    want = np.array([[4 * l2 * l2, 4 * l2 * l2], [4 * l2 * l2, 4 * l2 * l2 + 3 * l2 * l2]])
    np.testing.assert_allclose(traj.gram(its, sens), want, rtol=1e-14)
    y = traj.combine(its, [1.0, -2.0])
    np.testing.assert_allclose(y, [4.0 * np.exp(l2 - 2 * l2), 1.0 * np.exp(2 * l2), 1.0], rtol=1e-14)
    # non-positive newest iterate: kept as it is
    its[-1] = np.array([0.0, 1.0, -3.0])
    assert traj.combine(its, [1.0, 1.0])[0] == 0.0 and traj.combine(its, [1.0, 1.0])[2] == -3.0


@pytest.mark.parametrize("lams,h", [([0.5, 0.9, 0.97], 6), ([0.3, 0.6, 0.8, 0.95, 0.99], 12), ([0.7], 3)])
@pytest.mark.parametrize("horizon", [1, 10, 80])
def test_extrapolation_exact_on_linear_recurrences(lams, h, horizon):
    """With more history than distinct eigenvalues the Ritz values are the
    eigenvalues and the jump lands on the iterate `horizon` steps ahead."""
    x, sens = recurrence(2000, lams, h, 7)
    k0 = 3
    its = [x(k) for k in range(k0, k0 + h + 1)]
    coef, ritz = extrapolation_coefficients(traj.gram(its, sens), horizon, rel_tol=1e-12)
    y = traj.combine(its, coef)
    want = x(k0 + h + horizon)
    np.testing.assert_allclose(np.log(y), np.log(want), rtol=0, atol=1e-7)
    for lam in lams:
        assert np.min(np.abs(ritz - lam)) < 1e-6


def test_extrapolation_horizon_zero_and_degenerate_history():
    x, sens = recurrence(300, [0.9], 4, 3)
    its = [x(k) for k in range(5)]
    coef, _ = extrapolation_coefficients(traj.gram(its, sens), 0)
    assert np.all(coef == 0.0)
    flat = [np.ones(300)] * 5  # no motion: zero Gram matrix
    coef, ritz = extrapolation_coefficients(traj.gram(flat, sens), 20)
    assert np.all(coef == 0.0) and ritz.size == 0
    with pytest.raises(ValueError):
        extrapolation_coefficients(np.zeros((1, 1)), 5)


def test_truncation_drops_rounding_noise():
    """FP32 iterates of a one-mode recurrence: the extra history directions
    are rounding noise and are dropped; the jump stays accurate."""
    x, sens = recurrence(4000, [0.9], 8, 11)
    its = [x(k).astype(np.float32) for k in range(2, 11)]
    coef, ritz = extrapolation_coefficients(traj.gram(its, sens), 30, rel_tol=1e-9)
    y = traj.combine(its, coef)
    np.testing.assert_allclose(np.log(y), np.log(x(10 + 30)), rtol=0, atol=2e-5)
    assert ritz.size < 7


def evolve(L, lam, z0, j):
    """Iterate j of the diagonal log-linear recurrence started from z0."""
    return np.exp(L + lam ** j * (np.log(z0) - L))


@pytest.mark.parametrize("lams", [[0.5, 0.9, 0.97], [0.2, 0.6, 0.85, 0.95, 0.99]])
def test_two_stretch_jump_exact_on_linear_recurrence(lams):
    """Stretch A of a log-linear recurrence, an arbitrary jump, stretch B of
    the same recurrence from the jump's result: the final jump from the end
    of stretch A lands exactly on the recurrence's later iterate (the
    combined basis spans every mode).  A wrong pair (the jump treated as a
    J-step) or a missing rebase to stretch A's end fails this."""
    from kevernals_amd.trajectory import two_stretch_coefficients

    r = rng(5)
    n = 3000
    L = np.log(0.02 + 0.05 * r.random(n))
    lam = np.asarray(lams)[r.integers(0, len(lams), n)]
    x0 = np.exp(L + 0.5 * r.standard_normal(n))
    sens = 0.5 + r.random(n)
    A = [evolve(L, lam, x0, j) for j in range(2, 11)]  # iterates 2..10
    z = np.exp(L + 0.3 * r.standard_normal(n))  # the jump's result: any state
    B = [evolve(L, lam, z, j) for j in range(0, 9)]
    its = A + B
    for horizon in (5, 40, 90):
        coef, ritz = two_stretch_coefficients(traj.gram(its, sens), len(A), horizon)
        y = traj.combine(its, coef)
        np.testing.assert_allclose(np.log(y), np.log(evolve(L, lam, x0, 10 + horizon)), rtol=0, atol=2e-6)
    with pytest.raises(ValueError):  # two iterates in stretch B: no J-pair there
        two_stretch_coefficients(traj.gram(its[:11], sens), 9, 10)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_native_solve_matches_lapack(seed):
    """The library's Jacobi solve (kv_ritz_coefficients, behind
    trajectory.ritz_coefficients) against numpy's LAPACK eigensolver in the
    oracle, on Gram matrices of noisy recurrences with one and two stretches:
    the Ritz values and the coefficients agree to rounding."""
    from kevernals_amd.trajectory import ritz_coefficients

    r = rng(100 + seed)
    x, sens = recurrence(3000, [0.4, 0.7, 0.9, 0.97, 0.995], 17, seed)
    its = [x(k) * (1.0 + 1e-7 * r.standard_normal(3000)) for k in range(18)]
    G = traj.gram(its, sens)
    for pairs, target, horizon in (([(i, i + 1) for i in range(16)], 15, 60),
                                   ([(i, i + 1) for i in range(7)] + [(i, i + 1) for i in range(9, 16)], 6, 90)):
        got, ritz = ritz_coefficients(G, pairs, target, horizon)
        want, ritz_w = traj.ritz_coefficients(G, pairs, target, horizon)
        assert ritz.shape == ritz_w.shape
        np.testing.assert_allclose(ritz, ritz_w, rtol=0, atol=1e-9)
        scale = np.abs(want).max()
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-7 * scale)
