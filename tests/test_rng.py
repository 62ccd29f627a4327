"""The seeded transmission-data sampler's restatement (oracle/rng.py, the
checker of kv_poisson_data) pinned by independent references:

* Philox4x64-10 blocks against numpy's Philox bit generator;
* exp, log and log k! against numpy / scipy to a few ulps;
* the inversion branch (means below 10) against scipy's Poisson quantile
  function on the same uniforms;
* both branches against the Poisson distribution: chi-square goodness of fit
  and the mean at several means;
* the clamp and log of the data transform by hand."""
import numpy as np
import pytest
from scipy import stats
from scipy.special import gammaln

from oracle import rng

M = 2**64 - 1


def test_philox_blocks_equal_numpy_philox():
    key = np.array([123456789, 0x6B76], dtype=np.uint64)
    for c1 in (0, 1, 7):
        # numpy increments the counter before each block: start one below
        start = np.array([M, c1 - 1 if c1 else M, 0 if c1 else M, 0 if c1 else M], dtype=np.uint64)
        raw = np.random.Philox(key=key, counter=start).random_raw(4 * 5)
        n = 5
        w = rng.philox4x64_10(np.arange(n, dtype=np.uint64), np.full(n, c1, np.uint64), np.zeros(n, np.uint64),
                              np.zeros(n, np.uint64), np.full(n, key[0]), np.full(n, key[1]))
        np.testing.assert_array_equal(np.stack(w, 1).ravel(), raw)
    # a counter with high bits set (carries inside the 128-bit products)
    start = np.array([2**63 + 41, 3, 0, 0], dtype=np.uint64)
    raw = np.random.Philox(key=key, counter=start).random_raw(4)
    w = rng.philox4x64_10([2**63 + 42], [3], [0], [0], [key[0]], [key[1]])
    np.testing.assert_array_equal(np.concatenate(w), raw)


def test_log_exp_logfactorial_against_the_math_library():
    r = np.random.Generator(np.random.PCG64(5))
    x = np.exp(r.uniform(-40, 40, 20000))
    ulp = np.abs(rng.kv_log(x) - np.log(x)) / np.spacing(np.abs(np.log(x)))
    assert ulp.max() <= 4
    y = r.uniform(-740, 705, 20000)
    assert (np.abs(rng.kv_exp(y) - np.exp(y)) / np.spacing(np.exp(y))).max() <= 2
    assert rng.kv_exp(np.array([-800.0]))[0] == 0.0 and rng.kv_exp(np.array([800.0]))[0] == np.inf
    k = np.arange(0, 50000, dtype=np.float64)
    ref = gammaln(k + 1)
    assert np.all(np.abs(rng.kv_log_factorial(k) - ref) <= 4e-15 * np.maximum(1.0, ref))


def test_inversion_branch_is_the_poisson_quantile():
    n = 20000
    lam = np.random.Generator(np.random.PCG64(6)).uniform(0.05, 9.99, n)
    idx = np.arange(n, dtype=np.uint64) + np.uint64(1000)
    k = rng.poisson(lam, idx, 99)
    u = rng.Streams(idx, 99).next(np.ones(n, bool))  # the branch's one uniform
    q = stats.poisson.ppf(u, lam)
    cdf = stats.poisson.cdf(q, lam)
    safe = np.abs(u - cdf) > 1e-12  # away from a CDF step
    assert safe.mean() > 0.999
    np.testing.assert_array_equal(k[safe], q[safe])


@pytest.mark.parametrize("lam", [0.7, 4.0, 9.99, 10.0, 23.5, 250.0, 4000.0])
def test_counts_follow_the_poisson_distribution(lam):
    n = 100000
    k = rng.poisson(np.full(n, lam), np.arange(n, dtype=np.uint64), int(lam * 1000) + 3)
    assert np.all(k == np.floor(k)) and k.min() >= 0
    # mean within 4.5 standard errors
    assert abs(k.mean() - lam) < 4.5 * np.sqrt(lam / n)
    # chi-square over bins holding >= 50 expected counts, tails pooled
    lo, hi = stats.poisson.ppf([1e-4, 1 - 1e-4], lam)
    edges = np.arange(lo, hi + 2)
    obs = np.array([(k < edges[1]).sum()] + [(k == e).sum() for e in edges[1:-2]] + [(k >= edges[-2]).sum()])
    cdf = stats.poisson.cdf(edges[1:-1] - 1, lam)
    probs = np.diff(np.concatenate([[0.0], cdf, [1.0]]))
    exp_counts = n * probs
    assert np.all(exp_counts > 0)
    keep = exp_counts >= 50
    o = np.concatenate([obs[keep], [obs[~keep].sum()]])
    e = np.concatenate([exp_counts[keep], [exp_counts[~keep].sum()]])
    if e[-1] == 0:
        o, e = o[:-1], e[:-1]
    chi2 = ((o - e) ** 2 / e).sum()
    p = stats.chi2.sf(chi2, len(o) - 1)
    assert p > 1e-4, (lam, chi2, len(o), p)


def test_data_transform_clamps_and_logs():
    # i0 = 100: ell = 50 gives a mean of ~2e-20 -> count 0 -> clamped to 1
    out = rng.poisson_data(np.array([50.0, 0.0], np.float32), 100.0, 1)
    assert out[0] == np.float32(np.log(100.0))
    # ell = 0: mean i0, counts above i0 clamp to i0 -> 0
    many = rng.poisson_data(np.zeros(4000, np.float32), 100.0, 2)
    assert many.min() == 0.0 and np.mean(many == 0.0) > 0.4
    assert np.all(many >= 0)
