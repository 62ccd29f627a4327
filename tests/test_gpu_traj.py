"""Trajectory extrapolation on the GPU (kv_traj_gram, kv_traj_combine)
against the FP64 oracle on real mode-A MLEM histories.

Bar: the Gram matrix within 1e-6 of sqrt(G_ii G_jj) per entry (FP32 log-ratios
staged as sqrt(w)*d in FP32, products and sums in FP64) and bitwise
reproducible; the combination within 2e-6 relative; a jump along a real MLEM
run lands far closer to the later iterate than the iterate it starts from.
"""
import numpy as np
import pytest

from conftest import rng

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def K():
    import kevernals_amd

    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return kevernals_amd


def mlem_run(K, g, iters, seed):
    from kevernals_amd.learned import transmission_data
    from kevernals_amd.phantom import phantom

    ctx = K.Context(0, g)
    mu = torch.from_numpy(phantom(g, seed)).cuda()
    data = transmission_data(ctx, mu, 1e4, seed)
    x = torch.ones(g.vol_shape, device="cuda")
    out = [x.clone()]
    for _ in range(iters):
        ctx.mlem_iteration(x, data, 1.0, K.MODE_A)
        out.append(x.clone())
    torch.cuda.synchronize()
    return ctx, data, out


@pytest.fixture(scope="module")
def run(K):
    from kevernals_amd import geometry as G

    g = G.arc_geometry((70, 37, 11), (0.6, 0.6, 2.0), (150, 90), (0.5, 0.5), 7, 15.0)
    return mlem_run(K, g, 40, 31)


def sensitivity(K, ctx):
    g = ctx.geometry
    s = torch.empty(g.vol_shape, device="cuda")
    ctx.back_project(torch.ones(g.proj_shape, device="cuda"), s, K.MODE_A)
    torch.cuda.synchronize()
    return s.cpu().numpy()


@pytest.mark.parametrize("first,h", [(2, 1), (5, 6), (10, 13), (3, 32)])
def test_gram_matches_oracle(K, run, first, h):
    from oracle import traj

    ctx, _, its = run
    xs = its[first:first + h + 1]
    got = ctx.traj_gram(xs)
    want = traj.gram([x.cpu().numpy() for x in xs], sensitivity(K, ctx))
    scale = np.sqrt(np.outer(np.diag(want), np.diag(want)))
    err = np.abs(got - want) / scale
    print(f"[traj] gram h={h}: max err {err.max():.2e} of sqrt(Gii Gjj)", flush=True)
    assert err.max() <= 1e-6
    np.testing.assert_array_equal(got, got.T)
    np.testing.assert_array_equal(ctx.traj_gram(xs), got)  # deterministic


def test_gram_ignores_non_positive_voxels(K, run):
    from oracle import traj

    ctx, _, its = run
    xs = [x.clone() for x in its[8:13]]
    r = rng(4)
    for x in xs[1:3]:
        flat = x.view(-1)
        idx = torch.from_numpy(r.integers(0, flat.numel(), 300)).cuda()
        flat[idx] = 0.0
    xs[2].view(-1)[:50] = -1.0
    got = ctx.traj_gram(xs)
    want = traj.gram([x.cpu().numpy() for x in xs], sensitivity(K, ctx))
    assert np.all(np.isfinite(got))
    scale = np.sqrt(np.outer(np.diag(want), np.diag(want)))
    assert (np.abs(got - want) / scale).max() <= 1e-6


@pytest.mark.parametrize("h", [1, 7, 32])
def test_combine_matches_oracle(K, run, h):
    from oracle import traj

    ctx, _, its = run
    xs = its[40 - h - 1:40] if h < 32 else its[4:37]
    coef = rng(h).standard_normal(h) * 3.0
    out = torch.empty_like(xs[0])
    ctx.traj_combine(xs, coef, out)
    torch.cuda.synchronize()
    want = traj.combine([x.cpu().numpy() for x in xs], coef)
    np.testing.assert_allclose(out.cpu().numpy(), want, rtol=2e-6, atol=0)
    # in place over the newest iterate
    last = xs[-1].clone()
    ctx.traj_combine(xs[:-1] + [last], coef, last)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(last.cpu().numpy(), out.cpu().numpy())


def extreme_pair(shape, seed):
    """Iterate pairs (x0, x1) covering every path of the passes' log-ratio:
    ratios near 1 (the series; 0), around b/a = 3/5 and 5/3 (its edges; 1,
    2), far from 1 (the hardware log of the quotient; 3), a subnormal or
    barely normal x0 under a moderate x1 (its reciprocal leaves the normal
    range; 4), magnitudes whose sum leaves it at the top (5), zero or
    negative entries (ratio 0; 6), and both subnormal (7)."""
    r = rng(seed)
    n = int(np.prod(shape))
    cat = r.integers(0, 8, n)
    x0 = np.exp(r.uniform(-20.0, 20.0, n))
    x1 = x0 * np.where(cat == 0, 1.0 + r.uniform(-1e-3, 1e-3, n), 1.0)
    x1 = np.where(cat == 1, x0 * r.uniform(0.55, 0.65, n), x1)
    x1 = np.where(cat == 2, x0 * r.uniform(1.6, 1.75, n), x1)
    x1 = np.where(cat == 3, x0 * np.exp(r.uniform(-8.0, 8.0, n)), x1)
    x0 = np.where(cat == 4, 10.0 ** r.uniform(-40.0, -36.0, n), x0)
    x1 = np.where(cat == 4, 10.0 ** r.uniform(-6.0, -4.0, n), x1)  # d <= 83: exp(d) stays finite
    x0 = np.where(cat == 5, r.uniform(1.2e38, 1.7e38, n), x0)
    x1 = np.where(cat == 5, x0 * r.uniform(0.8, 1.25, n), x1)
    x0 = np.where(cat == 7, r.uniform(1e-44, 1e-38, n), x0)
    x1 = np.where(cat == 7, x0 * r.uniform(0.5, 2.0, n), x1)
    x0 = x0.astype(np.float32)
    x1 = x1.astype(np.float32)
    zero = cat == 6
    x0[zero & (r.random(n) < 0.5)] = 0.0
    x1[zero & (r.random(n) < 0.5)] = -1.0
    return x0.reshape(shape), x1.reshape(shape), cat.reshape(shape)


def test_log_ratio_paths_against_oracle(K, run):
    """Every path of the log-ratio (series, quotient log, log1p for extreme
    magnitudes, zero) through both passes, against the FP64 oracle: the
    one-ratio Gram entry to 1e-6, and the combination with c = 1 (y =
    x1 * exp(d), d = log(x1 / x0)) per voxel to 2e-6 + 5e-7 |d| relative
    (a few units in the last place of d, amplified by |d|) where y is a
    normal float; everywhere finite, and x1 itself where a ratio is not
    positive."""
    from oracle import traj

    ctx, _, its = run
    shape = its[0].shape
    x0, x1, cat = extreme_pair(shape, 91)
    assert all(np.any(cat == c) for c in range(8))
    xs = [torch.from_numpy(x0).cuda(), torch.from_numpy(x1).cuda()]
    got = ctx.traj_gram(xs)
    want = traj.gram([x0, x1], sensitivity(K, ctx))
    print(f"[traj] extreme pairs: gram {got[0, 0]:.9e} vs {want[0, 0]:.9e}", flush=True)
    assert abs(got[0, 0] - want[0, 0]) <= 1e-6 * want[0, 0]
    out = torch.empty_like(xs[0])
    ctx.traj_combine(xs, [1.0], out)
    torch.cuda.synchronize()
    y = out.cpu().numpy().astype(np.float64)
    assert np.all(np.isfinite(y))
    ref = traj.combine([x0, x1], [1.0])
    d = traj.log_ratios([x0, x1])[0].reshape(shape)
    f32 = np.finfo(np.float32)
    ok = (np.abs(ref) >= f32.tiny) & (np.abs(ref) < f32.max)
    rel = np.abs(y[ok] - ref[ok]) / np.abs(ref[ok])
    bound = 2e-6 + 5e-7 * np.abs(d[ok])
    worst = {c: float((rel / bound)[cat[ok] == c].max()) for c in range(8) if np.any(cat[ok] == c)}
    print(f"[traj] extreme pairs: combine worst error / bound by path {worst}", flush=True)
    assert np.all(rel <= bound)
    for c in (0, 1, 2, 3, 4, 5):
        assert np.any(ok & (cat == c)), f"path {c} unchecked"
    keep = (x0 <= 0) | (x1 <= 0)
    np.testing.assert_array_equal(out.cpu().numpy()[keep], x1[keep])


def test_jump_along_mlem_lands_near_later_iterate(K):
    """History of MLEM-A iterations 6..20 on a 96x96x24 / 9-view problem,
    jump 40 iterations ahead: the estimate is several times closer to
    iteration 60 than iteration 20 is (relative RMSE)."""
    from kevernals_amd import geometry as G
    from kevernals_amd.trajectory import Trajectory

    g = G.arc_geometry((96, 96, 24), (0.5, 0.5, 2.0), (192, 160), (0.5, 0.5), 9, 15.0)
    ctx, _, its = mlem_run(K, g, 60, 77)
    tr = Trajectory(ctx, g.vol_shape, 14)
    for x in its[:21]:
        tr.push(x)
    assert len(tr.iterates()) == 15
    y = torch.empty_like(its[0])
    info = tr.extrapolate(40, y)
    torch.cuda.synchronize()
    ref = its[60].double()

    def rel(v):
        return float((v.double() - ref).norm() / ref.norm())

    print(f"[traj] RMSE vs iteration 60: iteration 20 {rel(its[20]):.4f}, jump {rel(y):.4f}; "
          f"ritz {np.round(info.ritz, 3).tolist()}", flush=True)
    assert rel(y) < rel(its[20]) / 3


def test_rejects_bad_arguments(K, run):
    ctx, _, its = run
    with pytest.raises(K.KvError):
        ctx.traj_gram(its[:1])  # h = 0
    with pytest.raises(K.KvError):
        ctx.traj_gram(its[:35])  # h = 34 > 32
    with pytest.raises(K.KvError):
        ctx.traj_combine(its[:3], [1.0, float("nan")], its[0].clone())
    with pytest.raises(ValueError):
        ctx.traj_combine(its[:3], [1.0], its[0].clone())
