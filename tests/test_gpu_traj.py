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
