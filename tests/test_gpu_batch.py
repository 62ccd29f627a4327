"""Batched reconstructions (kv_batch_*): many uneven MLEM problems (mode R,
and mode A for the learned path) in shared persistent launches, eagerly or
replayed from a hipGraph.

Bar: every reconstruction bit-identical to the same number of unbatched
kv_mlem_iteration calls on its own (which are bit-exact against the
reference's MLEM loop, checked here once more against the oracle for one
problem), for every tile shape, for per-problem iteration counts (0 leaves a
volume untouched) and across graph replays with other device work between
them.
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


def problems():
    """Uneven sizes, none a multiple of the tile shapes: partial forward
    tiles (64x4, 32x8, 16x16, 8x32) and partial bricks (64x4x4, 64x4x8)."""
    from kevernals_amd import geometry as G

    specs = [
        ((70, 37, 11), (0.6, 0.6, 2.0), (150, 90), (0.5, 0.5), 7, 15.0),
        ((48, 64, 16), (0.5, 0.5, 2.0), (100, 120), (0.5, 0.5), 9, 20.0),
        ((33, 29, 5), (0.9, 0.9, 3.0), (61, 50), (0.6, 0.6), 3, 10.0),
        ((96, 40, 9), (0.4, 0.6, 2.0), (190, 77), (0.45, 0.6), 11, 25.0),
        ((64, 64, 24), (0.5, 0.5, 1.5), (128, 128), (0.5, 0.5), 5, 12.0),
    ]
    return [G.arc_geometry(v, s, d, p, n, a) for v, s, d, p, n, a in specs]


def data_for(K, g, seed):
    from kevernals_amd.phantom import phantom

    ctx = K.Context(0, g)
    out = torch.empty(g.proj_shape, device="cuda")
    ctx.forward_project(torch.from_numpy(phantom(g, seed)).cuda(), out, K.MODE_R)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def unbatched(K, g, data, x0, iters):
    ctx = K.Context(0, g)
    x = torch.from_numpy(np.ascontiguousarray(x0)).cuda()
    d = torch.from_numpy(data).cuda()
    for _ in range(iters):
        ctx.mlem_iteration(x, d, 1.0, K.MODE_R)
    torch.cuda.synchronize()
    return x.cpu().numpy()


@pytest.fixture(scope="module")
def setup(K):
    gs = problems()
    datas = [data_for(K, g, 40 + i) for i, g in enumerate(gs)]
    x0s = [(0.5 + 0.5 * rng(50 + i).random(g.vol_shape)).astype(np.float32) for i, g in enumerate(gs)]
    return gs, datas, x0s


def fresh_batch(K, setup):
    gs, datas, x0s = setup
    b = K.Batch(gs)
    for i in range(len(gs)):
        b.set_volume(i, x0s[i])
        b.set_data(i, datas[i])
    return b


def test_batch_matches_oracle_mlem(K, setup, oracle_mod):
    gs, datas, x0s = setup
    b = K.Batch(gs[2:3])
    b.set_volume(0, x0s[2])
    b.set_data(0, datas[2])
    b.mlem([3], use_graph=False)
    ref = oracle_mod.Oracle(gs[2]).mlem(datas[2], x0s[2], 3, 1.0)
    np.testing.assert_array_equal(b.get_volume(0), ref)


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_uneven_iterations_bit_exact(K, setup, use_graph):
    gs, datas, x0s = setup
    iters = [3, 1, 0, 2, 4]
    b = fresh_batch(K, setup)
    ms = b.mlem(iters, use_graph=use_graph)
    print(f"[batch] {len(gs)} problems iters {iters} graph={use_graph}: {ms:.3f} ms", flush=True)
    for i, g in enumerate(gs):
        got = b.get_volume(i)
        if iters[i] == 0:
            np.testing.assert_array_equal(got, x0s[i])
        else:
            np.testing.assert_array_equal(got, unbatched(K, g, datas[i], x0s[i], iters[i]), err_msg=f"problem {i}")


def test_batch_tile_shapes_do_not_change_results(K, setup):
    gs, datas, x0s = setup
    refs = [unbatched(K, g, datas[i], x0s[i], 2) for i, g in enumerate(gs)]
    for layout in (1, 2, 4, 8):
        for slices in (4, 8):
            b = fresh_batch(K, setup)
            for i in range(len(gs)):
                b.set_tiles(i, layout, slices)
                assert b.get_tiles(i) == (layout, slices)
            b.mlem(2, use_graph=False)
            for i in range(len(gs)):
                np.testing.assert_array_equal(b.get_volume(i), refs[i], err_msg=f"{layout}/{slices} problem {i}")
            b.close()
    # mixed shapes in one batch
    b = fresh_batch(K, setup)
    for i, (l, z) in enumerate([(4, 4), (1, 8), (2, 4), (4, 8), (2, 8)]):
        b.set_tiles(i, l, z)
    b.mlem(2, use_graph=True)
    for i in range(len(gs)):
        np.testing.assert_array_equal(b.get_volume(i), refs[i], err_msg=f"mixed problem {i}")


def test_batch_graph_replays_with_eager_work_between(K, setup):
    gs, datas, x0s = setup
    b = fresh_batch(K, setup)
    iters = [2, 1, 2, 1, 2]
    total = [0] * len(gs)
    for replay in range(3):
        b.mlem(iters, use_graph=True)
        total = [t + k for t, k in zip(total, iters)]
        # unrelated device work and synchronisation between replays
        junk = torch.randn(1 << 20, device="cuda")
        assert float((junk * 0).sum()) == 0.0
        torch.cuda.synchronize()
        for i, g in enumerate(gs):
            np.testing.assert_array_equal(b.get_volume(i), unbatched(K, g, datas[i], x0s[i], total[i]),
                                          err_msg=f"replay {replay} problem {i}")


def test_batch_autotune_keeps_volumes_and_caches(K, setup, tmp_path):
    gs, datas, x0s = setup
    b = fresh_batch(K, setup)
    cache = str(tmp_path / "tiles.txt")
    tiles = b.autotune(cache)
    assert all(l in (1, 2, 4, 8) and z in (4, 8) for l, z in tiles)
    for i in range(len(gs)):
        np.testing.assert_array_equal(b.get_volume(i), x0s[i])
    lines = open(cache).read().split()
    assert len(lines) >= 5
    b2 = fresh_batch(K, setup)
    assert b2.autotune(cache) == tiles  # read back from the cache
    b2.mlem(1)
    for i, g in enumerate(gs):
        np.testing.assert_array_equal(b2.get_volume(i), unbatched(K, g, datas[i], x0s[i], 1))


def test_batch_rejects_bad_arguments(K, setup):
    gs = setup[0]
    b = K.Batch(gs[:2])
    with pytest.raises(K.KvError):
        b.set_tiles(0, 3, 8)
    with pytest.raises(K.KvError):
        b.set_tiles(5, 1, 8)
    with pytest.raises(K.KvError):
        b.mlem([-1, 2])


def test_batch_mode_a_bit_exact(K, setup):
    """Mode-A batches (x <- x/s * A^T(data/Ax)) match unbatched MLEM-A."""
    gs, datas, x0s = setup
    b = K.Batch(gs, mode=K.MODE_A)
    for i in range(len(gs)):
        b.set_volume(i, x0s[i])
        b.set_data(i, datas[i])
    iters = [2, 0, 3, 1, 2]
    b.mlem(iters, use_graph=True)
    for i, g in enumerate(gs):
        ctx = K.Context(0, g)
        x = torch.from_numpy(x0s[i].copy()).cuda()
        d = torch.from_numpy(datas[i]).cuda()
        for _ in range(iters[i]):
            ctx.mlem_iteration(x, d, 1.0, K.MODE_A)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(b.get_volume(i), x.cpu().numpy(), err_msg=f"problem {i}")


def test_batched_gate_falls_back_to_plain(K, setup):
    """A network whose candidates lower the likelihood leaves every
    reconstruction of a batched gated run identical to plain batched MLEM-A;
    the per-reconstruction context and buffers address the batch's data."""
    from kevernals_amd.learned import batched_gated_mlem, init_params

    gs, datas, x0s = setup
    bad = init_params(3)
    bad[-1] = 5.0
    runs = []
    for params in (None, bad):
        b = K.Batch(gs, mode=K.MODE_A)
        for i in range(len(gs)):
            b.set_volume(i, x0s[i])
            b.set_data(i, np.maximum(datas[i], 0.0))
        vol_addr, data_addr = b.buffers(1)
        assert vol_addr and data_addr
        run = batched_gated_mlem(b, 6, params=params, every=2, first=2, patience=None)
        runs.append((run, [b.get_volume(i) for i in range(len(gs))]))
    (plain, vp), (gated, vg) = runs
    assert sum(gated.accepted) == 0 and sum(gated.rejected) == 3 * len(gs)
    for a, c in zip(vp, vg):
        np.testing.assert_array_equal(a, c)
