"""Each workload configuration at its own sizes on the GPU.

* 128x128x32 / 15 views: forward projection, backprojection and 10 MLEM
  iterations of recons::MLEM (Reconstructors.h:136-253) against the oracle
  on every voxel, bit for bit.
* 1024x1024x64 / 25 views: SART with relaxation 0.5 for 10 iterations
  (recons::ART, Reconstructors.h:19-134) -- the 10th iteration pinned by the
  oracle on sampled rays and voxels -- and shift-and-add at the volume's 64
  planes (Reconstructors.h:263-379) against the oracle on four planes.
* 32 uneven reconstructions between 256x256x32 and 1024x768x80, 15 MLEM
  iterations in the persistent batched launches (hipGraph): sampled
  reconstructions bit-identical to their unbatched runs, every
  reconstruction keeping a consistent uniform volume fixed.
* 512x512x48 / 21 views: the gated schedule with the trained CNN (the
  learned inference configuration): gate record and RMSE gain.
* 8 x 1024x1024x64 / 25 views, learned plus batched: the batched
  accelerated reconstruction bit-identical to the unbatched one for a
  reconstruction, every gate decision recorded.
"""
import os

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


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def say(msg):
    print(f"[configs] {msg}", flush=True)


def test_small_workload_mlem_10_iterations_bit_exact(K, oracle_mod):
    from kevernals_amd import geometry as G
    from kevernals_amd.phantom import phantom
    from oracle.rng import poisson_data

    g = G.small_plain()
    o = oracle_mod.Oracle(g, oracle_mod.threads())
    mu = phantom(g, 1)
    # reference-discretisation data from Poisson line integrals (I0 = 1e4)
    ell = o.fp_a(mu)
    ell_hat = poisson_data(ell, 1e4, 1001)
    fp = o.fp_r(mu)
    data = np.where(ell > 0, fp * ell_hat / np.maximum(ell, np.float32(1e-30)), 0).astype(np.float32)
    ctx = K.Context(0, g)
    x = torch.ones(g.vol_shape, device="cuda")
    d = dev(data)
    for _ in range(10):
        ctx.mlem_iteration(x, d, 1.0, K.MODE_R)
    say("oracle: 10 MLEM iterations at 128x128x32 ...")
    ref = o.mlem(data, np.ones(g.vol_shape, np.float32), 10, 1.0)
    np.testing.assert_array_equal(host(x), ref)


@pytest.fixture(scope="module")
def full(K):
    from kevernals_amd import geometry as G
    from kevernals_amd.phantom import phantom

    g = G.full_plain()
    ctx = K.Context(0, g)
    mu = dev(phantom(g, 2, microcalcifications=200))
    data = torch.empty(g.proj_shape, device="cuda")
    ctx.forward_project(mu, data, K.MODE_R)
    torch.cuda.synchronize()
    return g, ctx, data


def test_full_size_sart_10_iterations_sampled(K, oracle_mod, full):
    g, ctx, data = full
    o = oracle_mod.Oracle(g, oracle_mod.threads())
    x = torch.ones(g.vol_shape, device="cuda")
    for _ in range(9):
        ctx.sart_iteration(x, data, 0.5)
    x9 = host(x)
    fp9 = torch.empty(g.proj_shape, device="cuda")
    ctx.forward_project(x, fp9, K.MODE_R)
    ctx.sart_iteration(x, data, 0.5)
    x10, fp9, d = host(x), host(fp9), host(data)
    r = rng(77)
    rays = r.integers(0, g.nrays, 2000)
    say("oracle: forward projection of 2000 rays of iteration 9")
    np.testing.assert_array_equal(fp9.ravel()[rays], o.fp_r_rays(x9, rays))
    corr = (np.float32(0.5) * (d - fp9)).astype(np.float32)  # Reconstructors.h:90-93
    vox = r.integers(0, g.nvox, 2000)
    say("oracle: backprojection of 2000 voxels of iteration 10")
    want = (x9.ravel()[vox] + o.bp_r_voxels(corr, vox)).astype(np.float32)
    np.testing.assert_array_equal(x10.ravel()[vox], want)
    assert np.abs(x10 - 1.0).max() > 0.01  # the iterations did move the volume


def test_full_size_shift_and_add_64_planes(K, oracle_mod, full):
    g, ctx, data = full
    planes = g.plane_heights()
    assert len(planes) == 64
    sid = float(abs(g.det[0, 2] - np.float32(g.src[:, 2].mean(dtype=np.float32))))
    out = torch.empty((64, g.nv, g.nu), device="cuda")
    ctx.shift_and_add(data, out, g.src[:, 1], planes, sid, 0.0, g.sy)
    got = host(out)
    pick = [0, 21, 42, 63]
    say("oracle: shift-and-add on planes 0, 21, 42, 63")
    rc, ref = oracle_mod.saa(host(data), g.src[:, 1], planes[pick], sid, 0.0, g.sy)
    assert rc == 0
    np.testing.assert_array_equal(got[pick], ref)


def test_batched_uneven_32_at_real_sizes(K):
    from kevernals_amd import geometry as G

    gs = G.batched_uneven(3, 32)
    assert max(g.nvox for g in gs) > 40 << 20 and min(g.nvox for g in gs) < 8 << 20
    b = K.Batch(gs)
    gen = torch.Generator(device="cuda").manual_seed(3)
    datas = []
    for i, g in enumerate(gs):
        ctx = b.context(i)
        mu = 0.02 + 0.03 * torch.rand(g.vol_shape, device="cuda", generator=gen)
        p = torch.empty(g.proj_shape, device="cuda")
        ctx.forward_project(mu, p, K.MODE_R)
        torch.cuda.synchronize()
        datas.append(p.cpu().numpy())
        b.set_data(i, datas[-1])
        b.set_volume(i, np.ones(g.vol_shape, np.float32))
        del mu, p
    say("32 reconstructions: 15 batched MLEM iterations")
    b.mlem(15, use_graph=True)
    sizes = [g.nvox for g in gs]
    for i in sorted({int(np.argmin(sizes)), int(np.argmax(sizes)), 17}):
        g = gs[i]
        ctx = K.Context(0, g)
        x = torch.ones(g.vol_shape, device="cuda")
        d = dev(datas[i])
        for _ in range(15):
            ctx.mlem_iteration(x, d, 1.0, K.MODE_R)
        say(f"reconstruction {i} ({g.nx}x{g.ny}x{g.nz}, {g.nviews} views) vs unbatched")
        np.testing.assert_array_equal(b.get_volume(i), host(x), err_msg=f"reconstruction {i}")
        del ctx, x, d
    # consistent data of a uniform volume: one iteration of recons::MLEM
    # keeps it exactly on every voxel fed only by rays that see the volume
    # (ratio exactly 1, Reconstructors.h:207-213), shrinks the voxels that
    # some dropped ray feeds with a zero ratio and zeroes the unseen ones --
    # in every reconstruction of the batch
    masks = []
    for i, g in enumerate(gs):
        ctx = b.context(i)
        u = torch.full(g.vol_shape, 0.03, device="cuda")
        p = torch.empty(g.proj_shape, device="cuda")
        ctx.forward_project(u, p, K.MODE_R)
        seen = torch.empty(g.vol_shape, device="cuda")
        ctx.back_project(torch.ones_like(p), seen, K.MODE_R)
        touched = torch.empty(g.vol_shape, device="cuda")
        ctx.back_project((p == 0).float(), touched, K.MODE_R)
        torch.cuda.synchronize()
        masks.append(((seen != 0).cpu().numpy(), (touched != 0).cpu().numpy()))
        b.set_data(i, p.cpu().numpy())
        b.set_volume(i, np.full(g.vol_shape, 0.03, np.float32))
        del u, p, seen, touched
    b.mlem(1, use_graph=True)
    for i, g in enumerate(gs):
        v = b.get_volume(i)
        seen, touched = masks[i]
        keep = seen & ~touched
        assert keep.mean() > 0.3, i
        assert np.all(v[keep] == np.float32(0.03)), i
        assert np.all(v[~seen] == 0), i
        assert np.all(v[touched] <= np.float32(0.03)), i


def _committed_cnn(K):
    wpath = os.path.join(os.path.dirname(K.__file__), "weights", "cnn_schedule_to_100.npz")
    with np.load(wpath, allow_pickle=False) as f:
        return f["params"].astype(np.float32), float(f["scale"])


def test_learned_inference_configuration(K):
    """512x512x48 / 21 views (phantom seed 6000, I0 = 1e4): the gated
    schedule with the trained CNN.  Every step outside MLEM passes the
    likelihood gate (recorded in the trace), the result is several times
    closer to the 100-iteration MLEM-A result than plain MLEM at the same
    iteration count, and the CNN step, when kept, does not raise the RMSE by
    more than rounding."""
    from kevernals_amd import geometry as G
    from kevernals_amd.learned import Cnn, accelerated_mlem, rmse, transmission_data
    from kevernals_amd.phantom import phantom

    g = G.learned_infer()
    ctx = K.Context(0, g)
    data = transmission_data(ctx, dev(phantom(g, 6000)), 1e4, 6000)
    x = torch.ones(g.vol_shape, device="cuda")
    for k in range(1, 101):
        ctx.mlem_iteration(x, data, 1.0, K.MODE_A)
        if k == 18:
            x18 = x.clone()
    params, scale = _committed_cnn(K)
    alone = accelerated_mlem(ctx, data, reference=x)
    learned = accelerated_mlem(ctx, data, cnn=Cnn(ctx, params, scale), reference=x)
    kinds = [t[0] for t in learned.trace]
    assert kinds[-1] in ("final-", "final+", "cnn+")
    assert learned.accepted + learned.rejected == 2
    r_plain, r_alone, r_learned = rmse(x18, x), alone.trace[-1][3], learned.trace[-1][3]
    say(f"512x512x48: RMSE vs iteration 100: plain 18 iterations {r_plain:.4f}, schedule {r_alone:.4f}, "
        f"schedule + CNN {r_learned:.4f} ({kinds[-1]})")
    assert kinds[-1] != "final-" and r_alone < r_plain / 4
    if kinds[-1] == "cnn+":
        # the kept step did not lower the likelihood of the jump's result
        assert learned.trace[-1][4] >= ctx.log_likelihood(data, alone.x)
        assert r_learned < r_alone * 1.05
    else:
        np.testing.assert_array_equal(host(learned.x), host(alone.x))


def test_learned_batched_full_size(K):
    """8 x 1024x1024x64 / 25 views (phantoms from kv_phantom: 20 boxes and
    200 spheres of 0.3-1 mm at 0.1/mm on a 0.02/mm background, seeds
    5000-5007; Poisson data at I0 = 5e3) in one mode-A batch: the C-ABI
    batched accelerated reconstruction with the CNN step gives reconstruction
    3 bit-identically the unbatched result, and records one gate decision
    (0, 1 or 2 steps kept) per reconstruction."""
    from kevernals_amd import geometry as G
    from kevernals_amd.learned import transmission_data

    g = G.full_plain()
    n = 8
    b = K.Batch([g] * n, mode=K.MODE_A)
    blf = np.array(g.blf())
    fov = np.array([g.sx * g.nx, g.sy * g.ny, g.sz * g.nz])
    datas = []
    for i in range(n):
        r = rng(5000 + i)
        paves = [[*(blf + fov * r.uniform(0.1, 0.6, 3)), *(fov * r.uniform(0.1, 0.3, 3)), r.uniform(0.01, 0.05)]
                 for _ in range(20)]
        spheres = [[*(blf + fov * r.uniform(0.1, 0.9, 3)), 0.5 * r.uniform(0.3, 1.0), 0.1] for _ in range(200)]
        ctx = b.context(i)
        mu = torch.empty(g.vol_shape, device="cuda")
        ctx.phantom(mu, 0.02, paves, spheres)
        data = transmission_data(ctx, mu, 5e3, 5000 + i)
        ctx.copy(b.buffers(i)[1], data, data.numel() * 4)
        torch.cuda.synchronize()
        if i == 3:
            datas.append(data)
        del mu
        b.set_volume(i, np.ones(g.vol_shape, np.float32))
    params, scale = _committed_cnn(K)
    d_params = dev(params)
    say("8 x 1024x1024x64: batched accelerated reconstruction")
    kept = b.mlem_accelerated(params=d_params, scale=scale)
    assert len(kept) == n and all(k in (0, 1, 2) for k in kept)
    say(f"steps kept per reconstruction: {kept}")
    ctx = K.Context(0, g)
    x = torch.ones(g.vol_shape, device="cuda")
    k3 = ctx.mlem_accelerated(x, datas[0], params=d_params, scale=scale)
    assert k3 == kept[3]
    np.testing.assert_array_equal(b.get_volume(3), host(x))
