"""Full-size checks at 1024x1024x64 voxels / 2048^2 x 25 views, where the
whole-volume oracle is too slow: size-independent properties on every ray
and voxel, and exact oracle values on seeded samples."""
import numpy as np
import pytest

from conftest import rng

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def setup():
    import kevernals_amd as K
    from kevernals_amd import geometry as G

    g = G.full_plain()
    ctx = K.Context(0, g)
    return K, g, ctx


def test_uniform_volume_and_projections_at_full_size(setup):
    K, g, ctx = setup
    vol = torch.full(g.vol_shape, 0.03, device="cuda")
    fp = torch.empty(g.proj_shape, device="cuda")
    ctx.forward_project(vol, fp, K.MODE_R)
    torch.cuda.synchronize()
    hit = fp != 0
    frac = hit.float().mean().item()
    dev = ((fp[hit] - 0.03).abs() / 0.03).max().item()
    print(f"[full] fp_r uniform: {frac:.4f} of rays hit, max rel dev {dev:.2e}")
    assert frac > 0.3 and dev < 2e-6
    proj = torch.full(g.proj_shape, 1.5, device="cuda")
    bp = torch.empty(g.vol_shape, device="cuda")
    ctx.back_project(proj, bp, K.MODE_R)
    torch.cuda.synchronize()
    hit = bp != 0
    dev = ((bp[hit] - 1.5).abs() / 1.5).max().item()
    print(f"[full] bp_r uniform: {hit.float().mean().item():.4f} of voxels hit, max rel dev {dev:.2e}")
    assert dev < 2e-6


def test_deterministic_runs_are_bitwise_equal(setup):
    K, g, ctx = setup
    gen = torch.Generator(device="cuda").manual_seed(0)
    vol = torch.rand(g.vol_shape, device="cuda", generator=gen)
    a = torch.empty(g.proj_shape, device="cuda")
    b = torch.empty(g.proj_shape, device="cuda")
    ctx.forward_project(vol, a, K.MODE_R)
    ctx.forward_project(vol, b, K.MODE_R)
    torch.cuda.synchronize()
    assert torch.equal(a, b)
    x = torch.empty(g.vol_shape, device="cuda")
    y = torch.empty(g.vol_shape, device="cuda")
    ctx.back_project(a, x, K.MODE_R)
    ctx.back_project(a, y, K.MODE_R)
    torch.cuda.synchronize()
    assert torch.equal(x, y)


def test_samples_match_oracle_at_full_size(setup, oracle_mod):
    K, g, ctx = setup
    o = oracle_mod.Oracle(g)
    r = rng(40)
    vol = (0.02 + 0.03 * r.random(g.vol_shape, dtype=np.float32)).astype(np.float32)
    proj = (0.5 + r.random(g.proj_shape, dtype=np.float32)).astype(np.float32)
    fp = torch.empty(g.proj_shape, device="cuda")
    bp = torch.empty(g.vol_shape, device="cuda")
    ctx.forward_project(torch.from_numpy(vol).cuda(), fp, K.MODE_R)
    ctx.back_project(torch.from_numpy(proj).cuda(), bp, K.MODE_R)
    torch.cuda.synchronize()
    rays = r.integers(0, g.nrays, size=20000)
    vox = r.integers(0, g.nvox, size=4000)
    got_fp = fp.view(-1)[torch.from_numpy(rays).cuda()].cpu().numpy()
    got_bp = bp.view(-1)[torch.from_numpy(vox).cuda()].cpu().numpy()
    ref_fp = o.fp_r_rays(vol, rays)
    ref_bp = o.bp_r_voxels(proj, vox)
    print(f"[full] fp samples exact {np.mean(got_fp == ref_fp):.6f}, bp samples exact {np.mean(got_bp == ref_bp):.6f}")
    np.testing.assert_array_equal(got_fp, ref_fp)
    np.testing.assert_array_equal(got_bp, ref_bp)
    # voxel sequences of sampled rays
    ids = rays[:300].astype(np.int32)
    for rid, (idx, w, cnt) in zip(ids.tolist(), ctx.trace_rays(ids, cap=g.nx + g.ny + g.nz + 3, mode=K.MODE_R)):
        p, rem = divmod(rid, g.nu * g.nv)
        v, u = divmod(rem, g.nu)
        ri, rw = o.ray_trace(p, u, v)
        np.testing.assert_array_equal(idx, ri)
        np.testing.assert_array_equal(w, rw)


def test_adjoint_at_full_size(setup):
    K, g, ctx = setup
    gen = torch.Generator(device="cuda").manual_seed(1)
    x = torch.rand(g.vol_shape, device="cuda", generator=gen)
    y = torch.rand(g.proj_shape, device="cuda", generator=gen)
    Ax = torch.empty(g.proj_shape, device="cuda")
    Aty = torch.empty(g.vol_shape, device="cuda")
    ctx.forward_project(x, Ax, K.MODE_A)
    ctx.back_project(y, Aty, K.MODE_A)
    lhs = torch.dot(Ax.double().view(-1), y.double().view(-1)).item()
    rhs = torch.dot(x.double().view(-1), Aty.double().view(-1)).item()
    rel = abs(lhs - rhs) / abs(lhs)
    print(f"[full] adjoint <Ax,y>={lhs:.12e} <x,A^T y>={rhs:.12e} rel {rel:.2e}")
    assert rel <= 1e-6
