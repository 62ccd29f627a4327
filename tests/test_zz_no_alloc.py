"""kv_shift_and_add_noalloc (no device allocation; staging in a fixed
per-GPU array) against kv_shift_and_add through the C ABI: bit-identical
output on two calls in a row, the second with a smaller input, and the
staging limits rejected with KV_ERR_INVALID before any device work."""
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


def _saa_pair(K, ctx, g, proj, planes, sid, fz):
    nplanes = len(planes)
    ref = torch.full((nplanes, g.nv, g.nu), float("nan"), device="cuda")
    got = torch.full((nplanes, g.nv, g.nu), float("nan"), device="cuda")
    ctx.shift_and_add(proj, ref, g.src[:, 1], planes, sid, fz, g.sy)
    print(f"[no_alloc] {g.nu}x{g.nv}, {g.nviews} views, {nplanes} planes", flush=True)
    ctx.shift_and_add(proj, got, g.src[:, 1], planes, sid, fz, g.sy, noalloc=True)
    torch.cuda.synchronize()
    return ref.cpu().numpy(), got.cpu().numpy()


def test_noalloc_matches_allocating_entry_point_twice(K):
    from kevernals_amd import geometry as G

    g = G.small_plain()  # 256^2 detector, 15 views, 32 planes
    ctx = K.Context(0, g)
    proj = torch.from_numpy(rng(91).random(g.proj_shape).astype(np.float32)).cuda()
    planes = g.plane_heights()
    sid = float(abs(g.det[0, 2] - np.float32(g.src[:, 2].mean(dtype=np.float32))))
    ref, got = _saa_pair(K, ctx, g, proj, planes, sid, 0.0)
    assert np.count_nonzero(ref) > 1000
    np.testing.assert_array_equal(got, ref)
    # second call, smaller: fewer planes on a smaller detector, other fulcrum
    g2 = G.scaled_down(g, 2)
    ctx2 = K.Context(0, g2)
    proj2 = torch.from_numpy(rng(92).random(g2.proj_shape).astype(np.float32)).cuda()
    planes2 = g2.plane_heights()[::4]
    ref2, got2 = _saa_pair(K, ctx2, g2, proj2, planes2, sid, 3.0)
    assert np.count_nonzero(ref2) > 100
    np.testing.assert_array_equal(got2, ref2)
    # and the first context again, after the smaller call changed the staging
    ref3, got3 = _saa_pair(K, ctx, g, proj, planes[:3], sid, -2.0)
    np.testing.assert_array_equal(got3, ref3)


def test_noalloc_limits_rejected(K):
    from kevernals_amd.geometry import arc_geometry

    g = arc_geometry((8, 8, 4), (1.0, 1.0, 1.0), (8, 8), (1.0, 1.0), 25, 20.0)
    ctx = K.Context(0, g)
    proj = torch.zeros(g.proj_shape, device="cuda")
    out = torch.zeros(8, device="cuda")  # never written: the limits are checked first
    for nplanes in (4097, 2700):  # planes > 4096; 25 x 2700 pairs > 65536
        planes = np.linspace(-10.0, 10.0, nplanes).astype(np.float32)
        with pytest.raises(K.KvError) as e:
            ctx.shift_and_add(proj, out, g.src[:, 1], planes, 1000.0, 0.0, g.sy, noalloc=True)
        assert e.value.status == 1 and "staging limits" in str(e.value)
    # the precondition exits of the reference stay the same
    with pytest.raises(K.KvError) as e:
        ctx.shift_and_add(proj, out, g.src[:, 1], [1.0], 1000.0, 0.0, g.sy, noalloc=True)
    assert e.value.status == 5
