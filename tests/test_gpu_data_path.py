"""The data path on the GPU (kv_phantom, kv_log_scale).

Bar: kv_phantom bit-identical to the reference's PhantomMaker loop on its own
geometry classes (tests/golden/phantoms) and to the oracle on a full-size
volume's sampled slices; kv_log_scale within one float32 ulp of the FP64
oracle (the device's FP64 log and the host's may round the last FP64 bit
differently; the FP32 results agree except at float32 rounding ties).
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "phantoms")
CASES = json.load(open(os.path.join(GOLD, "cases.json")))


@pytest.fixture(scope="module")
def K():
    import kevernals_amd

    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return kevernals_amd


def ctx_for(K, shape, spacing, nviews=1):
    from kevernals_amd import geometry as G

    g = G.arc_geometry(tuple(shape), tuple(spacing), (8, 8), (1.0, 1.0), nviews, 10.0)
    return K.Context(0, g), g


@pytest.mark.parametrize("name", sorted(CASES))
def test_phantom_matches_the_reference_classes(K, name):
    c = CASES[name]
    ctx, g = ctx_for(K, c["shape"], c["spacing"])
    vol = torch.full(g.vol_shape, -1.0, device="cuda")
    ctx.phantom(vol, c["background"], c["paves"], c["spheres"])
    want = np.load(os.path.join(GOLD, f"{name}.npy"), allow_pickle=False)
    np.testing.assert_array_equal(vol.cpu().numpy(), want)


def test_phantom_at_full_size(K):
    """1024x1024x64 at 0.15x0.15x1 mm with the reference's example box and
    sphere scaled into the volume: sampled slices against the oracle, and
    the sphere's voxel count against its volume."""
    from oracle import data_path as D

    shape, spacing = (1024, 1024, 64), (0.15, 0.15, 1.0)
    paves = [[-25.0, -50.0, -20.0, 20.0, 40.0, 30.0, 0.03]]
    spheres = [[40.0, 30.0, 5.0, 15.0, 0.05], [0.0, 0.0, 0.0, 2.0, 0.1]]
    ctx, g = ctx_for(K, shape, spacing)
    vol = torch.empty(g.vol_shape, device="cuda")
    ctx.phantom(vol, 0.02, paves, spheres)
    got = vol.cpu().numpy()
    want = D.phantom_maker(shape, spacing, 0.02, paves, spheres)
    for z in (0, 17, 31, 32, 45, 63):
        np.testing.assert_array_equal(got[z], want[z], err_msg=f"slice {z}")
    inside = int((want == np.float32(np.float32(0.02) + np.float32(0.05))).sum())
    expect = 4.0 / 3.0 * np.pi * 15.0 ** 3 / (0.15 * 0.15 * 1.0)
    assert abs(inside - expect) / expect < 0.02


def test_log_scale_matches_oracle(K):
    from oracle import data_path as D

    r = np.random.Generator(np.random.PCG64(4))
    x = np.concatenate([r.standard_normal(1 << 20) * 10, r.uniform(0, 65535, 1 << 20), [0.0, -0.0, 1e-30, -1e-30,
                                                                                         3.4e38, -3.4e38]])
    x = x.astype(np.float32)
    d = torch.from_numpy(x).cuda()
    out = torch.empty_like(d)
    ctx = K.Context(0)
    for c in (1.0, 2.5):
        ctx.log_scale(d, out, c)
        torch.cuda.synchronize()
        got = out.cpu().numpy()
        want = D.log_scale(x, c)
        ulp = np.abs(got.view(np.int32).astype(np.int64) - want.view(np.int32).astype(np.int64))
        print(f"[data] log_scale c={c}: {np.mean(ulp == 0) * 100:.4f} % bit-exact, max {ulp.max()} ulp", flush=True)
        assert ulp.max() <= 1
    # in place
    y = d.clone()
    ctx.log_scale(y, y, 1.0)
    ctx.log_scale(d, out, 1.0)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(y.cpu().numpy(), out.cpu().numpy())


def test_rejects_bad_arguments(K):
    ctx, g = ctx_for(K, (8, 8, 4), (1.0, 1.0, 1.0))
    vol = torch.empty(g.vol_shape, device="cuda")
    with pytest.raises(K.KvError):  # one box announced, no list
        ctx._dev(ctx.lib.kv_phantom, 0.0, None, 1, None, 0, vol.data_ptr())
    with pytest.raises(K.KvError):
        ctx.log_scale(vol, vol, float("nan"))
