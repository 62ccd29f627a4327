"""HIP CNN kernels (MFMA, bf16 operands, FP32 accumulation) against the FP64
oracle of oracle/cnn.py.

Layer by layer, each stage is recomputed by the oracle from the GPU's own
stored input to that stage, so the only differences are FP32 vs FP64
summation: stored bf16 activations and gradients must agree to within one
bf16 ulp (a rounding flip at a tie) and almost all exactly; weight
gradients within 1e-4 of the layer's largest entry.  End to end, the bf16
network is compared with the unrounded FP64 network (the bf16 error is
stated, not a parity bar), Adam with the FP64 update, and a short training
run must lower the loss.
"""
import numpy as np
import pytest

from conftest import rng

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

SHAPE = (10, 19, 37)  # nz, ny, nx: partial bricks in every direction
SCALE = 20.0


@pytest.fixture(scope="module")
def K():
    import kevernals_amd

    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return kevernals_amd


def bf_of(u16):
    return (u16.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def ulp_diff(a, b):
    """Distance in bf16 ulps of two arrays of bf16-representable values."""
    ia = (a.astype(np.float32).view(np.int32) >> 16).astype(np.int64)
    ib = (b.astype(np.float32).view(np.int32) >> 16).astype(np.int64)
    ia = np.where(ia < 0, -(ia & 0x7FFF), ia)
    ib = np.where(ib < 0, -(ib & 0x7FFF), ib)
    return np.abs(ia - ib)


def inputs(seed=0, shape=SHAPE):
    r = rng(100 + seed)
    x = (0.02 + 0.06 * r.random(shape)).astype(np.float32)
    t = (x + 0.01 * r.standard_normal(shape)).astype(np.float32)
    from oracle import cnn

    p = cnn.init_params(seed).astype(np.float64)
    layers = cnn.split(p)
    for li in range(4):
        layers[li][1][:] = 0.05 * r.standard_normal(layers[li][1].shape)
    layers[3][0][:] *= 5.0  # a visible correction
    return x, t, cnn.join(layers).astype(np.float32)


def run_gradient(K, x, t, params):
    ctx = K.Context(0)
    nb = K.cnn_workspace_size(SHAPE, True)
    work = torch.zeros(nb, dtype=torch.uint8, device="cuda")
    grad = torch.zeros(K.CNN_NPARAMS, device="cuda")
    loss = ctx.cnn_gradient(torch.from_numpy(params).cuda(), SCALE, torch.from_numpy(x).cuda(),
                            torch.from_numpy(t).cuda(), work, grad)
    torch.cuda.synchronize()
    return loss, grad.cpu().numpy(), work.cpu().numpy()


def region(work, off, nbytes, dtype):
    return work[off:off + nbytes].view(dtype)


def assert_bf16_close(name, got, ref):
    d = ulp_diff(got, ref)
    exact = float(np.mean(d == 0))
    print(f"[cnn] {name}: exact {exact:.6f}, max ulp {int(d.max())}", flush=True)
    assert d.max() <= 1, name
    assert exact >= 0.995, name


def test_layers_match_oracle(K, oracle_mod):
    from oracle import cnn

    x, t, params = inputs(0)
    loss, grad, work = run_gradient(K, x, t, params)
    V = x.size
    offs = K.cnn_workspace_offsets(SHAPE)
    a1, a2, a3 = (bf_of(region(work, offs[i], V * 32, np.uint16)).reshape(SHAPE + (16,)) for i in range(3))
    dz4 = region(work, offs[3], V * 4, np.float32).reshape(SHAPE).astype(np.float64)
    dz3, dz2, dz1 = (bf_of(region(work, offs[i], V * 32, np.uint16)).reshape(SHAPE + (16,)) for i in (4, 5, 6))
    L = [(cnn.bf16(w), b) for w, b in cnn.split(params)]
    a0 = cnn.bf16(np.float32(SCALE) * x)[..., None]

    # forward
    assert_bf16_close("a1", a1, cnn.bf16(np.maximum(cnn.conv(L[0][0], a0) + L[0][1], 0)))
    assert_bf16_close("a2", a2, cnn.bf16(np.maximum(cnn.conv(L[1][0], a1) + L[1][1], 0)))
    assert_bf16_close("a3", a3, cnn.bf16(np.maximum(cnn.conv(L[2][0], a2) + L[2][1], 0)))
    assert a1.any() and a2.any() and a3.any()
    z4 = cnn.conv(L[3][0], a3)[..., 0] + L[3][1][0]
    res = SCALE * x.astype(np.float64) + z4 - SCALE * t
    ref_loss = float(np.mean(res * res))
    print(f"[cnn] loss gpu {loss:.9e} oracle {ref_loss:.9e}", flush=True)
    assert abs(loss - ref_loss) <= 1e-5 * ref_loss
    tol = 2.0 / V * 4e-6 * SCALE * max(np.abs(x).max(), np.abs(t).max())
    np.testing.assert_allclose(dz4, 2.0 * res / V, rtol=1e-5, atol=tol)

    # backward
    g4 = cnn.bf16(dz4.astype(np.float32))[..., None]
    assert_bf16_close("dz3", dz3, cnn.bf16(np.where(a3 > 0, cnn.conv_t(L[3][0], g4), 0.0)))
    assert_bf16_close("dz2", dz2, cnn.bf16(np.where(a2 > 0, cnn.conv_t(L[2][0], dz3), 0.0)))
    assert_bf16_close("dz1", dz1, cnn.bf16(np.where(a1 > 0, cnn.conv_t(L[1][0], dz2), 0.0)))
    assert dz1.any()

    # weight and bias gradients from the GPU's own stored operands
    ref = cnn.join([cnn.wgrad(dz1, a0), cnn.wgrad(dz2, a1), cnn.wgrad(dz3, a2), cnn.wgrad(g4, a3)])
    got_layers, ref_layers = cnn.split(grad), cnn.split(ref)
    for li, ((gw, gb), (rw, rb)) in enumerate(zip(got_layers, ref_layers)):
        scale = max(np.abs(rw).max(), np.abs(rb).max())
        err = max(np.abs(gw - rw).max(), np.abs(gb - rb).max()) / scale
        print(f"[cnn] layer {li + 1} weight/bias gradient: max err {err:.2e} of max |g|", flush=True)
        assert err <= 1e-4, li


def test_forward_matches_training_pass_and_oracle(K, oracle_mod):
    from oracle import cnn

    x, t, params = inputs(1)
    ctx = K.Context(0)
    work = torch.empty(K.cnn_workspace_size(SHAPE, False), dtype=torch.uint8, device="cuda")
    y = torch.empty(SHAPE, device="cuda")
    ctx.cnn_forward(torch.from_numpy(params).cuda(), SCALE, torch.from_numpy(x).cuda(), y, work)
    got = y.cpu().numpy().astype(np.float64)
    ref_bf, _ = cnn.forward(params, x, SCALE, bf16_round=True)
    ref64, _ = cnn.forward(params, x, SCALE, bf16_round=False)
    corr = np.abs(ref_bf - x)
    err = np.abs(got - ref_bf)
    print(f"[cnn] forward vs bf16 oracle: max err {err.max():.3e} (max correction {corr.max():.3e}); "
          f"bf16 vs fp64 network: max {np.abs(ref_bf - ref64).max():.3e}", flush=True)
    # a rounding flip of one bf16 activation (a tie decided differently by
    # FP32 and FP64 sums) spreads through the later layers' 3x3x3 windows to
    # a few percent of the outputs: bound the worst case relative to the
    # network's correction and require the typical voxel to agree closely
    assert err.max() <= 5e-3 * corr.max()
    assert np.median(err) <= 1e-5 * corr.max()
    # the inference pass equals the training pass's output
    loss, _, work_t = run_gradient(K, x, t, params)
    V = x.size
    dz4 = region(work_t, K.cnn_workspace_offsets(SHAPE)[3], V * 4, np.float32).reshape(SHAPE)
    res = dz4.astype(np.float64) * V / 2.0
    y_from_train = t.astype(np.float64) + res / SCALE
    np.testing.assert_allclose(got, y_from_train, rtol=0, atol=1e-5 * np.abs(got).max())
    # the output may not alias the input: rejected before any launch
    xd = torch.from_numpy(x).cuda()
    with pytest.raises(K.KvError):
        ctx.cnn_forward(torch.from_numpy(params).cuda(), SCALE, xd, xd, work)
    with pytest.raises(K.KvError):
        ctx.cnn_forward(torch.from_numpy(params).cuda(), SCALE, xd, y, work[:16])


@pytest.mark.parametrize("shape", [(9, 11, 21), (70, 9, 18)])
def test_forward_z_chunks_against_oracle(K, oracle_mod, shape):
    """Depths that exercise the convolution's z-walk: 9 slices (the last
    4-slice chunk starts on the last slice, so its new halo planes all lie
    past the volume) and 70 slices (a workgroup's 16 chunks end at slice 64
    and a second workgroup starts a fresh halo ring)."""
    from oracle import cnn

    x, _, params = inputs(3, shape)
    ctx = K.Context(0)
    work = torch.empty(K.cnn_workspace_size(shape, False), dtype=torch.uint8, device="cuda")
    y = torch.empty(shape, device="cuda")
    ctx.cnn_forward(torch.from_numpy(params).cuda(), SCALE, torch.from_numpy(x).cuda(), y, work)
    got = y.cpu().numpy().astype(np.float64)
    ref_bf, _ = cnn.forward(params, x, SCALE, bf16_round=True)
    corr = np.abs(ref_bf - x)
    err = np.abs(got - ref_bf)
    print(f"[cnn] forward {shape} vs bf16 oracle: max err {err.max():.3e} (max correction {corr.max():.3e})",
          flush=True)
    assert err.max() <= 5e-3 * corr.max()
    assert np.median(err) <= 1e-5 * corr.max()


def test_gradient_vs_fp64_network(K, oracle_mod):
    from oracle import cnn

    x, t, params = inputs(2)
    loss, grad, _ = run_gradient(K, x, t, params)
    l64, g64, _ = cnn.loss_and_grad(params, x, t, SCALE, bf16_round=False)
    cos = float(np.dot(grad, g64) / (np.linalg.norm(grad) * np.linalg.norm(g64)))
    rel = float(np.linalg.norm(grad - g64) / np.linalg.norm(g64))
    print(f"[cnn] bf16 gradient vs fp64 network: cosine {cos:.5f}, rel L2 {rel:.3e}; loss {loss:.6e} vs {l64:.6e}",
          flush=True)
    assert cos >= 0.995
    assert abs(loss - l64) <= 2e-2 * l64


def test_gradient_is_deterministic(K):
    x, t, params = inputs(3)
    l1, g1, _ = run_gradient(K, x, t, params)
    l2, g2, _ = run_gradient(K, x, t, params)
    assert l1 == l2
    np.testing.assert_array_equal(g1, g2)


def test_adam_matches_fp64(K, oracle_mod):
    from oracle import cnn

    r = rng(7)
    n = K.CNN_NPARAMS
    p = r.standard_normal(n).astype(np.float32)
    m = np.zeros(n, np.float32)
    v = np.zeros(n, np.float32)
    ctx = K.Context(0)
    dp, dm, dv = (torch.from_numpy(a.copy()).cuda() for a in (p, m, v))
    pr, mr, vr = p.astype(np.float64), m.astype(np.float64), v.astype(np.float64)
    for step in range(1, 6):
        g = r.standard_normal(n).astype(np.float32) * 10.0 ** r.uniform(-6, 0, n).astype(np.float32)
        ctx.cnn_adam(dp, torch.from_numpy(g).cuda(), dm, dv, step, 1e-3)
        pr, mr, vr = cnn.adam(pr, g, mr, vr, step)
    torch.cuda.synchronize()
    np.testing.assert_allclose(dp.cpu().numpy(), pr, rtol=1e-6, atol=1e-7)


def test_short_training_lowers_loss(K, oracle_mod):
    from oracle import cnn

    r = rng(9)
    clean = (0.03 + 0.02 * (np.indices(SHAPE).sum(0) % 7 == 0)).astype(np.float32)
    x = (clean + 0.004 * r.standard_normal(SHAPE)).astype(np.float32)
    ctx = K.Context(0)
    params = torch.from_numpy(cnn.init_params(5)).cuda()
    m = torch.zeros_like(params)
    v = torch.zeros_like(params)
    grad = torch.zeros_like(params)
    work = torch.empty(K.cnn_workspace_size(SHAPE, True), dtype=torch.uint8, device="cuda")
    xd, td = torch.from_numpy(x).cuda(), torch.from_numpy(clean).cuda()
    losses = []
    for step in range(1, 61):
        losses.append(ctx.cnn_gradient(params, SCALE, xd, td, work, grad))
        ctx.cnn_adam(params, grad, m, v, step, 1e-3)
    print(f"[cnn] training loss {losses[0]:.4e} -> {losses[-1]:.4e}", flush=True)
    assert losses[-1] < 0.5 * losses[0]  # the FP64 oracle run of the same loop reaches 0.18x


@pytest.mark.parametrize("shape", [(1, 5, 7), (4, 8, 16), (5, 9, 70), (9, 1, 3), (2, 17, 1)])
def test_forward_at_volume_edges(K, oracle_mod, shape):
    """Volumes thinner than a brick in every direction, one slice, one row,
    one column: halo voxels outside the volume read as zero (the staging's
    out-of-range buffer loads) and partial bricks are stored exactly."""
    from oracle import cnn

    x, _, params = inputs(3, shape)
    ctx = K.Context(0)
    work = torch.empty(K.cnn_workspace_size(shape, False), dtype=torch.uint8, device="cuda")
    y = torch.full(shape, float("nan"), device="cuda")
    ctx.cnn_forward(torch.from_numpy(params).cuda(), SCALE, torch.from_numpy(x).cuda(), y, work)
    got = y.cpu().numpy().astype(np.float64)
    assert np.isfinite(got).all()
    ref_bf, _ = cnn.forward(params, x, SCALE, bf16_round=True)
    corr = np.abs(ref_bf - x)
    err = np.abs(got - ref_bf)
    print(f"[cnn] {shape}: max err {err.max():.3e} of max correction {corr.max():.3e}", flush=True)
    assert err.max() <= 5e-3 * corr.max()
