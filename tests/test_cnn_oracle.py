"""Pins the FP64 CNN oracle (oracle/cnn.py) before it is trusted as the
checker of the HIP CNN kernels: convolutions against scipy.ndimage.correlate,
the backprop conv against the adjoint identity, weight gradients against
brute-force sums, the whole gradient against central finite differences,
bf16 rounding and Adam against torch's CPU implementations."""
import numpy as np
import pytest
from scipy import ndimage

from oracle import cnn


def rand(shape, seed):
    return np.random.Generator(np.random.PCG64(seed)).standard_normal(shape)


def test_conv_matches_scipy_correlate():
    a = rand((5, 6, 7, 3), 1)
    w = rand((2, 3, 27), 2)
    got = cnn.conv(w, a)
    for o in range(2):
        ref = sum(ndimage.correlate(a[..., i], w[o, i].reshape(3, 3, 3), mode="constant", cval=0.0)
                  for i in range(3))
        np.testing.assert_allclose(got[..., o], ref, rtol=1e-12, atol=1e-12)


def test_tap_order_is_z_y_x():
    a = np.zeros((3, 3, 3, 1))
    a[2, 1, 0, 0] = 1.0  # dz=+1, dy=0, dx=-1 from the centre
    w = np.zeros((1, 1, 27))
    w[0, 0, 9 * 2 + 3 * 1 + 0] = 5.0
    assert cnn.conv(w, a)[1, 1, 1, 0] == 5.0
    assert cnn.offsets()[9 * 2 + 3 * 1 + 0] == (1, 0, -1)


def test_conv_t_is_the_adjoint():
    a = rand((4, 5, 6, 16), 3)
    g = rand((4, 5, 6, 16), 4)
    w = rand((16, 16, 27), 5)
    lhs = np.sum(cnn.conv(w, a) * g)
    rhs = np.sum(a * cnn.conv_t(w, g))
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)


def test_wgrad_brute_force():
    a = rand((3, 4, 5, 2), 6)
    g = rand((3, 4, 5, 3), 7)
    gw, gb = cnn.wgrad(g, a)
    offs = cnn.offsets()
    nz, ny, nx = a.shape[:3]
    ref = np.zeros_like(gw)
    for o in range(3):
        for i in range(2):
            for t, (dz, dy, dx) in enumerate(offs):
                s = 0.0
                for z in range(nz):
                    for y in range(ny):
                        for x in range(nx):
                            zz, yy, xx = z + dz, y + dy, x + dx
                            if 0 <= zz < nz and 0 <= yy < ny and 0 <= xx < nx:
                                s += g[z, y, x, o] * a[zz, yy, xx, i]
                ref[o, i, t] = s
    np.testing.assert_allclose(gw, ref, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(gb, g.reshape(-1, 3).sum(0), rtol=1e-12)


def test_gradient_matches_finite_differences():
    x = 0.5 + np.abs(rand((5, 6, 7), 8)) * 0.2
    t = x + 0.05 * rand((5, 6, 7), 9)
    s = 3.0
    p = cnn.init_params(1).astype(np.float64)
    p[-1] = 0.05
    layers = cnn.split(p)
    for li in range(3):  # non-zero biases exercise every bias gradient
        layers[li][1][:] = 0.1 * rand(layers[li][1].shape, 20 + li)
    p = cnn.join(layers)
    loss, g, _ = cnn.loss_and_grad(p, x, t, s)
    idx = np.random.Generator(np.random.PCG64(10)).choice(p.size, 60, replace=False)
    # include every bias and the last layer
    idx = np.unique(np.concatenate([idx, [432 + k for k in range(16)], [p.size - 1, p.size - 2]]))
    h = 1e-6
    for k in idx:
        pp = p.copy()
        pp[k] += h
        lp = cnn.loss_and_grad(pp, x, t, s)[0]
        pp[k] -= 2 * h
        lm = cnn.loss_and_grad(pp, x, t, s)[0]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g).max()) + 1e-5 * abs(g[k]), (k, fd, g[k])


def test_identity_when_last_layer_is_zero():
    x = np.abs(rand((4, 4, 4), 11))
    p = cnn.init_params(2)
    layers = cnn.split(p)
    layers[3][0][:] = 0
    layers[3][1][:] = 0
    y, _ = cnn.forward(cnn.join(layers), x, 5.0)
    np.testing.assert_array_equal(y, x)


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    v = np.concatenate([rand(10000, 12).astype(np.float32) * 10.0 ** np.arange(-20, 20, 0.004)[:10000],
                        np.float32([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -2.5, 0.0])])
    v = v.astype(np.float32)
    ref = torch.from_numpy(v).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(cnn.bf16(v), ref)
    assert cnn.bf16(np.float32([1.0 + 2.0 ** -8]))[0] == 1.0  # tie to even


def test_adam_matches_torch():
    torch = pytest.importorskip("torch")
    p0 = rand(50, 13)
    grads = [rand(50, 14 + k) for k in range(4)]
    tp = torch.tensor(p0, dtype=torch.float64, requires_grad=True)
    opt = torch.optim.Adam([tp], lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    p, m, v = p0.copy(), np.zeros(50), np.zeros(50)
    for k, g in enumerate(grads):
        tp.grad = torch.tensor(g)
        opt.step()
        p, m, v = cnn.adam(p, g, m, v, k + 1)
    np.testing.assert_allclose(p, tp.detach().numpy(), rtol=1e-12, atol=1e-15)


def test_param_count():
    assert cnn.NPARAMS == 16 * 27 + 16 + 2 * (16 * 16 * 27 + 16) + 16 * 27 + 1
    assert cnn.init_params(0).size == cnn.NPARAMS
