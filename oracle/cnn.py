"""FP64 CPU oracle of the 3D residual CNN (forward, backprop, Adam).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the HIP CNN
kernels; the product path never uses it.  The network has no counterpart in
the reference, so this restatement is pinned by a library routine
(scipy.ndimage.correlate for every convolution), by brute-force sums on tiny
inputs and by central finite differences of the FP64 loss (tests/test_cnn_oracle.py).

Network (volumes float[nz][ny][nx], activations [nz][ny][nx][16]):
    a0 = s*x                                         (1 channel)
    a1 = relu(conv(W1, a0) + b1)      W1[16][1][27]
    a2 = relu(conv(W2, a1) + b2)      W2[16][16][27]
    a3 = relu(conv(W3, a2) + b3)      W3[16][16][27]
    z4 = conv(W4, a3) + b4            W4[1][16][27]
    y  = x + z4 / s
conv(W, a)[v][o] = sum_{t, i} W[o][i][t] * a[v + off(t)][i], zero outside the
volume, tap t = 9*(dz+1) + 3*(dy+1) + (dx+1) for offsets in {-1, 0, 1}.
Loss L = mean_v (s*y - s*target)^2.

With bf16=True the values are rounded to bfloat16 (round to nearest even)
where the GPU path stores or feeds them to the matrix cores: a0..a3, the
weights, dz4 and the backpropagated dz3..dz1; sums are FP64 here and FP32 on
the GPU.
"""
from __future__ import annotations

import numpy as np

C = 16
SIZES = [(16, 1), (16, 16), (16, 16), (1, 16)]  # (cout, cin) per layer
NPARAMS = sum(co * ci * 27 + co for co, ci in SIZES)


def offsets():
    """(dz, dy, dx) of tap t."""
    return [(t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1) for t in range(27)]


def bf16(a):
    """Round float32 values to bfloat16, nearest even; returned as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    b = f.view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
    return b.astype(np.uint32).view(np.float32).astype(np.float64)


def split(params):
    """Parameter blob -> [(W, b)] per layer, W[cout][cin][27]."""
    p = np.asarray(params, dtype=np.float64)
    out, k = [], 0
    for co, ci in SIZES:
        w = p[k:k + co * ci * 27].reshape(co, ci, 27)
        k += co * ci * 27
        b = p[k:k + co]
        k += co
        out.append((w, b))
    assert k == NPARAMS
    return out


def join(layers):
    return np.concatenate([np.concatenate([w.ravel(), b.ravel()]) for w, b in layers])


def init_params(seed: int = 0):
    """He-normal weights (fan-in 27*cin), zero biases; the last layer is
    scaled by 0.1 so the untrained network is close to the identity."""
    rng = np.random.Generator(np.random.PCG64(seed))
    layers = []
    for li, (co, ci) in enumerate(SIZES):
        w = rng.standard_normal((co, ci, 27)) * np.sqrt(2.0 / (27 * ci))
        if li == 3:
            w *= 0.1
        layers.append((w, np.zeros(co)))
    return join(layers).astype(np.float32)


def shift(a, d):
    """a[v + d] with zeros outside; a is [nz][ny][nx][c]."""
    out = np.zeros_like(a)
    src = []
    dst = []
    for ax, o in enumerate(d):
        n = a.shape[ax]
        if o >= 0:
            src.append(slice(o, n))
            dst.append(slice(0, n - o))
        else:
            src.append(slice(0, n + o))
            dst.append(slice(-o, n))
    out[tuple(dst)] = a[tuple(src)]
    return out


def conv(w, a):
    """conv(W, a)[v][o] = sum_{t,i} W[o][i][t] a[v + off(t)][i]."""
    co = w.shape[0]
    out = np.zeros(a.shape[:3] + (co,))
    for t, d in enumerate(offsets()):
        out += shift(a, d) @ w[:, :, t].T
    return out


def conv_t(w, g):
    """Adjoint of conv in its input: sum_{t,o} W[o][i][t] g[v - off(t)][o]."""
    ci = w.shape[1]
    out = np.zeros(g.shape[:3] + (ci,))
    for t, d in enumerate(offsets()):
        out += shift(g, tuple(-x for x in d)) @ w[:, :, t]
    return out


def wgrad(g, a):
    """dL/dW[o][i][t] = sum_v g[v][o] a[v + off(t)][i]; dL/db[o] = sum_v g[v][o]."""
    co, ci = g.shape[3], a.shape[3]
    gw = np.zeros((co, ci, 27))
    g2 = g.reshape(-1, co)
    for t, d in enumerate(offsets()):
        gw[:, :, t] = g2.T @ shift(a, d).reshape(-1, ci)
    return gw, g2.sum(0)


def forward(params, x, s, bf16_round=False):
    """Returns y and the intermediate values {a0..a3, z1..z4}."""
    r = bf16 if bf16_round else (lambda v: np.asarray(v, dtype=np.float64))
    layers = split(params)
    if bf16_round:
        layers = [(bf16(w), b) for w, b in layers]
    x = np.asarray(x, dtype=np.float64)
    a0 = r((np.float32(s) * x.astype(np.float32)) if bf16_round else s * x)[..., None]
    acts = {"a0": a0}
    a = a0
    for li in range(3):
        w, b = layers[li]
        z = conv(w, a) + b
        acts[f"z{li + 1}"] = z
        a = r(np.maximum(z, 0.0))
        acts[f"a{li + 1}"] = a
    w, b = layers[3]
    z4 = conv(w, a)[..., 0] + b[0]
    acts["z4"] = z4
    y = x + z4 / s
    return y, acts


def loss_and_grad(params, x, target, s, bf16_round=False):
    """L = mean (s*y - s*target)^2 and dL/dparams (blob order)."""
    r = bf16 if bf16_round else (lambda v: np.asarray(v, dtype=np.float64))
    layers = split(params)
    if bf16_round:
        layers = [(bf16(w), b) for w, b in layers]
    y, A = forward(params, x, s, bf16_round)
    n = y.size
    res = s * np.asarray(x, np.float64) + A["z4"] - s * np.asarray(target, np.float64)
    loss = float(np.mean(res * res))
    dz4 = r(2.0 * res / n)[..., None]
    grads = [None] * 4
    grads[3] = wgrad(dz4, A["a3"])
    g = dz4
    for li in (3, 2, 1):
        w, _ = layers[li]
        da = conv_t(w, g)
        g = r(np.where(A[f"a{li}"] > 0, da, 0.0))
        grads[li - 1] = wgrad(g, A[f"a{li - 1}"])
    return loss, join(grads), A


def adam(params, grad, m, v, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015, Algorithm 1) in FP64; returns new
    (params, m, v)."""
    p = np.asarray(params, np.float64)
    g = np.asarray(grad, np.float64)
    m = b1 * np.asarray(m, np.float64) + (1 - b1) * g
    v = b2 * np.asarray(v, np.float64) + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
