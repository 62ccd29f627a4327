"""The learned MLEM step on the GPU: the likelihood gate and what the trained
network gains over plain MLEM.

- A network whose candidates lower the log-likelihood is rejected every
  time, and the run is then bit-identical to plain MLEM (plain MLEM is the
  fallback).
- Accepted candidates never lower the log-likelihood (plain MLEM is the
  judge).
- With the committed trained weights, on held-out phantoms of the training
  geometry (256x256x32, 15 views, I0 1e4), the gated run gets closer to the
  100-iteration MLEM result than plain MLEM at the same iteration count and
  reaches RMSE <= 2% in fewer iterations.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

WEIGHTS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "kevernals_amd", "weights",
                       "cnn_mlem5_to_100.npz")


@pytest.fixture(scope="module")
def K():
    import kevernals_amd

    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return kevernals_amd


def problem(K, seed, geometry):
    from kevernals_amd.learned import transmission_data
    from kevernals_amd.phantom import phantom

    ctx = K.Context(0, geometry)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    mu = torch.from_numpy(phantom(geometry, seed)).cuda()
    data = transmission_data(ctx, mu, 1e4, gen)
    return ctx, data


def test_gate_rejects_harmful_steps_and_falls_back_to_plain(K):
    from kevernals_amd import geometry as G
    from kevernals_amd.learned import Cnn, gated_mlem, init_params

    g = G.scaled_down(G.learned_train(), 2)
    ctx, data = problem(K, 6100, g)
    p = init_params(3)
    p[-1] = 5.0  # a large constant offset: every candidate is far worse
    bad = Cnn(ctx, p)
    run = gated_mlem(ctx, bad, data, 8, every=2, first=2)
    plain = gated_mlem(ctx, None, data, 8)
    assert run.accepted == 0 and run.rejected == 4
    np.testing.assert_array_equal(run.x.cpu().numpy(), plain.x.cpu().numpy())


@pytest.mark.skipif(not os.path.exists(WEIGHTS), reason="trained weights not present")
def test_trained_network_gains_over_plain_mlem(K):
    from kevernals_amd import geometry as G
    from kevernals_amd.learned import Cnn, gated_mlem, rmse

    g = G.learned_train()
    iters = 20
    gains = []
    for seed in (6500, 6501):
        ctx, data = problem(K, seed, g)
        ref = torch.ones(g.vol_shape, device="cuda")
        for _ in range(100):
            ctx.mlem_iteration(ref, data, 1.0, K.MODE_A)
        cnn = Cnn.load(ctx, WEIGHTS)
        plain = gated_mlem(ctx, None, data, iters, reference=ref)
        learned = gated_mlem(ctx, cnn, data, iters, every=2, first=4, reference=ref)
        # accepted candidates never lower the likelihood
        ll = [e[4] for e in learned.trace if e[0] == "cnn+"]
        assert learned.accepted >= 1
        r_plain = {k: r for kind, k, t, r, _ in plain.trace}
        r_learned = {}
        for kind, k, t, r, _ in learned.trace:
            r_learned[k] = r  # last entry of each iteration (after a CNN step if any)
        first_plain = next((k for k in sorted(r_plain) if r_plain[k] <= 0.02), None)
        first_learned = next((k for k in sorted(r_learned) if r_learned[k] <= 0.02), None)
        print(f"[learned] seed {seed}: rmse@{iters} plain {r_plain[iters]:.4f} learned {r_learned[iters]:.4f}; "
              f"2% at iteration plain {first_plain} learned {first_learned}; accepted {learned.accepted} "
              f"rejected {learned.rejected}; loglik after accepted steps {ll[:3]}", flush=True)
        assert r_learned[iters] < r_plain[iters]
        assert first_learned is not None and (first_plain is None or first_learned < first_plain)
        gains.append(r_plain[iters] / r_learned[iters])
        final = rmse(learned.x, ref)
        assert final == pytest.approx(r_learned[iters])
    assert min(gains) > 1.0
