"""The learned MLEM step on the GPU: the likelihood gate and what the trained
network gains over plain MLEM.

- A network whose candidates lower the log-likelihood is rejected every
  time, and the run is then bit-identical to plain MLEM (plain MLEM is the
  fallback).
- Accepted candidates never lower the log-likelihood (plain MLEM is the
  arbiter).
- With the committed trained weights, on held-out phantoms of the training
  geometry (256x256x32, 15 views, I0 1e4), the accepted CNN step moves the
  estimate towards the 100-iteration MLEM result by more than four MLEM
  iterations' worth, and the gated run reaches 10% RMSE in about half the
  iterations of plain MLEM.
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
    mu = torch.from_numpy(phantom(geometry, seed)).cuda()
    data = transmission_data(ctx, mu, 1e4, seed)
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
    """The committed network (trained on 5-iteration -> 100-iteration pairs
    of this geometry) on held-out phantoms: its candidate after iteration 5
    is accepted and puts the estimate where plain MLEM only gets after about
    twice as many iterations; plain MLEM is no worse than the gated run at
    the end (the network predicts the 100-iteration result to ~10%, so the
    gain is early)."""
    from kevernals_amd import geometry as G
    from kevernals_amd.learned import Cnn, gated_mlem

    g = G.learned_train()
    for seed in (6500, 6501):
        ctx, data = problem(K, seed, g)
        ref = torch.ones(g.vol_shape, device="cuda")
        for _ in range(100):
            ctx.mlem_iteration(ref, data, 1.0, K.MODE_A)
        cnn = Cnn.load(ctx, WEIGHTS)
        plain = gated_mlem(ctx, None, data, 30, reference=ref)
        learned = gated_mlem(ctx, cnn, data, 30, every=2, first=5, reference=ref, patience=2)
        r_plain = [r for kind, k, t, r, _ in plain.trace]
        r_learned = {}
        for kind, k, t, r, _ in learned.trace:
            r_learned[k] = r  # last entry of each iteration (after its CNN step if any)
        lls = [e[4] for e in learned.trace if e[0] != "mlem"]
        assert learned.trace[5][0] == "cnn+"  # the first candidate is accepted
        assert learned.rejected <= 2  # patience stops the proposals
        it10 = next(k for k in sorted(r_learned) if r_learned[k] <= 0.10)
        p10 = next(i + 1 for i, r in enumerate(r_plain) if r <= 0.10)
        print(f"[learned] seed {seed}: after CNN {r_learned[5]:.4f} (plain@5 {r_plain[4]:.4f}); 10% at "
              f"iteration {it10} vs plain {p10}; @10 {r_learned[10]:.4f} vs {r_plain[9]:.4f}; @30 "
              f"{r_learned[30]:.4f} vs {r_plain[29]:.4f}; accepted {learned.accepted} rejected "
              f"{learned.rejected}; loglik at the gate {lls}", flush=True)
        assert r_learned[5] < r_plain[8]  # one CNN step is worth more than 4 MLEM iterations here
        assert r_learned[10] < 0.85 * r_plain[9]
        assert 2 * it10 <= p10 + 2
