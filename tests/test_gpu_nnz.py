"""The kernels' accepted-pair counts (kv_count_pairs, the numerators of the
roofline's algorithmic bytes) against the oracle's counts committed in
tests/golden/nnz/nnz.json: every ray and every voxel of the 128x128x32 and
1024x1024x64 workloads (Projector.cu:356-366 for the forward pass,
:710-729 for the backprojector)."""
import pytest

from test_nnz import fixture

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.mark.parametrize("name", ["small", "full"])
def test_kernel_pair_counts_equal_oracle(name):
    import kevernals_amd as K
    from kevernals_amd import geometry as G

    assert torch.cuda.is_available()
    g = G.small_plain() if name == "small" else G.full_plain()
    ctx = K.Context(0, g)
    fp, bp = ctx.count_pairs(K.MODE_R)
    fx = fixture()[name]
    print(f"[nnz] {name}: forward {fp} (oracle {fx['forward']}), backprojector {bp} (oracle {fx['backprojector']})",
          flush=True)
    assert fp == fx["forward"]
    assert bp == fx["backprojector"]
