"""The accepted (ray, voxel) pair counts that the roofline's algorithmic bytes
use, committed in tests/golden/nnz/nnz.json by tests/golden/make_nnz.py from
the CPU oracle: the 128x128x32 workload's counts are recounted here; the
GPU tests compare the kernels' own counts (kv_count_pairs) with both
workloads' fixtures."""
import json
import os

import pytest

from kevernals_amd import geometry as G

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "nnz", "nnz.json")


def fixture():
    with open(FIXTURE) as f:
        return json.load(f)


def test_fixture_holds_both_workloads():
    fx = fixture()
    for name in ("small", "full"):
        assert fx[name]["mode"] == "R"
        assert fx[name]["forward"] > 0 and fx[name]["backprojector"] > 0
    # about 80 accepted segments per detector ray at full size (Notes.md:8
    # reports 132 voxels per ray on the authors' own volume): a sanity bound
    g = G.full_plain()
    assert 10 < fx["full"]["forward"] / g.nrays < 200


def test_small_workload_recounted(oracle_mod):
    fx = fixture()["small"]
    o = oracle_mod.Oracle(G.small_plain(), oracle_mod.threads())
    assert o.fp_nnz() == fx["forward"]
    assert o.bp_nnz() == fx["backprojector"]
