"""The drop-in's reference-tree build (KV_USE_REFERENCE_TREE): the C++
Projector / recons:: compile and generate code against the reference's own
TomoGeometry, Result (outcome) and error-code headers
(source/modules/geometry/TomoGeometry.h, source/commons/Result.h,
source/modules/reconstruction/ProjectorErrorCode.h,
ReconstructorsErrorCode.h), as INTEGRATION.md's in-tree switch uses them.
Runs where the reference tree is present (this check reads it at its
location and copies nothing)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/source"


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference tree absent")
def test_dropin_compiles_inside_the_reference_tree():
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "kevernals_amd", "csrc"), "reftree-check"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
