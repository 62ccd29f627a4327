import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: long-running")


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle

    oracle.build()
    return oracle


def tiny_geometry(nviews=3, vol=(6, 5, 4), det=(12, 10), spacing=(1.0, 1.2, 2.0), pitch=(1.5, 1.7), arc=20.0,
                  sod=40.0, sid=70.0, shift=(0.3, -0.4)):
    from kevernals_amd.geometry import arc_geometry

    return arc_geometry(vol, spacing, det, pitch, nviews, arc, sod=sod, sid=sid, det_shift=shift)


def rng(seed=0):
    return np.random.Generator(np.random.PCG64(seed))
