"""The data path's oracle (oracle/data_path.py) on the CPU.

phantom_maker: bit-identical to the reference's PhantomMaker loop on its own
TomoVolume / Pave / Sphere classes (tests/golden/phantoms, written from the
reference build by tests/golden/make_phantoms.py, and recomputed live when
the build is present).  The cases fix the grid points (lower corners, not
centres), inclusive box bounds, the strict sphere test and the order of the
FP32 sums; a restatement that used voxel centres, exclusive bounds or <= on
the radius fails them (checked below by planting each mistake).

log_scale: the published vtkImageLogarithmicScale rule pinned by closed
forms: log(1 + (e^k - 1)) = k, odd symmetry, 0 at 0, the constant's
linearity, float32 rounding of the FP64 value.
"""
import ctypes
import json
import os

import numpy as np
import pytest

from oracle import data_path as D

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "phantoms")
REF_SO = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "libref_geometry.so")
CASES = json.load(open(os.path.join(GOLD, "cases.json")))


def oracle_case(c, **kw):
    return D.phantom_maker(c["shape"], c["spacing"], c["background"], c["paves"], c["spheres"], **kw)


@pytest.mark.parametrize("name", sorted(CASES))
def test_phantom_oracle_matches_the_reference_classes(name):
    want = np.load(os.path.join(GOLD, f"{name}.npy"), allow_pickle=False)
    got = oracle_case(CASES[name])
    np.testing.assert_array_equal(got, want)


@pytest.mark.skipif(not os.path.exists(REF_SO), reason="reference geometry build absent (oracle/_ref)")
def test_fixtures_match_the_reference_build_live():
    lib = ctypes.CDLL(REF_SO)
    lib.ref_phantom.restype = ctypes.c_int
    lib.ref_phantom.argtypes = [ctypes.c_int] * 3 + [ctypes.c_float] * 4 + [ctypes.c_void_p, ctypes.c_int,
                                                                             ctypes.c_void_p, ctypes.c_int,
                                                                             ctypes.c_void_p]
    for name, c in CASES.items():
        nx, ny, nz = c["shape"]
        out = np.zeros((nz, ny, nx), np.float32)
        pv = np.asarray(c["paves"], np.float32).reshape(-1, 7)
        sp = np.asarray(c["spheres"], np.float32).reshape(-1, 5)
        lib.ref_phantom(nx, ny, nz, *c["spacing"], c["background"], pv.ctypes.data, len(pv), sp.ctypes.data, len(sp),
                        out.ctypes.data)
        np.testing.assert_array_equal(out, np.load(os.path.join(GOLD, f"{name}.npy"), allow_pickle=False), name)


def test_fixtures_catch_plausible_mistakes(monkeypatch):
    """Voxel centres instead of lower corners, exclusive box bounds and <= on
    the radius each change at least one fixture."""
    want = {n: np.load(os.path.join(GOLD, f"{n}.npy"), allow_pickle=False) for n in CASES}
    grid = D.grid_points
    with monkeypatch.context() as m:
        m.setattr(D, "grid_points", lambda n, s: (grid(n, s) + np.float32(0.5 * s)).astype(np.float32))
        assert any(not np.array_equal(oracle_case(CASES[n]), want[n]) for n in CASES)
    with monkeypatch.context() as m:
        m.setattr(D, "almost_equal_relative", lambda a, b: np.zeros(np.broadcast(a, b).shape, bool))
        assert not np.array_equal(oracle_case(CASES["aligned"]), want["aligned"])
    c = CASES["sphere_edge"]
    le = c["spheres"] + []
    # a radius nudged up by one ulp admits the points at exactly the radius
    bumped = [[*s[:3], float(np.nextafter(np.float32(s[3]), np.float32(np.inf))), s[4]] for s in le]
    got = D.phantom_maker(c["shape"], c["spacing"], c["background"], c["paves"], bumped)
    assert not np.array_equal(got, want["sphere_edge"])


def test_log_scale_closed_forms():
    k = np.array([0.0, 0.5, 1.0, 2.0, 7.0])
    x = np.expm1(k)
    np.testing.assert_allclose(D.log_scale(x), k.astype(np.float32), rtol=1e-7, atol=0)
    np.testing.assert_array_equal(D.log_scale(-x), -D.log_scale(x))
    assert D.log_scale(np.array([0.0]))[0] == 0.0
    r = np.random.Generator(np.random.PCG64(2)).standard_normal(1000).astype(np.float32) * 100
    np.testing.assert_array_equal(D.log_scale(r, 2.0), (2.0 * np.sign(r) * np.log1p(np.abs(r.astype(np.float64))))
                                  .astype(np.float32))
    assert D.log_scale(r).dtype == np.float32
