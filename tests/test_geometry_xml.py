"""Reading the reference's XML geometry files (kevernals_amd/cpp/
TomoGeometryXml.cpp through kevernals_amd.geometry.from_xml) against the
reference's own reading (TomoGeometry.cpp:13-437 with its vendored tinyxml2,
built from its sources into oracle/_ref by oracle/Makefile, gathered as
Projector.cu:876-936).

Bar: identical kernel arguments, bit for bit, on every case the reference
reads (tests/golden/geometry_xml: the four workload geometries, a sub-region,
clamped and absent regions, odd pitches, extra sources, formatting variants:
CRLF, tabs, comments, CDATA, character references, number prefixes, signs,
exponents, fractional sizes), and the same accept/reject decision on every
malformed case it rejects; malformed files on which the reference crashes
are rejected with a message.  The expected values were recorded from the
reference build by tests/golden/make_geometry_xml.py; with the build present
they are also recomputed live.
"""
import ctypes
import json
import os

import numpy as np
import pytest

from kevernals_amd import geometry as G

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "geometry_xml")
REF_SO = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "libref_geometry.so")
EXPECTED = json.load(open(os.path.join(GOLD, "expected.json")))


def bits(a):
    return [format(int(v), "08x") for v in np.ascontiguousarray(a, np.float32).view(np.uint32).ravel()]


def ours(path):
    try:
        g = G.from_xml(path)
    except ValueError as e:
        return None, str(e)
    ints = [1, g.nx, g.ny, g.nz, g.nu, g.nv, g.nviews]
    return {"valid": 1, "ints": ints, "det": bits(g.det), "src": bits(g.src),
            "floats_head": bits(np.array([g.sx, g.sy, g.sz, g.pu, g.pv], np.float32))}, ""


@pytest.mark.parametrize("name", sorted(EXPECTED))
def test_reader_matches_the_reference_reading(name):
    got, err = ours(os.path.join(GOLD, f"{name}.xml"))
    want = EXPECTED[name]
    if not want["valid"]:
        assert got is None, f"{name}: the reference rejects this file"
        assert err
        return
    assert got is not None, err
    assert got["ints"] == want["ints"]
    assert got["floats_head"] == want["floats"][:5]
    assert got["det"] == want["det"] and got["src"] == want["src"]


def test_fulcrum_matches_the_reference_reading():
    lib = ctypes.CDLL(os.path.join(os.path.dirname(G.__file__), "lib", "libkevernals_dropin.so"))
    fn = lib.kvd_geometry_from_xml
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_char_p] + [ctypes.c_void_p] * 4 + [ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
    for name, want in EXPECTED.items():
        if not want["valid"]:
            continue
        ints = np.zeros(7, np.int32)
        floats = np.zeros(8, np.float32)
        fn(os.fsencode(os.path.join(GOLD, f"{name}.xml")), ints.ctypes.data, floats.ctypes.data, None, None, 0,
           None, 0)
        assert bits(floats) == want["floats"], name


@pytest.mark.skipif(not os.path.exists(REF_SO), reason="reference geometry build absent (oracle/_ref)")
@pytest.mark.parametrize("name", ["small", "train", "roi_sub", "roi_clamp", "roi_none", "format", "odd_pitch",
                                  "more_sources", "bad_roi_mismatch", "bad_number"])
def test_reader_matches_the_reference_build_live(name):
    lib = ctypes.CDLL(REF_SO)
    lib.ref_geometry_from_xml.restype = ctypes.c_int
    lib.ref_geometry_from_xml.argtypes = [ctypes.c_char_p] + [ctypes.c_void_p] * 4 + [ctypes.c_int]
    path = os.path.join(GOLD, f"{name}.xml")
    ints = np.zeros(7, np.int32)
    floats = np.zeros(8, np.float32)
    det = np.zeros((64, 3), np.float32)
    src = np.zeros((64, 3), np.float32)
    n = lib.ref_geometry_from_xml(os.fsencode(path), ints.ctypes.data, floats.ctypes.data, det.ctypes.data,
                                  src.ctypes.data, 64)
    got, err = ours(path)
    if n <= 0:
        assert got is None
        return
    assert got is not None, err
    assert got["ints"] == [int(v) for v in ints]
    assert got["det"] == bits(det[:n]) and got["src"] == bits(src[:n])


def write(tmp_path, text):
    p = tmp_path / "g.xml"
    p.write_text(text)
    return str(p)


def test_rejects_what_the_reference_crashes_on(tmp_path):
    """Inputs on which the reference dereferences null or indexes past a
    vector: this reader reports them instead."""
    base = open(os.path.join(GOLD, "small.xml")).read()
    cases = {
        "six values": base.replace("<roi>0 0 0 0 255 255</roi>", "<roi>0 0 255 255</roi>", 1),
        "has no text": base.replace("<center>0.0 0.0 -20.0</center>", "<center/>", 1),
        "empty region": base.replace("<roi>0 0 0 0 255 255</roi>", "<roi>200 200 0 0 100 100</roi>"),
        "pixel counts": base.replace("<pixelWidth>256</pixelWidth>", "<pixelWidth>-256</pixelWidth>"),
        "fewer sources": base.replace("<source>", "<skip>", 1).replace("</source>", "</skip>", 1)
        .replace("<skip>", "<!-- ", 1).replace("</skip>", " -->", 1),
    }
    for what, text in cases.items():
        with pytest.raises(ValueError) as e:
            G.from_xml(write(tmp_path, text))
        msg = str(e.value)
        assert msg and ("has no text" in msg or what in msg), (what, msg)


def test_comments_inside_lists_are_skipped(tmp_path):
    """The reference walks list children as nodes (a comment there crashes
    it); this reader walks elements, so a commented list reads as without
    the comment."""
    base = open(os.path.join(GOLD, "small.xml")).read()
    a = G.from_xml(write(tmp_path, base))
    b = G.from_xml(write(tmp_path, base.replace("<sources>", "<sources><!-- view 0 first -->", 1)))
    np.testing.assert_array_equal(a.src, b.src)
    np.testing.assert_array_equal(a.det, b.det)


def test_round_trip_keeps_values_and_reduces_sources(tmp_path):
    """to_xml then from_xml: volume, detector corners and (for a pitch whose
    product with the pixel count divides back exactly) the pitch come back
    unchanged; each source keeps its y while x and z become their FP32
    running means over the views (TomoTable.cpp:4-19)."""
    g = G.arc_geometry((96, 80, 24), (0.5, 0.5, 2.0), (200, 160), (0.5, 0.25), 9, 15.0)
    path = str(tmp_path / "rt.xml")
    G.to_xml(g, path, center=(0.0, 0.0, 0.0))
    h = G.from_xml(path)
    assert (h.nx, h.ny, h.nz, h.nu, h.nv, h.nviews) == (g.nx, g.ny, g.nz, g.nu, g.nv, g.nviews)
    assert (h.sx, h.sy, h.sz, h.pu, h.pv) == (g.sx, g.sy, g.sz, g.pu, g.pv)
    np.testing.assert_array_equal(h.det[:, :2], g.det[:, :2])
    np.testing.assert_array_equal(h.src[:, 1], g.src[:, 1])
    sx = sz = np.float32(0.0)
    for s in g.src:
        sx = np.float32(sx + s[0])
        sz = np.float32(sz + s[2])
    assert np.all(h.src[:, 0] == np.float32(sx / np.float32(g.nviews)))
    assert np.all(h.src[:, 2] == np.float32(sz / np.float32(g.nviews)))
    # the detector height: FP64 sum of the FP32 heights, mean in FP32
    assert np.all(h.det[:, 2] == np.float32(np.float32(g.det[:, 2].astype(np.float64).sum()) / np.float32(g.nviews)))
