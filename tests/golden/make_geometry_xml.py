"""Writes the XML geometry cases under tests/golden/geometry_xml/ and, when
the reference's geometry reading is built (oracle/_ref/libref_geometry.so,
from /root/reference by oracle/Makefile), records what it reads from each
into expected.json: the kernel arguments its projector would launch with
(float32 values as exact bit patterns), or that it rejects the file.

Only files the reference reads without crashing are recorded (it
dereferences null on some malformed inputs).

    python tests/golden/make_geometry_xml.py
"""
import ctypes
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "geometry_xml")
sys.path.insert(0, ROOT)

from kevernals_amd import geometry as G  # noqa: E402


def doc(vol=(128, 128, 32), scale=(0.5, 0.5, 2.0), center=(0.0, 0.0, 0.0), fulcrum=(0.0, 0.0, -20.0),
        sources=(), refs=(), width="153.6", height="153.6", pw="256", ph="256", rois=None, nl="\n",
        comments=False, raw=None):
    """The reference's file layout; every value is written as given (strings
    pass through), so cases can vary the text itself."""
    if raw:
        return raw

    def j(v):
        return v if isinstance(v, str) else " ".join(repr(float(x)) for x in v)

    res = vol if isinstance(vol, str) else f"{vol[2]} {vol[1]} {vol[0]}"
    sc = scale if isinstance(scale, str) else f"{scale[2]!r} {scale[1]!r} {scale[0]!r}"
    lines = ['<?xml version="1.0" encoding="UTF-8"?>', "<Geometry>"]
    if comments:
        lines.append("  <!-- acquisition geometry -->")
    lines += [f"  <Rotation><center>{j(fulcrum)}</center></Rotation>", "  <Grid>",
              f"    <center>{j(center)}</center>", f"    <scale>{sc}</scale>",
              f"    <resolution>{res}</resolution>", "  </Grid>"]
    if comments:
        lines.append("  <!-- sources, then detector -->")
    lines += ["  <XRay>", "    <sources>"] + [f"      <source>{j(s)}</source>" for s in sources]
    lines += ["    </sources>", "  </XRay>", "  <Camera>", f"    <totalWidth>{width}</totalWidth>",
              f"    <totalHeight>{height}</totalHeight>", f"    <pixelWidth>{pw}</pixelWidth>",
              f"    <pixelHeight>{ph}</pixelHeight>", "    <references>"]
    lines += [f"      <reference>{j(r)}</reference>" for r in refs]
    lines += ["    </references>", "  </Camera>", "  <Radios>"]
    if rois is None:
        rois = ["0 0 0 0 255 255"] * len(refs)
    lines += ["    <rois>"] + [f"      <roi>{r}</roi>" for r in rois] + ["    </rois>"]
    lines += ["  </Radios>", "</Geometry>", ""]
    return nl.join(lines)


def arc(g, center):
    c = np.asarray(center, np.float64)
    return ([tuple(s + c) for s in g.src.astype(np.float64)], [tuple(d + c) for d in g.det.astype(np.float64)])


def cases():
    small = G.small_plain()
    ctr = (10.5, -3.25, 7.0)
    s, r = arc(small, ctr)
    out = {}
    for name, g, c in (("small", small, ctr), ("full", G.full_plain(), (0.0, 0.0, 0.0)),
                       ("train", G.learned_train(), (1.0, 2.0, -3.0)), ("infer", G.learned_infer(), (0.0, 0.0, 0.0))):
        path = os.path.join(OUT, f"{name}.xml")
        G.to_xml(g, path, center=c, fulcrum=(0.0, 0.0, -20.0))
        out[name] = open(path).read()
    base = dict(center=ctr, sources=s, refs=r)
    out["roi_sub"] = doc(rois=["17 9 x x 200 180"] * 15, **base)
    out["roi_clamp"] = doc(rois=["-5 -7 0 0 300 999"] * 15, **base)
    out["roi_none"] = doc(rois=[], **base)
    out["format"] = doc(nl="\r\n", comments=True, width="+1.536e2", height="153.6mm", pw=" 256.9 ", ph="\t256\n",
                        center="&#49;0.5 -3.25 7", scale="2.0\t0.5 0.5", vol="32.9 128.5 128.99",
                        fulcrum="<![CDATA[0 0 -20]]>", sources=s, refs=r)
    out["more_sources"] = doc(center=ctr, sources=s + s[:2], refs=r)
    out["odd_pitch"] = doc(center=ctr, sources=s, refs=r, width="100.0", height="77.7", pw="333", ph="251",
                           rois=["3 5 0 0 300 240"] * 15)
    # rejected by the reference (it prints an error and returns)
    out["bad_zero_width"] = doc(width="0.0", **base)
    out["bad_pixel_zero"] = doc(pw="0", **base)
    out["bad_roi_mismatch"] = doc(rois=["0 0 0 0 255 255"] * 14 + ["1 0 0 0 255 255"], **base)
    out["bad_source_words"] = doc(center=ctr, sources=s[:3] + ["1.0 2.0"], refs=r)
    out["bad_no_sources"] = doc(center=ctr, sources=[], refs=r)
    out["bad_no_camera"] = doc(raw=doc(**base).replace("<Camera>", "<Kamera>").replace("</Camera>", "</Kamera>"))
    out["bad_unclosed"] = doc(raw=doc(**base).replace("</Grid>", ""))
    out["bad_number"] = doc(width="wide", **base)
    return out


def read_ref(lib, path, cap=64):
    ints = np.zeros(7, np.int32)
    floats = np.zeros(8, np.float32)
    det = np.zeros((cap, 3), np.float32)
    src = np.zeros((cap, 3), np.float32)
    n = lib.ref_geometry_from_xml(os.fsencode(path), ints.ctypes.data, floats.ctypes.data, det.ctypes.data,
                                  src.ctypes.data, cap)
    return n, ints, floats, det[:max(n, 0)], src[:max(n, 0)]


def bits(a):
    return [format(int(v), "08x") for v in np.ascontiguousarray(a, np.float32).view(np.uint32).ravel()]


def main():
    os.makedirs(OUT, exist_ok=True)
    for name, text in cases().items():
        with open(os.path.join(OUT, f"{name}.xml"), "w", newline="") as f:
            f.write(text)
    so = os.path.join(ROOT, "oracle", "_ref", "libref_geometry.so")
    if not os.path.exists(so):
        print("oracle/_ref/libref_geometry.so not built: expected.json not written")
        return 1
    lib = ctypes.CDLL(so)
    lib.ref_geometry_from_xml.restype = ctypes.c_int
    lib.ref_geometry_from_xml.argtypes = [ctypes.c_char_p] + [ctypes.c_void_p] * 4 + [ctypes.c_int]
    expected = {}
    for name in sorted(cases()):
        n, ints, floats, det, src = read_ref(lib, os.path.join(OUT, f"{name}.xml"))
        expected[name] = {"valid": int(ints[0]), "ints": [int(v) for v in ints], "floats": bits(floats),
                          "det": bits(det), "src": bits(src)} if n > 0 else {"valid": 0}
        print(name, "valid" if n > 0 else "rejected", ints.tolist())
    with open(os.path.join(OUT, "expected.json"), "w") as f:
        json.dump(expected, f, indent=1, sort_keys=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
