"""Writes the phantom cases of tests/golden/phantoms/: cases.json (volume
size, spacing, background, boxes, spheres) and, when the reference's
geometry layer is built (oracle/_ref/libref_geometry.so, from /root/reference
by oracle/Makefile), <case>.npy with the densities the reference's
PhantomMaker loop gives on its own TomoVolume grid with its own
Pave::Contains / Sphere::Contains (float32[nz][ny][nx]).

    python tests/golden/make_phantoms.py
"""
import ctypes
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "phantoms")


def cases():
    r = np.random.Generator(np.random.PCG64(17))
    many_p = [[float(v) for v in (*(r.uniform(-6, 4, 3)), *(r.uniform(0.5, 5, 3)), r.uniform(-1, 1))]
              for _ in range(300)]
    many_s = [[float(v) for v in (*(r.uniform(-6, 6, 3)), r.uniform(0.3, 3), r.uniform(-1, 1))]
              for _ in range(300)]
    return {
        # the reference's own example shapes (mainPhantomMakerTest.cpp:36-41),
        # default background 50 (PhantomMaker.h)
        "example": dict(shape=[40, 40, 20], spacing=[5.0, 5.0, 25.0], background=50.0,
                        paves=[[-25.0, -50.0, 207.0, 20.0, 40.0, 30.0, 100.0]],
                        spheres=[[60.0, -100.0, 220.0, 15.0, 300.0]]),
        # box faces exactly on grid points: inclusive bounds
        "aligned": dict(shape=[16, 12, 8], spacing=[0.5, 0.5, 2.0], background=1.0,
                        paves=[[-1.5, -2.0, -4.0, 2.0, 1.5, 4.0, 10.0], [0.0, 0.0, 0.0, 0.5, 0.5, 2.0, 0.25]],
                        spheres=[]),
        # grid points exactly at the radius: strict <
        "sphere_edge": dict(shape=[16, 12, 8], spacing=[0.5, 0.5, 2.0], background=0.0, paves=[],
                            spheres=[[0.0, 0.0, 0.0, 1.0, 7.0], [-2.0, -1.0, -4.0, 2.0, 3.0],
                                     [0.5, 0.5, 2.0, 0.5, 1.0]]),
        # more shapes than one LDS chunk, overlapping (order of the FP32 sums)
        "many": dict(shape=[24, 20, 10], spacing=[0.5, 0.5, 1.25], background=0.02, paves=many_p, spheres=many_s),
        # spacings that are not binary fractions (sequential sums round)
        "inexact": dict(shape=[64, 48, 16], spacing=[0.15, 0.15, 1.0], background=0.02,
                        paves=[[-2.25, -1.05, -3.0, 3.3, 2.1, 4.0, 0.03]],
                        spheres=[[0.6, -0.45, 1.0, 1.2, 0.05], [-3.0, 2.4, -2.0, 0.9, 0.08]]),
    }


def main():
    os.makedirs(OUT, exist_ok=True)
    cs = cases()
    with open(os.path.join(OUT, "cases.json"), "w") as f:
        json.dump(cs, f, indent=1)
    so = os.path.join(ROOT, "oracle", "_ref", "libref_geometry.so")
    if not os.path.exists(so):
        print("oracle/_ref/libref_geometry.so not built: expected volumes not written")
        return 1
    lib = ctypes.CDLL(so)
    lib.ref_phantom.restype = ctypes.c_int
    lib.ref_phantom.argtypes = [ctypes.c_int] * 3 + [ctypes.c_float] * 4 + [ctypes.c_void_p, ctypes.c_int,
                                                                             ctypes.c_void_p, ctypes.c_int,
                                                                             ctypes.c_void_p]
    for name, c in cs.items():
        nx, ny, nz = c["shape"]
        out = np.zeros((nz, ny, nx), np.float32)
        pv = np.asarray(c["paves"], np.float32).reshape(-1, 7)
        sp = np.asarray(c["spheres"], np.float32).reshape(-1, 5)
        n = lib.ref_phantom(nx, ny, nz, *c["spacing"], c["background"], pv.ctypes.data, len(pv), sp.ctypes.data,
                            len(sp), out.ctypes.data)
        assert n == out.size
        np.save(os.path.join(OUT, f"{name}.npy"), out, allow_pickle=False)
        print(name, out.shape, "distinct values:", len(np.unique(out)))
    return 0


if __name__ == "__main__":
    sys.exit(main())
