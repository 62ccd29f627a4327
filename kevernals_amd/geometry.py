"""Acquisition geometries of the tomosynthesis path.

A :class:`Geometry` is the kernel argument list of the reference projector
(KevLoq/Kevernals source/modules/reconstruction/Projector.cu:227-236):
volume size and voxel spacing, detector size and pixel pitch, and per view the
detector bottom-left corner (x, y) with its plane height z
(Projector.cu:917-923) and the source position (Projector.cu:930-936).  World
units are millimetres and the volume is centred on the origin
(source/modules/geometry/TomoGeometry.cpp:127-129).

The synthetic acquisitions used by tests and benchmarks put the sources on an
arc in the y-z plane around the isocentre (the reference moves sources along
y, Projector.cu:934, and shifts along y in shift-and-add,
Reconstructors.h:327-331) above a stationary flat detector.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Geometry:
    nx: int
    ny: int
    nz: int
    sx: float
    sy: float
    sz: float
    nu: int
    nv: int
    nviews: int
    pu: float
    pv: float
    det: np.ndarray = field(repr=False)  # (nviews, 3) float32
    src: np.ndarray = field(repr=False)  # (nviews, 3) float32

    def __post_init__(self):
        self.det = np.ascontiguousarray(self.det, dtype=np.float32).reshape(self.nviews, 3)
        self.src = np.ascontiguousarray(self.src, dtype=np.float32).reshape(self.nviews, 3)

    @property
    def vol_shape(self):
        """numpy shape of a volume: (nz, ny, nx), x fastest."""
        return (self.nz, self.ny, self.nx)

    @property
    def proj_shape(self):
        """numpy shape of a projection stack: (nviews, nv, nu), u fastest."""
        return (self.nviews, self.nv, self.nu)

    @property
    def nvox(self) -> int:
        return self.nx * self.ny * self.nz

    @property
    def nrays(self) -> int:
        return self.nviews * self.nu * self.nv

    def blf(self):
        """Bottom-left-front corner of the volume, as TomoGeometry.cpp:125-127
        computes it: -0.5 * (spacing * size) per axis, stored in float."""
        return tuple(
            float(np.float32(-0.5 * float(np.float32(np.float32(s) * np.float32(n)))))
            for s, n in ((self.sx, self.nx), (self.sy, self.ny), (self.sz, self.nz))
        )

    def voxel_positions(self):
        """Voxel lower-corner coordinates per axis, accumulated sequentially in
        float32 from the bottom-left-front corner (TomoVolume.cpp:197-215)."""
        out = []
        for b, s, n in zip(self.blf(), (self.sx, self.sy, self.sz), (self.nx, self.ny, self.nz)):
            pos = np.empty(n, dtype=np.float32)
            acc = np.float32(b)
            step = np.float32(s)
            for i in range(n):
                pos[i] = acc
                acc = np.float32(acc + step)
            out.append(pos)
        return out

    def plane_heights(self):
        """Slice heights used by shift-and-add: the volume's z positions
        (TomoGeometry::volumeZs, TomoVolume.cpp:209-215)."""
        return self.voxel_positions()[2]

    def scaled(self, **kw) -> "Geometry":
        d = dict(self.__dict__)
        d.update(kw)
        return Geometry(**d)


def arc_geometry(
    vol=(128, 128, 32),
    spacing=(0.5, 0.5, 2.0),
    det=(256, 256),
    pitch=(0.6, 0.6),
    nviews=15,
    arc_deg=15.0,
    sod=650.0,
    sid=1000.0,
    det_shift=(0.0, 0.0),
) -> Geometry:
    """Sources on a +-arc_deg arc of radius `sod` about the isocentre in the
    y-z plane; stationary detector at z = -(sid - sod), centred (optionally
    shifted by det_shift mm)."""
    nx, ny, nz = vol
    nu, nv = det
    pu, pv = pitch
    if nviews > 1:
        th = [math.radians(-arc_deg + 2.0 * arc_deg * k / (nviews - 1)) for k in range(nviews)]
    else:
        th = [0.0]
    src = np.array([[0.0, sod * math.sin(t), sod * math.cos(t)] for t in th], dtype=np.float64)
    zdet = -(sid - sod)
    d = np.array([[-nu * pu / 2.0 + det_shift[0], -nv * pv / 2.0 + det_shift[1], zdet]] * nviews, dtype=np.float64)
    return Geometry(nx, ny, nz, spacing[0], spacing[1], spacing[2], nu, nv, nviews, pu, pv,
                    d.astype(np.float32), src.astype(np.float32))


# ---- the workloads ----------------------------------------------------------

def small_plain() -> Geometry:
    """128x128x32 voxels of 0.5x0.5x2 mm, 256^2 detector at 0.6 mm, 15 views
    on +-15 deg, SID 1000 mm, source-to-isocentre 650 mm."""
    return arc_geometry((128, 128, 32), (0.5, 0.5, 2.0), (256, 256), (0.6, 0.6), 15, 15.0)


def full_plain() -> Geometry:
    """1024x1024x64 voxels of 0.15x0.15x1 mm, 2048^2 detector at 0.15 mm,
    25 views on +-20 deg."""
    return arc_geometry((1024, 1024, 64), (0.15, 0.15, 1.0), (2048, 2048), (0.15, 0.15), 25, 20.0)


def learned_train() -> Geometry:
    """Training pairs: 256x256x32 at 0.25x0.25x2 mm, 512^2 detector at 0.3 mm,
    15 views on +-15 deg."""
    return arc_geometry((256, 256, 32), (0.25, 0.25, 2.0), (512, 512), (0.3, 0.3), 15, 15.0)


def learned_infer() -> Geometry:
    """Learned inference: 512x512x48 at 0.25x0.25x(4/3) mm, 1024^2 detector at
    0.3 mm, 21 views on +-20 deg."""
    return arc_geometry((512, 512, 48), (0.25, 0.25, 64.0 / 48.0), (1024, 1024), (0.3, 0.3), 21, 20.0)


def batched_uneven(seed: int = 3, count: int = 32):
    """Uneven batch: volumes between 256x256x32 and 1024x768x80 (multiples of
    32/32/8) at 0.15x0.15x1 mm, detector 2*nx x 2*ny at 0.15 mm, views in
    {9..25} on +-U[10,25] deg."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for _ in range(count):
        nx = 32 * int(rng.integers(8, 33))
        ny = 32 * int(rng.integers(8, 25))
        nz = 8 * int(rng.integers(4, 11))
        nv_ = int(rng.integers(9, 26))
        arc = float(rng.uniform(10.0, 25.0))
        out.append(arc_geometry((nx, ny, nz), (0.15, 0.15, 1.0), (2 * nx, 2 * ny), (0.15, 0.15), nv_, arc))
    return out


def scaled_down(g: Geometry, f: int) -> Geometry:
    """Same acquisition with volume and detector sizes divided by f and
    spacings multiplied by f (same fields of view)."""
    return arc_geometry(
        (max(1, g.nx // f), max(1, g.ny // f), max(1, g.nz // f)),
        (g.sx * f, g.sy * f, g.sz * f),
        (max(1, g.nu // f), max(1, g.nv // f)),
        (g.pu * f, g.pv * f),
        g.nviews,
        float(np.degrees(np.arctan2(g.src[-1, 1], g.src[-1, 2]))) if g.nviews > 1 else 0.0,
        sod=float(np.hypot(g.src[0, 1], g.src[0, 2])),
        sid=float(np.hypot(g.src[0, 1], g.src[0, 2]) - g.det[0, 2]),
    )


# ---- the reference's XML geometry files -------------------------------------

def to_xml(g: Geometry, path: str, center=(0.0, 0.0, 0.0), fulcrum=(0.0, 0.0, 0.0)) -> None:
    """Write `g` as the reference's XML geometry file
    (source/modules/geometry/TomoGeometry.cpp:13-437): Grid/scale and
    Grid/resolution list (z, y, x) (the reader reverses them), sources and
    detector corners are written in a frame whose origin is `center` (the
    reader subtracts Grid/center), the detector is g's detector (totalWidth =
    nu * pu, pixelWidth = nu pixels) and the region read is all of it.
    Reading the file back gives g's values except where the reference's
    reading reduces them: the pitch becomes totalWidth / nu in FP32, and the
    sources' x and z become their means over the views (TomoTable.cpp:4-19)."""
    cx, cy, cz = (float(c) for c in center)

    def fmt(v):
        return repr(float(v))

    def xyz(p):
        return f"{fmt(p[0] + cx)} {fmt(p[1] + cy)} {fmt(p[2] + cz)}"

    lines = ['<?xml version="1.0" encoding="UTF-8"?>', "<Geometry>",
             f"  <Rotation><center>{fmt(fulcrum[0])} {fmt(fulcrum[1])} {fmt(fulcrum[2])}</center></Rotation>",
             "  <Grid>",
             f"    <center>{fmt(cx)} {fmt(cy)} {fmt(cz)}</center>",
             f"    <scale>{fmt(g.sz)} {fmt(g.sy)} {fmt(g.sx)}</scale>",
             f"    <resolution>{g.nz} {g.ny} {g.nx}</resolution>",
             "  </Grid>", "  <XRay>", "    <sources>"]
    lines += [f"      <source>{xyz(s)}</source>" for s in g.src.astype(np.float64)]
    lines += ["    </sources>", "  </XRay>", "  <Camera>",
              f"    <totalWidth>{fmt(np.float64(g.nu) * np.float64(g.pu))}</totalWidth>",
              f"    <totalHeight>{fmt(np.float64(g.nv) * np.float64(g.pv))}</totalHeight>",
              f"    <pixelWidth>{g.nu}</pixelWidth>", f"    <pixelHeight>{g.nv}</pixelHeight>",
              "    <references>"]
    lines += [f"      <reference>{xyz(d)}</reference>" for d in g.det.astype(np.float64)]
    lines += ["    </references>", "  </Camera>", "  <Radios>", "    <rois>"]
    lines += [f"      <roi>0 0 0 0 {g.nu - 1} {g.nv - 1}</roi>" for _ in range(g.nviews)]
    lines += ["    </rois>", "  </Radios>", "</Geometry>", ""]
    with open(path, "w") as f:
        f.write("\n".join(lines))


def from_xml(path: str) -> Geometry:
    """Read the reference's XML geometry file into the kernel arguments its
    projector would launch with (TomoGeometry.cpp:13-437 read by
    kevernals_amd/cpp/TomoGeometryXml.cpp, gathered as Projector.cu:876-940).
    Raises ValueError with the reason when the file is not a valid geometry."""
    import ctypes
    import os

    lib = ctypes.CDLL(os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libkevernals_dropin.so"))
    fn = lib.kvd_geometry_from_xml
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_char_p] + [ctypes.c_void_p] * 4 + [ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
    ints = np.zeros(7, np.int32)
    floats = np.zeros(8, np.float32)
    err = ctypes.create_string_buffer(512)
    np_ = fn(os.fsencode(path), ints.ctypes.data, floats.ctypes.data, None, None, 0, err, 512)
    if np_ <= 0:
        raise ValueError(f"{path}: {err.value.decode(errors='replace')}")
    det = np.zeros((np_, 3), np.float32)
    src = np.zeros((np_, 3), np.float32)
    fn(os.fsencode(path), ints.ctypes.data, floats.ctypes.data, det.ctypes.data, src.ctypes.data, np_, err, 512)
    _, nx, ny, nz, nu, nv, nviews = (int(v) for v in ints)
    return Geometry(nx, ny, nz, float(floats[0]), float(floats[1]), float(floats[2]), nu, nv, nviews,
                    float(floats[3]), float(floats[4]), det, src)
