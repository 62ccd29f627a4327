"""CPU oracle of the data path (kv_phantom, kv_log_scale).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the HIP
kernels; the product path never uses it.

phantom_maker restates PhantomMaker::GetPhantom
(source/modules/dataHandling/PhantomMaker.cpp:13-29) in numpy FP32: each
voxel's grid point is its lower corner, sequential FP32 sums of the spacing
from the bottom-left-front corner -0.5 * (n * spacing) (TomoVolume.cpp:193-229,
TomoGeometry.cpp:121-127); it starts at the background and adds, in list
order, each box that contains the point (Pave::Contains, Pave.h:51-59,
bounds inclusive up to AlmostEqualRelative, commons/Maths.h:25-34) and then
each sphere (Sphere::Contains, Sphere.h:37-40: |centre - point| < radius,
the norm sqrt((x*x + y*y) + z*z), Vector3D.cpp:12-15).  Pinned against the
reference's own classes (oracle/_ref, tests/golden/phantoms).

log_scale restates vtkImageLogarithmicScale (VTK 9.0.1, a dependency absent
from the reference tree; the reference calls it with constant 1,
mainReconstruct.cpp:120-124): c*log(1 + x) for x > 0, -c*log(1 - x)
otherwise, in FP64, stored as FP32.  Pinned by closed forms
(tests/test_data_path.py).
"""
from __future__ import annotations

import numpy as np

F = np.float32
EPS = F(np.finfo(np.float32).eps)


def grid_points(n, spacing):
    w = F(F(n) * F(spacing))
    pos = np.empty(n, np.float32)
    v = F(-0.5 * float(w))
    for i in range(n):
        pos[i] = v
        v = F(v + F(spacing))
    return pos


def almost_equal_relative(a, b):
    diff = np.abs(a - b)
    m = np.maximum(np.abs(a), np.abs(b))
    return diff <= EPS * m


def phantom_maker(shape_xyz, spacing_xyz, background=50.0, paves=(), spheres=()):
    """Volume float32[nz][ny][nx]."""
    nx, ny, nz = shape_xyz
    gx = grid_points(nx, spacing_xyz[0])[None, None, :]
    gy = grid_points(ny, spacing_xyz[1])[None, :, None]
    gz = grid_points(nz, spacing_xyz[2])[:, None, None]
    d = np.full((nz, ny, nx), F(background), np.float32)
    for x, y, z, wx, wy, wz, dens in (np.asarray(p, np.float32) for p in paves):
        ex, ey, ez = F(x + wx), F(y + wy), F(z + wz)
        inside = (((gx > x) | almost_equal_relative(gx, x)) & ((gx < ex) | almost_equal_relative(gx, ex)) &
                  ((gy > y) | almost_equal_relative(gy, y)) & ((gy < ey) | almost_equal_relative(gy, ey)) &
                  ((gz > z) | almost_equal_relative(gz, z)) & ((gz < ez) | almost_equal_relative(gz, ez)))
        d = np.where(inside, (d + dens).astype(np.float32), d)
    for cx, cy, cz, r, dens in (np.asarray(s, np.float32) for s in spheres):
        dx, dy, dz = F(cx) - gx, F(cy) - gy, F(cz) - gz
        dist = np.sqrt((dx * dx + dy * dy) + dz * dz).astype(np.float32)
        d = np.where(dist < r, (d + dens).astype(np.float32), d)
    return d.astype(np.float32)


def log_scale(x, c=1.0):
    x = np.asarray(x, np.float64)
    out = np.empty_like(x)
    pos = x > 0.0
    out[pos] = c * np.log(x[pos] + 1.0)
    out[~pos] = -c * np.log(1.0 - x[~pos])  # x = 0 gives -0.0, as the filter
    return out.astype(np.float32)
