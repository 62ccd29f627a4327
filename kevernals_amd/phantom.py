"""Seeded phantoms and transmission data.

Phantoms follow PhantomMaker semantics (KevLoq/Kevernals
source/modules/dataHandling/PhantomMaker.cpp:9-51): a background density
plus the density of every shape that contains the voxel's lower corner
(TomoVolume.cpp:219-230 grid positions), shapes adding up.

Transmission data (I ~ Poisson(I0 * exp(-l)), l the mode-A line integral in
mm, l_hat = -ln(clamp(I, 1, I0) / I0)) is sampled on the device by
kv_poisson_data (Context.poisson_data, learned.transmission_data).
"""
from __future__ import annotations

import numpy as np

from .geometry import Geometry


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(int(seed)))


def _rotation(rng: np.random.Generator) -> np.ndarray:
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array(
        [
            [1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
            [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
            [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)],
        ]
    )


def shapes(g: Geometry, seed: int, n_ellipsoids: int = 20, mu_range=(0.01, 0.05), n_spheres: int = 0,
           sphere_mu: float = 0.1, sphere_diam=(0.3, 1.0)):
    """Random shape list: ellipsoids centred in the inner 80% of the field of
    view with semi-axes in [5%, 25%] of it and a random rotation, plus
    spheres (microcalcifications) of diameter in sphere_diam mm."""
    rng = _rng(seed)
    blf = np.array(g.blf())
    fov = np.array([g.sx * g.nx, g.sy * g.ny, g.sz * g.nz])
    out = []
    for _ in range(n_ellipsoids):
        c = blf + fov * (0.1 + 0.8 * rng.random(3))
        a = fov * rng.uniform(0.05, 0.25, size=3)
        R = _rotation(rng)
        mu = rng.uniform(*mu_range)
        out.append(("ellipsoid", c, a, R, mu))
    for _ in range(n_spheres):
        c = blf + fov * (0.1 + 0.8 * rng.random(3))
        r = 0.5 * rng.uniform(*sphere_diam)
        out.append(("sphere", c, np.array([r, r, r]), np.eye(3), sphere_mu))
    return out


def rasterise(g: Geometry, shape_list, background: float = 0.02) -> np.ndarray:
    """Density volume (nz, ny, nx) float32: background plus every containing
    shape's density, tested at voxel lower corners in float64."""
    xs, ys, zs = (p.astype(np.float64) for p in g.voxel_positions())
    vol = np.full(g.vol_shape, background, dtype=np.float64)
    for kind, c, a, R, mu in shape_list:
        rmax = float(np.max(a))
        lo = c - rmax
        hi = c + rmax
        ix = np.nonzero((xs >= lo[0]) & (xs <= hi[0]))[0]
        iy = np.nonzero((ys >= lo[1]) & (ys <= hi[1]))[0]
        iz = np.nonzero((zs >= lo[2]) & (zs <= hi[2]))[0]
        if len(ix) == 0 or len(iy) == 0 or len(iz) == 0:
            continue
        X = xs[ix][None, None, :] - c[0]
        Y = ys[iy][None, :, None] - c[1]
        for z0 in range(0, len(iz), 8):
            zz = iz[z0 : z0 + 8]
            Z = zs[zz][:, None, None] - c[2]
            if kind == "sphere":
                inside = X * X + Y * Y + Z * Z <= a[0] * a[0]
            else:
                # body coordinates q = R^T (p - c)
                q0 = (R[0, 0] * X + R[1, 0] * Y + R[2, 0] * Z) / a[0]
                q1 = (R[0, 1] * X + R[1, 1] * Y + R[2, 1] * Z) / a[1]
                q2 = (R[0, 2] * X + R[1, 2] * Y + R[2, 2] * Z) / a[2]
                inside = q0 * q0 + q1 * q1 + q2 * q2 <= 1.0
            sub = vol[zz[0] : zz[-1] + 1, iy[0] : iy[-1] + 1, ix[0] : ix[-1] + 1]
            sub += mu * inside
    return vol.astype(np.float32)


def phantom(g: Geometry, seed: int, microcalcifications: int = 0) -> np.ndarray:
    """Seeded phantom: 0.02/mm background, 20 ellipsoids adding U[0.01,0.05]/mm,
    plus optional 0.1/mm microcalcification spheres of 0.3-1 mm."""
    return rasterise(g, shapes(g, seed, n_spheres=microcalcifications))
