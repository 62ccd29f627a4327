"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the
CPU-baseline leg of bench.py, as the checker.  The product path never uses it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "build", "liboracle.so")


class OrGeom(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("sx", ctypes.c_float), ("sy", ctypes.c_float), ("sz", ctypes.c_float),
        ("nu", ctypes.c_int32), ("nv", ctypes.c_int32), ("np", ctypes.c_int32),
        ("pu", ctypes.c_float), ("pv", ctypes.c_float),
        ("det", ctypes.POINTER(ctypes.c_float)), ("src", ctypes.POINTER(ctypes.c_float)),
    ]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        P = ctypes.c_void_p
        I = ctypes.c_int
        G = ctypes.POINTER(OrGeom)
        sig = {
            "or_r_ray_alphas": (I, [G, I, I, I, P, I]),
            "or_r_ray_trace": (I, [G, I, I, I, P, P, I]),
            "or_r_fp": (None, [G, P, P, I]),
            "or_r_bp": (None, [G, P, P, I]),
            "or_r_fp_f64": (None, [G, P, P, I]),
            "or_r_bp_f64": (None, [G, P, P, I]),
            "or_r_fp_rows": (None, [G, P, P, I, I, I, I]),
            "or_r_bp_slices": (None, [G, P, P, I, I, I]),
            "or_r_bp_voxel_trace": (I, [G, I, I, I, P, P, I]),
            "or_r_fp_rays": (None, [G, P, I, P, P, I]),
            "or_r_bp_voxels": (None, [G, P, I, P, P, I]),
            "or_r_fp_nnz": (ctypes.c_int64, [G, I]),
            "or_r_bp_nnz": (ctypes.c_int64, [G, I]),
            "or_r_art": (None, [G, P, P, I, ctypes.c_float, I]),
            "or_r_mlem": (None, [G, P, P, I, ctypes.c_float, I]),
            "or_saa": (I, [I, I, I, P, P, I, P, ctypes.c_float, ctypes.c_float, ctypes.c_float, P]),
            "or_almost_equal_relative": (I, [ctypes.c_float, ctypes.c_float, ctypes.c_float]),
            "or_a_fp": (None, [G, P, P, I]),
            "or_a_bp": (None, [G, P, P, I]),
            "or_a_fp_f64": (None, [G, P, P, I]),
            "or_a_bp_f64": (None, [G, P, P, I]),
            "or_a_fp_brute": (None, [G, P, P, I]),
            "or_a_ray_trace": (I, [G, I, I, I, P, P, I]),
            "or_a_mlem_f64": (None, [G, P, P, I, I]),
            "or_a_loglik_f64": (ctypes.c_double, [G, P, P, I]),
        }
        for k, (r, a) in sig.items():
            f = getattr(L, k)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def threads() -> int:
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


class Oracle:
    """Oracle bound to one geometry (a kevernals_amd.geometry.Geometry)."""

    def __init__(self, g, nthreads: int | None = None):
        self.g = g
        self.det = np.ascontiguousarray(g.det, dtype=np.float32)
        self.src = np.ascontiguousarray(g.src, dtype=np.float32)
        fp = ctypes.POINTER(ctypes.c_float)
        self.og = OrGeom(g.nx, g.ny, g.nz, g.sx, g.sy, g.sz, g.nu, g.nv, g.nviews, g.pu, g.pv,
                         self.det.ctypes.data_as(fp), self.src.ctypes.data_as(fp))
        self.nt = nthreads or threads()
        self.L = lib()

    def _g(self):
        return ctypes.byref(self.og)

    @staticmethod
    def _p(a):
        assert a.flags["C_CONTIGUOUS"]
        return a.ctypes.data

    def _vol(self, dtype=np.float32):
        return np.zeros(self.g.vol_shape, dtype=dtype)

    def _proj(self, dtype=np.float32):
        return np.zeros(self.g.proj_shape, dtype=dtype)

    # mode R
    def ray_alphas(self, p, u, v):
        cap = self.g.nx + self.g.ny + self.g.nz + 3
        a = np.zeros(cap, dtype=np.float32)
        n = self.L.or_r_ray_alphas(self._g(), p, u, v, self._p(a), cap)
        assert n >= 0
        return a[:n]

    def ray_trace(self, p, u, v, mode="R"):
        cap = self.g.nx + self.g.ny + self.g.nz + 3
        idx = np.zeros(cap, dtype=np.int64)
        w = np.zeros(cap, dtype=np.float32)
        f = self.L.or_r_ray_trace if mode == "R" else self.L.or_a_ray_trace
        n = f(self._g(), p, u, v, self._p(idx), self._p(w), cap)
        assert n >= 0
        return idx[:n], w[:n]

    def bp_voxel_trace(self, x, y, z):
        cap = 121 * self.g.nviews
        pix = np.zeros(cap, dtype=np.int64)
        w = np.zeros(cap, dtype=np.float32)
        n = self.L.or_r_bp_voxel_trace(self._g(), x, y, z, self._p(pix), self._p(w), cap)
        assert n >= 0
        return pix[:n], w[:n]

    def fp_r(self, vol):
        vol = np.ascontiguousarray(vol, dtype=np.float32)
        out = self._proj()
        self.L.or_r_fp(self._g(), self._p(vol), self._p(out), self.nt)
        return out

    def bp_r(self, proj):
        proj = np.ascontiguousarray(proj, dtype=np.float32)
        out = self._vol()
        self.L.or_r_bp(self._g(), self._p(proj), self._p(out), self.nt)
        return out

    def fp_r_rays(self, vol, rays):
        vol = np.ascontiguousarray(vol, dtype=np.float32)
        rays = np.ascontiguousarray(rays, dtype=np.int64)
        out = np.zeros(len(rays), dtype=np.float32)
        self.L.or_r_fp_rays(self._g(), self._p(vol), len(rays), self._p(rays), self._p(out), self.nt)
        return out

    def bp_r_voxels(self, proj, vox):
        proj = np.ascontiguousarray(proj, dtype=np.float32)
        vox = np.ascontiguousarray(vox, dtype=np.int64)
        out = np.zeros(len(vox), dtype=np.float32)
        self.L.or_r_bp_voxels(self._g(), self._p(proj), len(vox), self._p(vox), self._p(out), self.nt)
        return out

    def fp_r64(self, vol):
        vol = np.ascontiguousarray(vol, dtype=np.float64)
        out = self._proj(np.float64)
        self.L.or_r_fp_f64(self._g(), self._p(vol), self._p(out), self.nt)
        return out

    def bp_r64(self, proj):
        proj = np.ascontiguousarray(proj, dtype=np.float64)
        out = self._vol(np.float64)
        self.L.or_r_bp_f64(self._g(), self._p(proj), self._p(out), self.nt)
        return out

    def fp_r_rows(self, vol, p, v0, v1):
        vol = np.ascontiguousarray(vol, dtype=np.float32)
        out = self._proj()
        self.L.or_r_fp_rows(self._g(), self._p(vol), self._p(out), p, v0, v1, self.nt)
        return out

    def bp_r_slices(self, proj, z0, z1):
        proj = np.ascontiguousarray(proj, dtype=np.float32)
        out = self._vol()
        self.L.or_r_bp_slices(self._g(), self._p(proj), self._p(out), z0, z1, self.nt)
        return out

    def fp_nnz(self):
        return int(self.L.or_r_fp_nnz(self._g(), self.nt))

    def bp_nnz(self):
        return int(self.L.or_r_bp_nnz(self._g(), self.nt))

    def art(self, data, vol0, iters, relax):
        vol = np.array(vol0, dtype=np.float32, order="C", copy=True)
        data = np.ascontiguousarray(data, dtype=np.float32)
        self.L.or_r_art(self._g(), self._p(data), self._p(vol), iters, relax, self.nt)
        return vol

    def mlem(self, data, vol0, iters, relax):
        vol = np.array(vol0, dtype=np.float32, order="C", copy=True)
        data = np.ascontiguousarray(data, dtype=np.float32)
        self.L.or_r_mlem(self._g(), self._p(data), self._p(vol), iters, relax, self.nt)
        return vol

    # mode A
    def fp_a(self, vol, brute=False):
        vol = np.ascontiguousarray(vol, dtype=np.float32)
        out = self._proj()
        (self.L.or_a_fp_brute if brute else self.L.or_a_fp)(self._g(), self._p(vol), self._p(out), self.nt)
        return out

    def bp_a(self, proj):
        proj = np.ascontiguousarray(proj, dtype=np.float32)
        out = self._vol()
        self.L.or_a_bp(self._g(), self._p(proj), self._p(out), self.nt)
        return out

    def fp_a64(self, vol):
        vol = np.ascontiguousarray(vol, dtype=np.float64)
        out = self._proj(np.float64)
        self.L.or_a_fp_f64(self._g(), self._p(vol), self._p(out), self.nt)
        return out

    def bp_a64(self, proj):
        proj = np.ascontiguousarray(proj, dtype=np.float64)
        out = self._vol(np.float64)
        self.L.or_a_bp_f64(self._g(), self._p(proj), self._p(out), self.nt)
        return out

    def mlem_a64(self, data, vol0, iters):
        vol = np.array(vol0, dtype=np.float64, order="C", copy=True)
        data = np.ascontiguousarray(data, dtype=np.float64)
        self.L.or_a_mlem_f64(self._g(), self._p(data), self._p(vol), iters, self.nt)
        return vol

    def loglik_a64(self, data, vol):
        data = np.ascontiguousarray(data, dtype=np.float64)
        vol = np.ascontiguousarray(vol, dtype=np.float64)
        return float(self.L.or_a_loglik_f64(self._g(), self._p(data), self._p(vol), self.nt))


def saa(proj, src_y, plane_z, sid, fulcrum_z, voxel_dy):
    """recons::ShiftAndAdd restated (Reconstructors.h:263-379)."""
    proj = np.ascontiguousarray(proj, dtype=np.float32)
    np_, nv, nu = proj.shape
    sy = np.ascontiguousarray(src_y, dtype=np.float32)
    pz = np.ascontiguousarray(plane_z, dtype=np.float32)
    out = np.zeros((len(pz), nv, nu), dtype=np.float32)
    rc = lib().or_saa(nu, nv, np_, proj.ctypes.data, sy.ctypes.data, len(pz), pz.ctypes.data, sid, fulcrum_z,
                      voxel_dy, out.ctypes.data)
    return rc, out


def almost_equal_relative(a, b, eps=np.finfo(np.float32).eps):
    return bool(lib().or_almost_equal_relative(a, b, eps))
