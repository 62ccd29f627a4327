"""Python binding of libkevernals_amd (the C ABI in include/kevernals_amd.h).

This is the host-side mirror used by tests and benchmarks; the product is the
HIP library.  Device buffers are passed as torch tensors (device memory and
streams are plumbing here) or raw device addresses.  There is no CPU
fallback: without the built library, or without a GPU, calls raise.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

from .geometry import Geometry  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libkevernals_amd.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "kevernals_amd.h")

MODE_R = 0
MODE_A = 1

KV_OK = 0
STATUS_NAMES = {
    0: "KV_OK",
    1: "KV_ERR_INVALID",
    2: "KV_ERR_DEVICE",
    3: "KV_ERR_NOMEM",
    4: "KV_ERR_STATE",
    5: "KV_ERR_PRECONDITION",
}


class KvError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class _KvGeometry(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("sx", ctypes.c_float), ("sy", ctypes.c_float), ("sz", ctypes.c_float),
        ("nu", ctypes.c_int32), ("nv", ctypes.c_int32), ("np", ctypes.c_int32),
        ("pu", ctypes.c_float), ("pv", ctypes.c_float),
        ("det", ctypes.POINTER(ctypes.c_float)), ("src", ctypes.POINTER(ctypes.c_float)),
    ]


_lib = None

_P = ctypes.c_void_p
_I = ctypes.c_int
_F = ctypes.c_float
_SIGS = {
    "kv_create": (_I, [_I, ctypes.POINTER(_P)]),
    "kv_destroy": (_I, [_P]),
    "kv_set_stream": (_I, [_P, _P]),
    "kv_get_stream": (_P, [_P]),
    "kv_synchronize": (_I, [_P]),
    "kv_last_error": (ctypes.c_char_p, []),
    "kv_version": (ctypes.c_char_p, []),
    "kv_set_geometry": (_I, [_P, ctypes.POINTER(_KvGeometry)]),
    "kv_malloc": (_I, [_P, ctypes.c_size_t, ctypes.POINTER(_P)]),
    "kv_free": (_I, [_P, _P]),
    "kv_memcpy_h2d": (_I, [_P, _P, _P, ctypes.c_size_t]),
    "kv_memcpy_d2h": (_I, [_P, _P, _P, ctypes.c_size_t]),
    "kv_memset": (_I, [_P, _P, _I, ctypes.c_size_t]),
    "kv_forward_project": (_I, [_P, _P, _P, _I]),
    "kv_back_project": (_I, [_P, _P, _P, _I]),
    "kv_mlem_iteration": (_I, [_P, _P, _P, _F, _I]),
    "kv_mlem_forward": (_I, [_P, _P, _P, _F, _I]),
    "kv_mlem_backward": (_I, [_P, _P, _I]),
    "kv_count_pairs": (_I, [_P, _I, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]),
    "kv_sart_iteration": (_I, [_P, _P, _P, _F]),
    "kv_shift_and_add": (_I, [_P, _P, _P, _I, _P, _F, _F, _F, _P]),
    "kv_shift_and_add_noalloc": (_I, [_P, _P, _P, _I, _P, _F, _F, _F, _P]),
    "kv_log_likelihood": (_I, [_P, _P, _P, ctypes.POINTER(ctypes.c_double)]),
    "kv_log_likelihoods": (_I, [_P, _P, _P, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]),
    "kv_sensitivity": (_I, [_P, _P]),
    "kv_trace_rays": (_I, [_P, _I, _I, _P, _I, _P, _P, _P]),
    "kv_forward_project_host": (_I, [_P, _P, _P, _I]),
    "kv_back_project_host": (_I, [_P, _P, _P, _I]),
    "kv_batch_create": (_I, [_I, _I, _P, _I, ctypes.POINTER(_P)]),
    "kv_batch_context": (_I, [_P, _I, ctypes.POINTER(_P)]),
    "kv_batch_buffers": (_I, [_P, _I, ctypes.POINTER(_P), ctypes.POINTER(_P)]),
    "kv_memcpy_d2d": (_I, [_P, _P, _P, ctypes.c_size_t]),
    "kv_batch_destroy": (_I, [_P]),
    "kv_batch_set_volume": (_I, [_P, _I, _P]),
    "kv_batch_set_data": (_I, [_P, _I, _P]),
    "kv_batch_get_volume": (_I, [_P, _I, _P]),
    "kv_batch_mlem": (_I, [_P, _P, _F, _I, ctypes.POINTER(_F)]),
    "kv_batch_set_tiles": (_I, [_P, _I, _I, _I]),
    "kv_batch_get_tiles": (_I, [_P, _I, ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    "kv_batch_autotune": (_I, [_P, ctypes.c_char_p]),
    "kv_cnn_workspace_size": (_I, [_I, _I, _I, _I, ctypes.POINTER(ctypes.c_size_t)]),
    "kv_cnn_workspace_offsets": (_I, [_I, _I, _I, ctypes.POINTER(ctypes.c_size_t)]),
    "kv_cnn_forward": (_I, [_P, _P, _F, _P, _P, _I, _I, _I, _P, ctypes.c_size_t]),
    "kv_cnn_gradient": (_I, [_P, _P, _F, _P, _P, _I, _I, _I, _P, ctypes.c_size_t, _P,
                             ctypes.POINTER(ctypes.c_double)]),
    "kv_cnn_adam": (_I, [_P, _P, _P, _P, _P, _I, _F]),
    "kv_phantom": (_I, [_P, _F, _P, _I, _P, _I, _P]),
    "kv_log_scale": (_I, [_P, _P, _P, ctypes.c_int64, ctypes.c_double]),
    "kv_traj_gram": (_I, [_P, _P, _I, _P]),
    "kv_poisson_data": (_I, [_P, _P, ctypes.c_int64, _F, ctypes.c_uint64, _P]),
    "kv_traj_combine": (_I, [_P, _P, _I, _P, _P]),
    "kv_ritz_coefficients": (_I, [_P, _I, _P, _I, _I, _I, ctypes.c_double, ctypes.c_double, _P, _P,
                                  ctypes.POINTER(_I)]),
    "kv_mlem_accelerated": (_I, [_P, _P, _P, _P, _P, _F, ctypes.POINTER(_I)]),
    "kv_batch_mlem_accelerated": (_I, [_P, _P, _P, _F, _P]),
}


def load_library(path: Optional[str] = None) -> ctypes.CDLL:
    """Load libkevernals_amd.so (built by __graft_entry__.build()).  Raises
    if it is missing: there is no fallback implementation."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"libkevernals_amd.so not built at {p}; run __graft_entry__.build()")
    lib = ctypes.CDLL(p)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def declared_symbols(header: str = HEADER_PATH):
    """Entry points declared in the public header."""
    import re

    text = open(header).read()
    return sorted(set(re.findall(r"^\s*(?:kv_status|void|const char)\s*\*?\s*(kv_\w+)\s*\(", text, re.M)))


def _ptr(x) -> int:
    """Device address of a torch tensor, or an int address."""
    if x is None:
        return 0
    if isinstance(x, int):
        return x
    if hasattr(x, "data_ptr"):
        if not x.is_contiguous():
            raise ValueError("tensor must be contiguous")
        return x.data_ptr()
    raise TypeError(f"cannot take a device address of {type(x)}")


def _nbytes(x) -> int:
    return x.numel() * x.element_size()


CNN_NPARAMS = 14737
LL_MAX = 3  # volumes per log_likelihoods call
TRAJ_MAX = 32


def cnn_workspace_size(shape, training: bool) -> int:
    """Bytes of CNN workspace for a volume of shape (nz, ny, nx)."""
    lib = load_library()
    nz, ny, nx = shape
    n = ctypes.c_size_t()
    st = lib.kv_cnn_workspace_size(nx, ny, nz, int(training), ctypes.byref(n))
    if st != KV_OK:
        raise KvError(st, lib.kv_last_error().decode())
    return n.value


def cnn_workspace_offsets(shape):
    """Byte offsets of a1, a2, a3, dz4, dz3, dz2, dz1 in a training workspace."""
    lib = load_library()
    nz, ny, nx = shape
    o = (ctypes.c_size_t * 7)()
    st = lib.kv_cnn_workspace_offsets(nx, ny, nz, o)
    if st != KV_OK:
        raise KvError(st, lib.kv_last_error().decode())
    return list(o)


class _KvAccelSchedule(ctypes.Structure):
    _fields_ = [("start", ctypes.c_int32), ("first", ctypes.c_int32), ("mid", ctypes.c_int32),
                ("second", ctypes.c_int32), ("horizon", ctypes.c_int32)]


def _schedule_struct(schedule) -> _KvAccelSchedule:
    if schedule is None:
        return _KvAccelSchedule(2, 10, 40, 8, 100)
    return _KvAccelSchedule(schedule.start, schedule.first, schedule.mid, schedule.second, schedule.horizon)


def ritz_coefficients_native(gram: np.ndarray, pairs, target: int, horizon: int, rel_tol: float,
                             ritz_max: float):
    """kv_ritz_coefficients: (coefficients over the h log-ratios, Ritz values)."""
    lib = load_library()
    G = np.ascontiguousarray(gram, dtype=np.float64)
    h = G.shape[0]
    pr = np.ascontiguousarray(np.asarray(pairs, dtype=np.int32).reshape(-1, 2))
    coef = np.zeros(h)
    ritz = np.zeros(len(pr))
    nr = _I()
    st = lib.kv_ritz_coefficients(_host_ptr(G), h, _host_ptr(pr), len(pr), int(target), int(horizon), rel_tol,
                                  ritz_max, _host_ptr(coef), _host_ptr(ritz), ctypes.byref(nr))
    if st != KV_OK:
        raise KvError(st, lib.kv_last_error().decode())
    return coef, ritz[:nr.value].copy()


def _host_ptr(a: np.ndarray) -> int:
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return a.ctypes.data


class Context:
    """One kv_context: a device, a stream, a geometry and its tables.

    With follow_torch_stream (the default) every device operation is
    enqueued on torch's current stream, so buffers from torch's caching
    allocator (including temporaries released right after the call) are
    reused only after the kernels that read them."""

    def __init__(self, device: int = 0, geometry: Optional[Geometry] = None, follow_torch_stream: bool = True):
        self.lib = load_library()
        h = _P()
        self._check(self.lib.kv_create(device, ctypes.byref(h)))
        self.handle = h
        self._owned = True
        self.geometry = None
        self.follow_torch_stream = follow_torch_stream
        self._stream = None
        if geometry is not None:
            self.set_geometry(geometry)

    @classmethod
    def borrow(cls, handle, geometry: Geometry, owner=None) -> "Context":
        """Wrap a context owned by someone else (a batch's reconstruction):
        it is never destroyed from here; `owner` is kept alive."""
        self = cls.__new__(cls)
        self.lib = load_library()
        self.handle = handle
        self._owned = False
        self._owner = owner
        self.geometry = geometry
        self.follow_torch_stream = True
        self._stream = None
        return self

    def _on_torch_stream(self):
        if not self.follow_torch_stream:
            return
        import torch

        s = torch.cuda.current_stream().cuda_stream
        if s != self._stream:
            self._check(self.lib.kv_set_stream(self.handle, s))
            self._stream = s

    def _check(self, status: int):
        if status != KV_OK:
            raise KvError(status, self.lib.kv_last_error().decode())

    def close(self):
        if getattr(self, "handle", None):
            if getattr(self, "_owned", True):
                self.lib.kv_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- plumbing ---------------------------------------------------------
    def copy(self, dst, src, nbytes: int):
        """Device-to-device copy on the context's stream."""
        self._dev(self.lib.kv_memcpy_d2d, _ptr(dst), _ptr(src), nbytes)
    def set_stream(self, stream_handle: Optional[int]):
        """Enqueue on an explicit hipStream_t (0 = null stream) from now on;
        stops following torch's current stream."""
        self.follow_torch_stream = False
        self._check(self.lib.kv_set_stream(self.handle, stream_handle or 0))
        self._stream = stream_handle or 0

    def use_torch_stream(self):
        """Follow torch's current stream (the default)."""
        self.follow_torch_stream = True
        self._on_torch_stream()

    def synchronize(self):
        self._check(self.lib.kv_synchronize(self.handle))

    def set_geometry(self, g: Geometry):
        det = np.ascontiguousarray(g.det, dtype=np.float32)
        src = np.ascontiguousarray(g.src, dtype=np.float32)
        kg = _KvGeometry(g.nx, g.ny, g.nz, g.sx, g.sy, g.sz, g.nu, g.nv, g.nviews, g.pu, g.pv,
                         det.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                         src.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
        self._check(self.lib.kv_set_geometry(self.handle, ctypes.byref(kg)))
        self.geometry = g

    def _dev(self, fn, *args):
        self._on_torch_stream()
        self._check(fn(self.handle, *args))

    # -- operators (device buffers) ---------------------------------------
    def forward_project(self, vol, proj, mode: int = MODE_R):
        self._dev(self.lib.kv_forward_project, _ptr(vol), _ptr(proj), mode)

    def back_project(self, proj, vol, mode: int = MODE_R):
        self._dev(self.lib.kv_back_project, _ptr(proj), _ptr(vol), mode)

    def mlem_iteration(self, vol, data, relax: float = 1.0, mode: int = MODE_R):
        self._dev(self.lib.kv_mlem_iteration, _ptr(vol), _ptr(data), relax, mode)

    def mlem_forward(self, vol, data, relax: float = 1.0, mode: int = MODE_R):
        self._dev(self.lib.kv_mlem_forward, _ptr(vol), _ptr(data), relax, mode)

    def mlem_backward(self, vol, mode: int = MODE_R):
        self._dev(self.lib.kv_mlem_backward, _ptr(vol), mode)

    def count_pairs(self, mode: int = MODE_R):
        """(forward pairs, backprojector pairs) accepted by the two passes."""
        a, b = ctypes.c_int64(), ctypes.c_int64()
        self._dev(self.lib.kv_count_pairs, mode, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def sart_iteration(self, vol, data, relax: float = 0.5):
        self._dev(self.lib.kv_sart_iteration, _ptr(vol), _ptr(data), relax)

    def sensitivity(self, vol):
        self._dev(self.lib.kv_sensitivity, _ptr(vol))

    def shift_and_add(self, proj, out, src_y, plane_z, sid: float, fulcrum_z: float, voxel_dy: float,
                      noalloc: bool = False):
        """kv_shift_and_add, or with noalloc=True kv_shift_and_add_noalloc
        (no device allocation; staging limits KV_SAA_NOALLOC_MAX_*)."""
        sy = np.ascontiguousarray(src_y, dtype=np.float32)
        pz = np.ascontiguousarray(plane_z, dtype=np.float32)
        fn = self.lib.kv_shift_and_add_noalloc if noalloc else self.lib.kv_shift_and_add
        self._dev(fn, _ptr(proj), _host_ptr(sy), len(pz), _host_ptr(pz), sid, fulcrum_z, voxel_dy, _ptr(out))

    def log_likelihood(self, data, vol) -> float:
        out = ctypes.c_double()
        self._dev(self.lib.kv_log_likelihood, _ptr(data), _ptr(vol), ctypes.byref(out))
        return out.value

    def log_likelihoods(self, data, vols) -> list:
        """log_likelihood of up to LL_MAX volumes with one ray walk (each
        value bit-identical to log_likelihood of that volume)."""
        n = len(vols)
        ptrs = (ctypes.c_void_p * n)(*[_ptr(v) for v in vols])
        out = (ctypes.c_double * n)()
        self._dev(self.lib.kv_log_likelihoods, _ptr(data), ptrs, n, out)
        return list(out)

    def trace_rays(self, rays, cap: int = 1024, mode: int = MODE_R):
        rays = np.ascontiguousarray(rays, dtype=np.int32)
        n = len(rays)
        idx = np.zeros((n, cap), dtype=np.int64)
        w = np.zeros((n, cap), dtype=np.float32)
        counts = np.zeros(n, dtype=np.int32)
        self._check(self.lib.kv_trace_rays(self.handle, mode, n, _host_ptr(rays), cap, _host_ptr(idx),
                                           _host_ptr(w), _host_ptr(counts)))
        return [(idx[i, : min(counts[i], cap)], w[i, : min(counts[i], cap)], int(counts[i])) for i in range(n)]

    # -- host-buffer variants --------------------------------------------
    # -- learned step (3D residual CNN) --------------------------------------
    def cnn_forward(self, params, scale: float, x, y, work):
        """y = x + f(scale*x)/scale for a volume x of shape (nz, ny, nx)."""
        nz, ny, nx = x.shape
        self._dev(self.lib.kv_cnn_forward, _ptr(params), scale, _ptr(x), _ptr(y), nx, ny, nz, _ptr(work),
                  _nbytes(work))

    def cnn_gradient(self, params, scale: float, x, target, work, grad) -> float:
        """Gradient of mean((scale*(y - target))^2) into grad; returns the loss."""
        nz, ny, nx = x.shape
        loss = ctypes.c_double()
        self._dev(self.lib.kv_cnn_gradient, _ptr(params), scale, _ptr(x), _ptr(target), nx, ny, nz, _ptr(work),
                  _nbytes(work), _ptr(grad), ctypes.byref(loss))
        return loss.value

    def cnn_adam(self, params, grad, m, v, step: int, lr: float = 1e-3):
        self._dev(self.lib.kv_cnn_adam, _ptr(params), _ptr(grad), _ptr(m), _ptr(v), step, lr)

    def phantom(self, vol, background: float = 50.0, paves=(), spheres=()):
        """PhantomMaker::GetPhantom on this geometry's voxel grid into vol:
        paves are (x, y, z, wx, wy, wz, density) boxes (bottom-left-front
        corner and size, mm), spheres (cx, cy, cz, radius, density).
        Synchronous."""
        pv = np.ascontiguousarray(np.asarray(paves, dtype=np.float32).reshape(-1, 7))
        sp = np.ascontiguousarray(np.asarray(spheres, dtype=np.float32).reshape(-1, 5))
        self._dev(self.lib.kv_phantom, float(background), _host_ptr(pv) if len(pv) else None, len(pv),
                  _host_ptr(sp) if len(sp) else None, len(sp), _ptr(vol))

    def poisson_data(self, ell, out, i0: float, seed: int):
        """out = -ln(clamp(I, 1, i0)/i0), I ~ Poisson(i0 exp(-ell)) per ray,
        from the library's seeded Philox4x64-10 / PTRS sampler."""
        self._dev(self.lib.kv_poisson_data, _ptr(ell), int(ell.numel()), float(i0), int(seed) & (2**64 - 1),
                  _ptr(out))

    def log_scale(self, src, dst, c: float = 1.0):
        """dst = c*log(1 + src) (sign-symmetric), the reference's data
        transform (vtkImageLogarithmicScale, constant c)."""
        self._dev(self.lib.kv_log_scale, _ptr(src), _ptr(dst), int(src.numel()), float(c))

    def traj_gram(self, iterates) -> np.ndarray:
        """Gram matrix (h x h, FP64) of the log-ratios of h+1 consecutive
        mode-A MLEM iterates (device volumes, oldest first) under the weight
        sens * newest iterate."""
        h = len(iterates) - 1
        ptrs = (ctypes.c_void_p * (h + 1))(*[_ptr(x) for x in iterates])
        out = np.zeros((max(h, 0), max(h, 0)), dtype=np.float64)
        self._dev(self.lib.kv_traj_gram, ctypes.cast(ptrs, ctypes.c_void_p), h, _host_ptr(out))
        return out

    def traj_combine(self, iterates, coef, out):
        """out = newest iterate * exp(sum_i coef[i] * log(x_{i+1}/x_i))."""
        h = len(iterates) - 1
        ptrs = (ctypes.c_void_p * (h + 1))(*[_ptr(x) for x in iterates])
        c = np.ascontiguousarray(coef, dtype=np.float64)
        if c.size != h:
            raise ValueError(f"{h} log-ratios need {h} coefficients, got {c.size}")
        self._dev(self.lib.kv_traj_combine, ctypes.cast(ptrs, ctypes.c_void_p), h, _host_ptr(c), _ptr(out))

    def mlem_accelerated(self, vol, data, schedule=None, params=None, scale: float = 20.0) -> int:
        """kv_mlem_accelerated: the two-stretch jump schedule (default
        {2, 10, 40, 8, 100}) on vol (initial volume in, result out) with the
        likelihood-gated CNN step when params (a device tensor of
        CNN_NPARAMS floats) is given.  Returns the steps kept (0, 1, 2)."""
        sch = _schedule_struct(schedule)
        kept = _I()
        self._dev(self.lib.kv_mlem_accelerated, _ptr(vol), _ptr(data), ctypes.byref(sch),
                  _ptr(params) if params is not None else None, scale, ctypes.byref(kept))
        return kept.value

    def forward_project_host(self, vol: np.ndarray, mode: int = MODE_R) -> np.ndarray:
        g = self.geometry
        vol = np.ascontiguousarray(vol, dtype=np.float32)
        out = np.empty(g.proj_shape, dtype=np.float32)
        self._check(self.lib.kv_forward_project_host(self.handle, _host_ptr(vol), _host_ptr(out), mode))
        return out

    def back_project_host(self, proj: np.ndarray, mode: int = MODE_R) -> np.ndarray:
        g = self.geometry
        proj = np.ascontiguousarray(proj, dtype=np.float32)
        out = np.empty(g.vol_shape, dtype=np.float32)
        self._check(self.lib.kv_back_project_host(self.handle, _host_ptr(proj), _host_ptr(out), mode))
        return out


def _geometry_struct(g: Geometry, keep: list) -> _KvGeometry:
    det = np.ascontiguousarray(g.det, dtype=np.float32)
    src = np.ascontiguousarray(g.src, dtype=np.float32)
    keep += [det, src]
    return _KvGeometry(g.nx, g.ny, g.nz, g.sx, g.sy, g.sz, g.nu, g.nv, g.nviews, g.pu, g.pv,
                       det.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                       src.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))


class Batch:
    """A kv_batch: many mode-R MLEM reconstructions of uneven sizes advanced
    together by persistent forward/backprojection launches, optionally
    replayed from a captured hipGraph.  Host arrays in and out; every call is
    synchronous."""

    def __init__(self, geometries, device: int = 0, mode: int = MODE_R):
        self.lib = load_library()
        self.geometries = list(geometries)
        self.mode = mode
        keep: list = []
        arr = (_KvGeometry * len(self.geometries))(*[_geometry_struct(g, keep) for g in self.geometries])
        h = _P()
        self._check(self.lib.kv_batch_create(device, len(self.geometries), ctypes.cast(arr, _P), mode,
                                             ctypes.byref(h)))
        self.handle = h

    def context(self, i: int) -> "Context":
        """Reconstruction i's context (owned by the batch)."""
        h = _P()
        self._check(self.lib.kv_batch_context(self.handle, i, ctypes.byref(h)))
        return Context.borrow(h, self.geometries[i], owner=self)

    def buffers(self, i: int):
        """Device addresses (volume, data) of reconstruction i."""
        v, d = _P(), _P()
        self._check(self.lib.kv_batch_buffers(self.handle, i, ctypes.byref(v), ctypes.byref(d)))
        return v.value, d.value

    def _check(self, status: int):
        if status != KV_OK:
            raise KvError(status, self.lib.kv_last_error().decode())

    def close(self):
        if getattr(self, "handle", None):
            self.lib.kv_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return len(self.geometries)

    def set_volume(self, i: int, vol: np.ndarray):
        vol = np.ascontiguousarray(vol, dtype=np.float32)
        if vol.shape != tuple(self.geometries[i].vol_shape):
            raise ValueError(f"volume {i}: shape {vol.shape} != {self.geometries[i].vol_shape}")
        self._check(self.lib.kv_batch_set_volume(self.handle, i, _host_ptr(vol)))

    def set_data(self, i: int, data: np.ndarray):
        data = np.ascontiguousarray(data, dtype=np.float32)
        if data.shape != tuple(self.geometries[i].proj_shape):
            raise ValueError(f"data {i}: shape {data.shape} != {self.geometries[i].proj_shape}")
        self._check(self.lib.kv_batch_set_data(self.handle, i, _host_ptr(data)))

    def get_volume(self, i: int) -> np.ndarray:
        out = np.empty(self.geometries[i].vol_shape, dtype=np.float32)
        self._check(self.lib.kv_batch_get_volume(self.handle, i, _host_ptr(out)))
        return out

    def mlem(self, iters, relax: float = 1.0, use_graph: bool = True) -> float:
        """Run iters[i] MLEM iterations on reconstruction i (an int applies to
        all); returns the device time in ms."""
        if isinstance(iters, int):
            iters = [iters] * len(self)
        it = np.ascontiguousarray(iters, dtype=np.int32)
        if it.shape != (len(self),):
            raise ValueError("one iteration count per reconstruction")
        ms = _F()
        self._check(self.lib.kv_batch_mlem(self.handle, _host_ptr(it), relax, int(bool(use_graph)), ctypes.byref(ms)))
        return ms.value

    def mlem_accelerated(self, schedule=None, params=None, scale: float = 20.0):
        """kv_batch_mlem_accelerated on every reconstruction (mode-A batch);
        returns the steps kept per reconstruction."""
        sch = _schedule_struct(schedule)
        kept = np.zeros(len(self), dtype=np.int32)
        self._check(self.lib.kv_batch_mlem_accelerated(self.handle, ctypes.byref(sch),
                                                        _ptr(params) if params is not None else None, scale,
                                                        _host_ptr(kept)))
        return kept.tolist()

    def set_tiles(self, i: int, fp_layout: int, bp_slices: int):
        self._check(self.lib.kv_batch_set_tiles(self.handle, i, fp_layout, bp_slices))

    def get_tiles(self, i: int):
        a, b = _I(), _I()
        self._check(self.lib.kv_batch_get_tiles(self.handle, i, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def autotune(self, cache_path: Optional[str] = None):
        self._check(self.lib.kv_batch_autotune(self.handle, cache_path.encode() if cache_path else None))
        return [self.get_tiles(i) for i in range(len(self))]
