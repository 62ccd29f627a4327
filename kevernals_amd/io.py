"""The data formats on either side of the projection path: TIFF output as the
reference's vtkTIFFWriter dumps (Reconstructors.h:49-52 and its other
verbose dumps) and DICOM projection input as its DICOMReader
(DICOMReader.cpp:26-202), through libkevernals_dropin.so (host code).
"""
import ctypes
import os

import numpy as np

_LIB = None


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libkevernals_dropin.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built (run __graft_entry__.build())")
        lib = ctypes.CDLL(path)
        lib.kvd_write_tiff.restype = ctypes.c_int
        lib.kvd_write_tiff.argtypes = [ctypes.c_char_p, ctypes.c_void_p] + [ctypes.c_int] * 5 + \
            [ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
        lib.kvd_dicom_read.restype = ctypes.c_void_p
        lib.kvd_dicom_read.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                       ctypes.c_char_p, ctypes.c_int]
        lib.kvd_image_info.restype = ctypes.c_int
        lib.kvd_image_info.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        lib.kvd_image_data.restype = ctypes.c_void_p
        lib.kvd_image_data.argtypes = [ctypes.c_void_p]
        lib.kvd_image_free.restype = None
        lib.kvd_image_free.argtypes = [ctypes.c_void_p]
        lib.kvd_geometry_roi_from_xml.restype = ctypes.c_int
        lib.kvd_geometry_roi_from_xml.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int]
        _LIB = lib
    return _LIB


_TYPES = {np.dtype(np.uint8): 3, np.dtype(np.uint16): 5, np.dtype(np.float32): 10}


def write_tiff(path, image, spacing=(1.0, 1.0), compression="packbits"):
    """Write image ([z][y][x] or [z][y][x][c], VTK memory order: x fastest,
    y pointing up) as vtkTIFFWriter does: one z slice -> rows from the top
    (highest y) down; several -> one page per slice, rows in memory order."""
    a = np.ascontiguousarray(image)
    if a.ndim == 2:
        a = a[None]
    if a.ndim == 3:
        a = a[..., None]
    if a.ndim != 4 or a.dtype not in _TYPES:
        raise ValueError("image must be [z][y][x][c] of uint8, uint16 or float32")
    if compression not in ("packbits", "none"):
        raise ValueError("compression is 'packbits' or 'none'")
    nz, ny, nx, nc = a.shape
    err = ctypes.create_string_buffer(512)
    rc = _lib().kvd_write_tiff(os.fsencode(path), a.ctypes.data, _TYPES[a.dtype], nc, nx, ny, nz, float(spacing[0]),
                               float(spacing[1]), 1 if compression == "packbits" else 0, err, 512)
    if rc:
        raise ValueError(err.value.decode(errors="replace"))


class Image:
    """A float image with a VTK extent (inclusive) and spacing."""

    def __init__(self, data, extent, spacing):
        self.data, self.extent, self.spacing = data, tuple(extent), tuple(spacing)


class _Owner:
    def __init__(self, handle):
        self.handle = handle

    def __del__(self):
        _lib().kvd_image_free(self.handle)


def _read(path, directory, remove, roi):
    lib = _lib()
    err = ctypes.create_string_buffer(1024)
    rm = np.ascontiguousarray(np.asarray(remove if remove is not None else [], np.int32))
    r = None if roi is None else np.ascontiguousarray(np.asarray(roi, np.int32))
    if r is not None and r.shape != (4,):
        raise ValueError("roi is (x, y, width, height)")
    h = lib.kvd_dicom_read(os.fsencode(path), int(directory), rm.ctypes.data if rm.size else None, int(rm.size),
                           r.ctypes.data if r is not None else None, err, 1024)
    if not h:
        raise ValueError(err.value.decode(errors="replace"))
    owner = _Owner(h)  # frees the image when the last view of it goes
    ext = np.zeros(6, np.int32)
    sp = np.zeros(3, np.float64)
    lib.kvd_image_info(h, ext.ctypes.data, sp.ctypes.data)
    shape = (ext[5] - ext[4] + 1, ext[3] - ext[2] + 1, ext[1] - ext[0] + 1)
    n = int(np.prod(shape))
    buf = (ctypes.c_float * n).from_address(lib.kvd_image_data(h))
    buf._owner = owner
    data = np.frombuffer(buf, np.float32).reshape(shape)
    return Image(data, ext.tolist(), sp.tolist())


def read_dicom(path, roi=None):
    """One DICOM file as DICOMReader::Read: frames along z, rows bottom-up,
    cropped to roi = (x, y, width, height) detector pixels (None: whole)."""
    return _read(path, False, None, roi)


def read_dicom_directory(path, roi=None, indices_to_remove=()):
    """Every .dcm/.DCM file of a directory as DICOMReader::ReadDirectory:
    ordered by acquisition time, the given indices of that order removed,
    each cropped to roi and appended along z."""
    return _read(path, True, list(indices_to_remove), roi)


def roi_from_xml(path):
    """The detector region (x, y, width, height) of an XML geometry file."""
    roi = np.zeros(4, np.int32)
    err = ctypes.create_string_buffer(512)
    if _lib().kvd_geometry_roi_from_xml(os.fsencode(path), roi.ctypes.data, err, 512):
        raise ValueError(err.value.decode(errors="replace"))
    return tuple(int(v) for v in roi)
