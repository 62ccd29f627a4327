"""The reference's DICOM ingest steps on decoded frames (test infrastructure).

Restates KevLoq/Kevernals source/modules/dataHandling/DICOMReader.cpp for
images already decoded to stored values ([frame][row][column], first file
row first), as vtk-dicom's vtkDICOMReader hands them to it with its default
bottom-up memory row order:

* to_vtk: frames -> a VTK image [z][y][x], y = rows - 1 - file row;
* crop: vtkExtractVOI with x in [x0, x0 + w - 1], y in [y0, y0 + h - 1], all
  z (DICOMReader.cpp:61-78), clipped to the image, extent kept;
* order: the acquisition time is read only from files carrying a PatientID
  (DICOMReader.cpp:148-152), the images are sorted by it (:177-179); files of
  equal time keep file-name order (the reference leaves it to
  directory_iterator and std::sort);
* remove: the given indices of that order, largest first (:181-191);
* append: crops stacked along z (vtkImageAppend, :192-233).

Pinned by tests/test_dicom.py: the order of the fixture series is written
out by hand (cases.json), crops are checked against direct index arithmetic
on the known frames, and removal against the reference's erase loop run on
a Python list.
"""
import numpy as np


def to_vtk(frames):
    return np.ascontiguousarray(np.asarray(frames, np.float32)[:, ::-1, :])


def crop(frames, x0, y0, w, h):
    """(image [z][y][x], extent) of one decoded file cropped to the region."""
    img = to_vtk(frames)
    nz, ny, nx = img.shape
    xa, xb = max(x0, 0), min(x0 + w - 1, nx - 1)
    ya, yb = max(y0, 0), min(y0 + h - 1, ny - 1)
    if w <= 0 or h <= 0 or xa > xb or ya > yb:
        raise ValueError("the detector region lies outside the image")
    return img[:, ya:yb + 1, xa:xb + 1].copy(), (xa, xb, ya, yb, 0, nz - 1)


def order(files):
    """files: list of dicts {name, patient, time}; returns the names in the
    order the images are appended."""
    by_name = sorted(files, key=lambda f: f["name"])
    return [f["name"] for f in sorted(by_name, key=lambda f: f["time"] if f["patient"] else "")]


def remove(names, indices):
    out = list(names)
    for i in sorted(indices, reverse=True):
        if i < 0 or i >= len(out):
            raise ValueError("index to remove out of range")
        del out[i]
    return out


def read_directory(decoded, files, indices, roi):
    """decoded: {name: frames}; files as in order(); roi = (x, y, w, h)."""
    names = remove(order(files), indices)
    crops = [crop(decoded[n], *roi) for n in names]
    ext = crops[0][1]
    if any(c[1][:4] != ext[:4] for c in crops):
        raise ValueError("images differ in size")
    data = np.concatenate([c[0] for c in crops], axis=0)
    return data, (ext[0], ext[1], ext[2], ext[3], 0, data.shape[0] - 1)
