"""DICOM ingest (kevernals_amd/cpp/DICOMReader.cpp through kevernals_amd.io)
against known pixel values and the reference's reading steps
(oracle/dicom.py).

The fixtures (tests/golden/dicom, written by tests/golden/make_dicom.py from
the standard's encoding rules with pixel values chosen there) cover the
transfer syntaxes the reader accepts (explicit and implicit VR little endian,
deflated, big endian, a bare dataset), 8/16/32-bit cells, BitsStored below
BitsAllocated with junk in the other bits, signed values, sequences of
undefined length before the pixel data, spacing from PixelSpacing or
ImagerPixelSpacing, and the inputs it must refuse.  The reference reads
through vtk-dicom, which is not part of it; no DICOM file ships with the
reference.  Parity unpinned against vtk-dicom itself: the bottom-up row order
(its documented default) and the handling of bits outside BitsStored (taken
from PS3.5 here).
"""
import json
import os
import shutil

import numpy as np
import pytest

from kevernals_amd import io as KIO
from oracle import dicom as D

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "dicom")
CASES = json.load(open(os.path.join(GOLD, "cases.json")))
EXPECTED = np.load(os.path.join(GOLD, "expected.npz"))
FILES = sorted(k for k in CASES if k != "series")


def frames(name):
    return EXPECTED[name.replace("/", "__")]


@pytest.mark.parametrize("name", FILES)
def test_file_reads_to_known_values(name):
    c = CASES[name]
    path = os.path.join(GOLD, name)
    if not c["ok"]:
        with pytest.raises(ValueError) as e:
            KIO.read_dicom(path)
        assert c["why"] in str(e.value) or "cannot read" in str(e.value).lower()
        return
    img = KIO.read_dicom(path)
    f = frames(name)
    want = D.to_vtk(f)
    assert img.data.dtype == np.float32
    np.testing.assert_array_equal(img.data, want)
    assert img.extent == (0, f.shape[2] - 1, 0, f.shape[1] - 1, 0, f.shape[0] - 1)
    assert img.spacing == (c["spacing"][0], c["spacing"][1], 1.0)


def test_rows_are_bottom_up_and_values_are_exact():
    f = frames("explicit_le_u16_12bit.dcm")
    img = KIO.read_dicom(os.path.join(GOLD, "explicit_le_u16_12bit.dcm"))
    # the first row of frame 0 in the file is the highest y of slice 0
    np.testing.assert_array_equal(img.data[0, -1], f[0, 0].astype(np.float32))
    assert f.max() < 4096 and img.data.max() == f.max()  # the junk bits above BitsStored are gone
    s = frames("signed_12bit_high_15.dcm")
    assert s.min() < 0
    np.testing.assert_array_equal(KIO.read_dicom(os.path.join(GOLD, "signed_12bit_high_15.dcm")).data, D.to_vtk(s))


@pytest.mark.parametrize("roi", [(1, 2, 3, 2), (0, 0, 5, 6), (3, 4, 10, 10), (-2, -1, 4, 3)])
def test_crop_matches_index_arithmetic(roi):
    name = "explicit_le_u16_12bit.dcm"
    f = frames(name)
    img = KIO.read_dicom(os.path.join(GOLD, name), roi=roi)
    data, ext = D.crop(f, *roi)
    np.testing.assert_array_equal(img.data, data)
    assert img.extent == ext
    x0, y0, w, h = roi
    rows = f.shape[1]
    for z in range(f.shape[0]):
        for y in range(max(y0, 0), min(y0 + h, rows)):
            for x in range(max(x0, 0), min(x0 + w, f.shape[2])):
                assert img.data[z, y - ext[2], x - ext[0]] == f[z, rows - 1 - y, x]


def test_crop_outside_the_image_is_an_error():
    with pytest.raises(ValueError):
        KIO.read_dicom(os.path.join(GOLD, "explicit_le_u16_12bit.dcm"), roi=(10, 10, 4, 4))


def series_files():
    return [dict(name=n, **v) for n, v in CASES["series"]["files"].items() if n.endswith((".dcm", ".DCM"))]


def test_series_order_is_the_references():
    assert D.order(series_files()) == CASES["series"]["order"]
    # reading the time from every file (not only those with a PatientID)
    # would put c.DCM last
    naive = [f["name"] for f in sorted(sorted(series_files(), key=lambda f: f["name"]), key=lambda f: f["time"])]
    assert naive != CASES["series"]["order"]


@pytest.mark.parametrize("indices", [[], [1], [0, 3], [3, 0, 1]])
@pytest.mark.parametrize("roi", [None, (2, 1, 5, 4)])
def test_directory_matches_oracle(indices, roi):
    decoded = {f["name"]: frames("series/" + f["name"]) for f in series_files()}
    full = (0, 0, 1 << 20, 1 << 20)
    img = KIO.read_dicom_directory(os.path.join(GOLD, "series"), roi=roi, indices_to_remove=indices)
    data, ext = D.read_directory(decoded, series_files(), indices, roi or full)
    np.testing.assert_array_equal(img.data, data)
    assert img.extent == ext


def test_removal_follows_the_erase_loop():
    names = CASES["series"]["order"]
    for idx in ([2], [0, 1], [3, 1]):
        ref = list(names)
        for i in sorted(idx, reverse=True):
            ref.pop(i)
        assert D.remove(names, idx) == ref


def test_directory_errors(tmp_path):
    with pytest.raises(ValueError):
        KIO.read_dicom_directory(os.path.join(GOLD, "series"), indices_to_remove=[4])
    with pytest.raises(ValueError):
        KIO.read_dicom_directory(str(tmp_path / "absent"))
    with pytest.raises(ValueError):
        KIO.read_dicom_directory(str(tmp_path))  # no DICOM files
    bad = tmp_path / "bad"
    shutil.copytree(os.path.join(GOLD, "series"), bad)
    shutil.copy(os.path.join(GOLD, "reject_not_dicom.dcm"), bad / "z.dcm")
    with pytest.raises(ValueError):
        KIO.read_dicom_directory(str(bad))


def test_roi_from_the_geometry_file():
    gx = os.path.join(HERE, "golden", "geometry_xml")
    # <roi>17 9 x x 200 180</roi>: bottom-left (17, 9), top-right (200, 180)
    assert KIO.roi_from_xml(os.path.join(gx, "roi_sub.xml")) == (17, 9, 184, 172)
    assert KIO.roi_from_xml(os.path.join(gx, "small.xml")) == (0, 0, 256, 256)


def test_mutated_files_are_decoded_or_rejected(tmp_path):
    """Truncated and byte-flipped fixtures either decode or raise ValueError
    (the reader never reads past its buffers; the same mutations run clean
    under AddressSanitizer in a host build)."""
    rng = np.random.default_rng(11)
    srcs = [open(os.path.join(GOLD, n), "rb").read() for n in FILES if CASES[n]["ok"]]
    p = str(tmp_path / "m.dcm")
    for it in range(1500):
        b = bytearray(srcs[it % len(srcs)])
        if it % 3 == 0:
            b = b[:int(rng.integers(0, len(b)))]
        else:
            for _ in range(int(rng.integers(1, 8))):
                b[int(rng.integers(0, len(b)))] = int(rng.integers(0, 256))
        open(p, "wb").write(bytes(b))
        try:
            img = KIO.read_dicom(p)
            assert img.data.ndim == 3
        except ValueError:
            pass
