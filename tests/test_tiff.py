"""TIFF output (kevernals_amd/cpp/TiffWriter.cpp through kevernals_amd.io)
against the oracle (oracle/tiff.py) and the reference's own file.

The oracle restates vtkTIFFWriter (VTK 9.0.1) over libtiff 4.0.x.  Pins:
* resources/testImage.tiff of the reference (tests/golden/tiff, a 460x360
  RGBA PackBits file of that writer): re-encoding its pixels reproduces it
  byte for byte (directory layout, tag types, out-of-line value order, strip
  sizing, resolution rational, extra-samples value);
* libtiff's PackBits encoder as built into Pillow (libtiff 4.7): identical
  strips on rows built to reach every encoder state (2-byte runs between
  literals, literals and runs past 128 bytes); a plain PackBits encoder, one
  without the 2-byte-run merge and one with a 129-byte literal cap each fail;
* an independent decoder (Pillow) reads every page of the product's files
  back to the input values, rows in vtkTIFFWriter's order.
Parity unpinned: the page tags of the multi-slice path (restated from
vtkTIFFWriter::WriteVolume; the reference ships no multi-page file).
"""
import io as pyio
import os

import numpy as np
import pytest

from kevernals_amd import io as KIO
from oracle import tiff as T

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "tiff", "testImage.tiff")
REF_FILE = "/root/reference/resources/testImage.tiff"

PIL = pytest.importorskip("PIL.Image")


def _fixture():
    data = open(FIXTURE, "rb").read()
    tags, arr = T.read_tiff(data)[0]
    return data, tags, arr


def test_oracle_reproduces_the_reference_file():
    data, tags, arr = _fixture()
    assert np.array_equal(arr, np.asarray(PIL.open(FIXTURE)))
    # VTK memory order has y up: the file's first row is the highest y
    assert T.write_vtk_tiff(arr[::-1][None], spacing=(10 / 71, 10 / 71)) == data


@pytest.mark.skipif(not os.path.exists(REF_FILE), reason="reference tree absent")
def test_fixture_is_the_reference_file():
    assert open(REF_FILE, "rb").read() == open(FIXTURE, "rb").read()


def _stress_rows():
    rng = np.random.default_rng(0)
    rows = [bytes([1, 2, 2, 3, 4, 4, 5, 6, 6, 7] * 20),           # 2-byte runs between literals
            bytes(rng.integers(0, 256, 300, dtype=np.uint8)),        # literals past 128 bytes
            bytes([9] * 300 + [1, 2, 3, 5, 5, 6]),                   # a run past 128 bytes
            bytes(rng.integers(0, 3, 400, dtype=np.uint8)),          # mixed short runs
            bytes([7] * 129 + [8] + [8] * 130 + [1, 1, 2]),
            bytes([3, 3] + list(range(125)) + [4, 4, 5] + [6] * 2 + [7]),
            b"", b"\x05", b"\x05\x05"]
    return rows


def _pillow_strip(row):
    buf = pyio.BytesIO()
    PIL.fromarray(np.frombuffer(row, np.uint8)[None, :]).save(buf, format="TIFF", compression="packbits")
    tags, _ = T.read_tiff(buf.getvalue())[0]
    d = buf.getvalue()
    return b"".join(d[o:o + c] for o, c in zip(tags[273], tags[279]))


def test_packbits_matches_libtiff_and_catches_mistakes():
    from PIL import features
    if not features.check("libtiff"):
        pytest.skip("Pillow without libtiff")
    rows = [r for r in _stress_rows() if r]
    for r in rows:
        enc = T.packbits_encode_row(r)
        assert enc == _pillow_strip(r)
        assert T.packbits_decode(enc, len(r)) == r
    src = open(T.__file__).read()
    for old, new in [("if n == 1 and out[-2] == 0xFF and out[lastlit] < 126:", "if False:"),
                     ("if out[lastlit] == 127:", "if out[lastlit] == 128:")]:
        assert old in src
        ns = {}
        exec(compile(src.replace(old, new), "mutant", "exec"), ns)
        assert any(ns["packbits_encode_row"](r) != _pillow_strip(r) for r in rows), new


def _cases():
    rng = np.random.default_rng(1)
    vol = rng.standard_normal((5, 37, 23)).astype(np.float32)
    vol[:, :, :8] = 0
    return {
        "float_volume": (vol, (0.5, 0.25)),
        "float_slice_wide_rows": (rng.standard_normal((1, 9, 2100)).astype(np.float32), (0.14, 0.14)),
        "u16_volume": (rng.integers(0, 4, (3, 40, 300)).astype(np.uint16), (1.0, 1.0)),
        "u8_rgb_slice": (rng.integers(0, 3, (1, 20, 30, 3)).astype(np.uint8), (2.0, 3.0)),
        "runs_volume": (np.repeat(rng.standard_normal((2, 3, 1)).astype(np.float32), 3000, axis=2), (1, 1)),
        "one_pixel": (np.array([[[1.5]]], np.float32), (1, 1)),
        "projections_4_views": (rng.random((4, 64, 48)).astype(np.float32), (0.14, 0.14)),
    }


@pytest.mark.parametrize("name", sorted(_cases()))
def test_product_matches_oracle_and_decodes(tmp_path, name):
    a, sp = _cases()[name]
    p = str(tmp_path / f"{name}.tiff")
    KIO.write_tiff(p, a, sp)
    data = open(p, "rb").read()
    assert data == T.write_vtk_tiff(a, sp)
    im = PIL.open(p)
    pages = []
    for k in range(getattr(im, "n_frames", 1)):
        im.seek(k)
        pages.append(np.asarray(im))
    got = np.stack(pages).reshape(a.shape)
    want = a if a.shape[0] > 1 else a[:, ::-1]  # one slice: top row first
    assert np.array_equal(got, want)


def test_uncompressed_and_page_tags(tmp_path):
    a = np.arange(2 * 3 * 4, dtype=np.float32).reshape(2, 3, 4)
    p = str(tmp_path / "u.tiff")
    KIO.write_tiff(p, a, (0.5, 0.5), compression="none")
    data = open(p, "rb").read()
    assert data == T.write_vtk_tiff(a, (0.5, 0.5), compression=T.COMPRESSION_NONE)
    pages = T.read_tiff(data)
    assert [t[297] for t, _ in pages] == [[0, 2], [1, 2]]
    assert all(t[254] == [2] and t[339] == [3] and t[282] == [20, 1] and t[296] == [3] for t, _ in pages)
    assert all(274 not in t for t, _ in pages)  # the volume path sets no orientation
    np.testing.assert_array_equal(np.stack([x[..., 0] for _, x in pages]), a)


def test_rejects_what_the_writer_does_not_handle(tmp_path):
    p = str(tmp_path / "x.tiff")
    with pytest.raises(ValueError):
        KIO.write_tiff(p, np.zeros((2, 3, 4, 3), np.float32))  # multi-component volume
    with pytest.raises(ValueError):
        KIO.write_tiff(p, np.zeros((2, 3, 4), np.float64))
    with pytest.raises(ValueError):
        KIO.write_tiff(str(tmp_path / "missing" / "x.tiff"), np.zeros((1, 2, 2), np.float32))
