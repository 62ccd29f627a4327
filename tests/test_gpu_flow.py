"""The reference's main flow (mainReconstruct.cpp:70-219) end to end on the
GPU: an XML geometry with a detector region, raw projections as a directory
of DICOM files (one per view, acquisition times out of file-name order, two
extra files removed by index), the region cropped on reading, the
log(1 + x) data transform, MLEM, and the result written as TIFF.  Every
stage is compared with the oracle bit for bit.
"""
import importlib.util
import os

import numpy as np
import pytest

from conftest import rng

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

HERE = os.path.dirname(os.path.abspath(__file__))
GX = os.path.join(HERE, "golden", "geometry_xml")


def encoder():
    spec = importlib.util.spec_from_file_location("make_dicom", os.path.join(HERE, "golden", "make_dicom.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def test_dicom_directory_to_tiff_through_mlem(tmp_path, oracle_mod):
    import kevernals_amd as K
    from kevernals_amd import geometry as G
    from kevernals_amd import io as KIO
    from kevernals_amd.phantom import phantom
    from oracle import data_path as DP
    from oracle import tiff as T

    assert torch.cuda.is_available(), "GPU tests need a GPU"
    xml = os.path.join(GX, "roi_sub.xml")
    g = G.from_xml(xml)
    x0, y0, w, h = KIO.roi_from_xml(xml)
    assert (w, h) == (g.nu, g.nv) and (x0, y0) == (17, 9)
    full = (256, 256)  # the detector of roi_sub.xml (pixelWidth / pixelHeight)
    o = oracle_mod.Oracle(g)
    proj = o.fp_r(phantom(g, 12))
    raw = np.clip(np.round(proj * 4000.0), 0, 16383).astype(np.int64)  # raw detector values in the region

    # one DICOM file per view: the full detector, rows top-down in the file
    # (the reader stores them bottom-up, y = rows - 1 - file row)
    enc = encoder()
    r = rng(3)
    dcm = tmp_path / "dcm_data_files"
    dcm.mkdir()
    order = r.permutation(g.nviews + 2)  # file names out of acquisition order
    for k in range(g.nviews + 2):
        frame = r.integers(0, 16384, (full[1], full[0]))
        if 0 < k <= g.nviews:
            frame[y0:y0 + h, x0:x0 + w] = raw[k - 1]
        body = enc.dataset(enc.Enc(), frame[::-1][None], 16, 14, 13, time=f"{100000 + k:06d}")
        (dcm / f"p{order[k]:02d}.dcm").write_bytes(enc.meta(enc.TS_EXPLICIT) + body)
    img = KIO.read_dicom_directory(str(dcm), roi=(x0, y0, w, h), indices_to_remove=[0, g.nviews + 1])
    assert img.extent == (x0, x0 + w - 1, y0, y0 + h - 1, 0, g.nviews - 1)
    np.testing.assert_array_equal(img.data, raw.astype(np.float32))

    # log(1 + x) on the GPU (mainReconstruct.cpp:113-124)
    ctx = K.Context(0, g)
    d_raw = torch.from_numpy(np.ascontiguousarray(img.data)).cuda()
    d_data = torch.empty_like(d_raw)
    ctx.log_scale(d_raw, d_data, 1.0)
    data = DP.log_scale(raw.astype(np.float32))
    np.testing.assert_array_equal(d_data.cpu().numpy(), data)

    # MLEM from the reference's default initial volume (ones), 3 iterations
    x = torch.ones(g.vol_shape, device="cuda")
    for _ in range(3):
        ctx.mlem_iteration(x, d_data, 1.0, K.MODE_R)
    vol = x.cpu().numpy()
    np.testing.assert_array_equal(vol, o.mlem(data, np.ones(g.vol_shape, np.float32), 3, 1.0))

    # the result as the reference's TIFF output
    out = tmp_path / "MLEMReconstructedImage.tiff"
    KIO.write_tiff(str(out), vol, spacing=(g.sx, g.sy))
    assert out.read_bytes() == T.write_vtk_tiff(vol, (g.sx, g.sy))
    pages = T.read_tiff(out.read_bytes())
    np.testing.assert_array_equal(np.stack([p[..., 0] for _, p in pages]), vol)
