"""The C++ drop-in (Projector / recons:: over the C ABI) driven as the
reference's callers drive it, against the oracle on the geometry the drop-in
marshals: source x and z averaged over views in FP32 (TomoTable.cpp:11-18),
common detector height (TomoGeometry.cpp:345), per-view source y and detector
corner (Projector.cu:917-936)."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHECK = os.path.join(ROOT, "kevernals_amd", "lib", "kv_dropin_check")


def marshalled(g):
    """The kernel geometry the drop-in builds from TomoGeometry accessors."""
    from kevernals_amd.geometry import Geometry

    sx = np.float32(0.0)
    sz = np.float32(0.0)
    for p in range(g.nviews):
        sx = np.float32(sx + g.src[p, 0])
        sz = np.float32(sz + g.src[p, 2])
    sx = np.float32(sx / np.float32(g.nviews))
    sz = np.float32(sz / np.float32(g.nviews))
    dz = np.float32(np.float32(sum(float(v) for v in g.det[:, 2])) / np.float32(g.nviews))
    src = np.stack([np.full(g.nviews, sx), g.src[:, 1], np.full(g.nviews, sz)], axis=1).astype(np.float32)
    det = np.stack([g.det[:, 0], g.det[:, 1], np.full(g.nviews, dz)], axis=1).astype(np.float32)
    return Geometry(g.nx, g.ny, g.nz, g.sx, g.sy, g.sz, g.nu, g.nv, g.nviews, g.pu, g.pv, det, src), sz, dz


def check_dumps(tmp_path, gm, o, vol, proj):
    """Verbose ART/MLEM (steps run apart, TIFF dumps of each): the same final
    volumes as the fused iterations, and every dumped image equal to the
    oracle's value at that step (pages decoded by the TIFF oracle's reader)."""
    from oracle import tiff as T

    def rd(name, shape):
        return np.fromfile(tmp_path / name, dtype=np.float32).reshape(shape)

    def tif(path):
        pages = T.read_tiff(open(path, "rb").read())
        return np.stack([p[..., 0] for _, p in pages])

    np.testing.assert_array_equal(rd("mlem_verbose.bin", gm.vol_shape), rd("mlem.bin", gm.vol_shape))
    np.testing.assert_array_equal(rd("art_verbose.bin", gm.vol_shape), rd("art.bin", gm.vol_shape))
    for which, iters, relax in (("mlem", 3, 1.0), ("art", 2, 0.5)):
        d = tmp_path / f"dump_{which}"
        np.testing.assert_array_equal(tif(d / "initialVolume.tiff"), vol)
        run = o.mlem if which == "mlem" else o.art
        x = vol
        for it in range(iters):
            fp = o.fp_r(x)
            np.testing.assert_array_equal(tif(d / f"currentProjectionStep{it}.tiff"), fp)
            r = np.float32(relax)
            if which == "mlem":
                corr = np.where(fp == 0, fp, r * (proj / np.where(fp == 0, 1, fp))).astype(np.float32)
            else:
                corr = (r * (proj - fp)).astype(np.float32)
            np.testing.assert_array_equal(tif(d / f"currentCorrectedProjectionStep{it}.tiff"), corr)
            x = run(proj, vol, it + 1, relax)
            np.testing.assert_array_equal(tif(d / f"reconstructionstep{it}.tiff"), x)


def test_dropin_matches_oracle(tmp_path, oracle_mod):
    from kevernals_amd import geometry as G
    from kevernals_amd.phantom import phantom

    assert os.path.exists(CHECK), "drop-in not built"
    g = G.scaled_down(G.small_plain(), 2)
    gm, sz, dz = marshalled(g)
    o = oracle_mod.Oracle(gm)
    vol = phantom(gm, 5)
    proj = o.fp_r(phantom(gm, 6))
    fz = 0.0
    with open(tmp_path / "geom.txt", "w") as f:
        f.write(f"{g.nx} {g.ny} {g.nz} {g.sx!r} {g.sy!r} {g.sz!r} {g.nu} {g.nv} {g.nviews} {g.pu!r} {g.pv!r} {fz}\n")
        for p in range(g.nviews):
            vals = [*g.src[p], *g.det[p]]
            f.write(" ".join(f"{float(v):.9g}" for v in vals) + "\n")
    vol.tofile(tmp_path / "vol.bin")
    proj.tofile(tmp_path / "proj.bin")
    r = subprocess.run([CHECK, str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    status = dict(line.split() for line in open(tmp_path / "status.txt"))
    assert all(v == "0" for k, v in status.items() if k != "saa_one_slice")
    assert status["saa_one_slice"] == "5"  # ReconstructorsErrorCode::ShiftAndAdd

    def rd(name, shape):
        return np.fromfile(tmp_path / name, dtype=np.float32).reshape(shape)

    np.testing.assert_array_equal(rd("fp.bin", gm.proj_shape), o.fp_r(vol))
    np.testing.assert_array_equal(rd("bp.bin", gm.vol_shape), o.bp_r(proj))
    np.testing.assert_array_equal(rd("mlem.bin", gm.vol_shape), o.mlem(proj, vol, 3, 1.0))
    np.testing.assert_array_equal(rd("mlem_default.bin", gm.vol_shape),
                                  o.mlem(proj, np.ones(gm.vol_shape, np.float32), 2, 1.0))
    np.testing.assert_array_equal(rd("art.bin", gm.vol_shape), o.art(proj, vol, 2, 0.5))
    check_dumps(tmp_path, gm, o, vol, proj)
    # single-source projection: every view from view 0's source position
    g1 = G.Geometry(gm.nx, gm.ny, gm.nz, gm.sx, gm.sy, gm.sz, gm.nu, gm.nv, gm.nviews, gm.pu, gm.pv, gm.det,
                    np.repeat(gm.src[:1], gm.nviews, axis=0))
    np.testing.assert_array_equal(rd("single.bin", gm.proj_shape), oracle_mod.Oracle(g1).fp_r(vol))
    planes = gm.plane_heights()
    sid = np.float32(abs(np.float32(dz) - np.float32(sz)))
    rc, saa = oracle_mod.saa(proj, gm.src[:, 1], planes, float(sid), fz, gm.sy)
    assert rc == 0
    np.testing.assert_array_equal(rd("saa.bin", (len(planes), gm.nv, gm.nu)), saa)


def test_dropin_from_xml_matches_oracle(tmp_path, oracle_mod):
    """The reference's calling sequence end to end: TomoGeometry read from an
    XML geometry file (frame origin off the volume centre, fulcrum below
    it), then Projector and recons:: on the GPU; bit-exact against the oracle
    on the kernel arguments that file yields (from_xml, pinned against the
    reference's own reading in tests/test_geometry_xml.py)."""
    from kevernals_amd import geometry as G
    from kevernals_amd.phantom import phantom

    assert os.path.exists(CHECK), "drop-in not built"
    g = G.scaled_down(G.small_plain(), 2)
    fz = -12.0
    G.to_xml(g, str(tmp_path / "geom.xml"), center=(10.5, -3.25, 7.0), fulcrum=(0.0, 0.0, fz))
    gm = G.from_xml(str(tmp_path / "geom.xml"))
    o = oracle_mod.Oracle(gm)
    vol = phantom(gm, 5)
    proj = o.fp_r(phantom(gm, 6))
    vol.tofile(tmp_path / "vol.bin")
    proj.tofile(tmp_path / "proj.bin")
    r = subprocess.run([CHECK, str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    status = dict(line.split() for line in open(tmp_path / "status.txt"))
    assert status and all(v == "0" for v in status.values())

    def rd(name, shape):
        return np.fromfile(tmp_path / name, dtype=np.float32).reshape(shape)

    np.testing.assert_array_equal(rd("fp.bin", gm.proj_shape), o.fp_r(vol))
    np.testing.assert_array_equal(rd("bp.bin", gm.vol_shape), o.bp_r(proj))
    np.testing.assert_array_equal(rd("mlem.bin", gm.vol_shape), o.mlem(proj, vol, 3, 1.0))
    np.testing.assert_array_equal(rd("art.bin", gm.vol_shape), o.art(proj, vol, 2, 0.5))
    planes = gm.plane_heights()
    sid = np.float32(abs(np.float32(gm.det[0, 2]) - np.float32(gm.src[0, 2])))
    rc, saa = oracle_mod.saa(proj, gm.src[:, 1], planes, float(sid), fz, gm.sy)
    assert rc == 0
    np.testing.assert_array_equal(rd("saa.bin", (len(planes), gm.nv, gm.nu)), saa)


def test_dropin_accelerated_mlem_matches_python_schedule(tmp_path):
    """recons::MLEMAccelerated (C++ over kv_mlem_accelerated) against
    learned.accelerated_mlem on the marshalled geometry: the same volume bit
    for bit and the same steps kept, without and with the committed CNN."""
    import torch

    import kevernals_amd as K
    from kevernals_amd import geometry as G
    from kevernals_amd.learned import Cnn, accelerated_mlem, transmission_data
    from kevernals_amd.phantom import phantom

    assert os.path.exists(CHECK), "drop-in not built"
    g = G.arc_geometry((96, 80, 16), (0.5, 0.5, 2.0), (192, 160), (0.5, 0.5), 9, 15.0)
    gm, sz, dz = marshalled(g)
    ctx = K.Context(0, gm)
    data = transmission_data(ctx, torch.from_numpy(phantom(gm, 8600)).cuda(), 1e4, 8600)
    wpath = os.path.join(os.path.dirname(K.__file__), "weights", "cnn_schedule_to_100.npz")
    with np.load(wpath, allow_pickle=False) as f:
        params, scale = f["params"].astype(np.float32), float(f["scale"])
    assert scale == 20.0  # the drop-in check passes the default scale
    with open(tmp_path / "geom.txt", "w") as f:
        f.write(f"{g.nx} {g.ny} {g.nz} {g.sx!r} {g.sy!r} {g.sz!r} {g.nu} {g.nv} {g.nviews} {g.pu!r} {g.pv!r} 0.0\n")
        for p in range(g.nviews):
            f.write(" ".join(f"{float(v):.9g}" for v in [*g.src[p], *g.det[p]]) + "\n")
    phantom(gm, 5).tofile(tmp_path / "vol.bin")
    np.zeros(gm.proj_shape, np.float32).tofile(tmp_path / "proj.bin")
    data.cpu().numpy().tofile(tmp_path / "ell.bin")
    params.tofile(tmp_path / "cnn.bin")
    r = subprocess.run([CHECK, str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    status = dict(line.split() for line in open(tmp_path / "status.txt"))
    assert status["mlem_accelerated"] == "0" and status["mlem_accelerated_cnn"] == "0"
    for name, cnn in (("accel", None), ("accel_cnn", Cnn(ctx, params, scale))):
        run = accelerated_mlem(ctx, data, cnn=cnn)
        got = np.fromfile(tmp_path / f"{name}.bin", dtype=np.float32).reshape(gm.vol_shape)
        np.testing.assert_array_equal(got, run.x.cpu().numpy(), err_msg=name)
        assert int(status[f"{name}_kept"]) == run.accepted
