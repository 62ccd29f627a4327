"""GPU parity: the HIP kernels (through the C ABI) against the CPU oracle on
the same seeded inputs.

Bars: mode-R voxel index sequences and weights bit-exact; mode-R ray sums and
backprojections within 1e-5 relative (they are in fact bit-exact: same FP32
operation sequence); fused MLEM/SART iterations bit-exact against the
reference's composed host loops; mode A within FP32 summation-order
tolerance (rtol 1e-5) and adjoint to 1e-6; shift-and-add bit-exact.
"""
import numpy as np
import pytest

from conftest import rng, tiny_geometry

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


@pytest.fixture(scope="module")
def K():
    import kevernals_amd

    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return kevernals_amd


@pytest.fixture(scope="module")
def O(oracle_mod):
    return oracle_mod


def small():
    from kevernals_amd import geometry as G

    return G.small_plain()


def half_small():
    from kevernals_amd import geometry as G

    return G.scaled_down(G.small_plain(), 2)


def phantom_of(g, seed):
    from kevernals_amd.phantom import phantom

    return phantom(g, seed)


def report(name, a, b):
    m = b != 0
    rel = np.abs(a - b)[m] / np.abs(b[m]) if m.any() else np.zeros(1)
    exact = float(np.mean(a == b))
    print(f"[parity] {name}: exact {exact:.6f}  max rel {rel.max():.3e}", flush=True)
    return exact, float(rel.max())


# ------------------------------------------------------------ mode R -------
def test_fp_r_small_config_parity(K, O):
    g = small()
    mu = phantom_of(g, 1)
    ctx = K.Context(0, g)
    out = torch.empty(g.proj_shape, device="cuda")
    ctx.forward_project(dev(mu), out, K.MODE_R)
    got = host(out)
    ref = O.Oracle(g).fp_r(mu)
    exact, rel = report("fp_r 128x128x32/15", got, ref)
    assert rel <= 1e-5
    assert exact == 1.0  # same FP32 operation sequence


def test_fp_r_voxel_sequences_bit_exact(K, O):
    g = small()
    ctx = K.Context(0, g)
    o = O.Oracle(g)
    r = rng(21)
    rays = r.integers(0, g.nrays, size=3000).astype(np.int32)
    traces = ctx.trace_rays(rays, cap=g.nx + g.ny + g.nz + 3, mode=K.MODE_R)
    nonempty = 0
    for rid, (idx, w, cnt) in zip(rays.tolist(), traces):
        p, rem = divmod(rid, g.nu * g.nv)
        v, u = divmod(rem, g.nu)
        ri, rw = o.ray_trace(p, u, v)
        assert cnt == len(ri), (p, u, v)
        np.testing.assert_array_equal(idx, ri)
        np.testing.assert_array_equal(w, rw)
        nonempty += cnt > 0
    assert nonempty > 500


def test_bp_r_small_config_parity(K, O):
    g = small()
    proj = (0.5 + rng(22).random(g.proj_shape)).astype(np.float32)
    ctx = K.Context(0, g)
    out = torch.empty(g.vol_shape, device="cuda")
    ctx.back_project(dev(proj), out, K.MODE_R)
    got = host(out)
    z0, z1 = 12, 16  # bounded oracle work: four full slices at the real size
    ref = O.Oracle(g).bp_r_slices(proj, z0, z1)
    exact, rel = report("bp_r 128x128x32/15 slices 12-15", got[z0:z1], ref[z0:z1])
    assert rel <= 1e-5
    assert exact == 1.0


def test_bp_r_half_config_full_parity(K, O):
    g = half_small()
    proj = (0.5 + rng(23).random(g.proj_shape)).astype(np.float32)
    ctx = K.Context(0, g)
    out = torch.empty(g.vol_shape, device="cuda")
    ctx.back_project(dev(proj), out, K.MODE_R)
    exact, rel = report("bp_r 64x64x16/15", host(out), O.Oracle(g).bp_r(proj))
    assert rel <= 1e-5 and exact == 1.0


def test_fused_mlem_matches_reference_loop_bit_exact(K, O):
    g = half_small()
    o = O.Oracle(g)
    mu = phantom_of(g, 2)
    data = o.fp_r(mu)
    x0 = np.ones(g.vol_shape, np.float32)
    ctx = K.Context(0, g)
    x = dev(x0)
    d = dev(data)
    for _ in range(3):
        ctx.mlem_iteration(x, d, 1.0, K.MODE_R)
    ref = o.mlem(data, x0, 3, 1.0)
    exact, rel = report("mlem_r x3 64x64x16", host(x), ref)
    assert exact == 1.0


def test_fused_art_matches_reference_loop_bit_exact(K, O):
    g = half_small()
    o = O.Oracle(g)
    mu = phantom_of(g, 3)
    data = o.fp_r(mu)
    x0 = np.full(g.vol_shape, 0.01, np.float32)
    ctx = K.Context(0, g)
    x = dev(x0)
    for _ in range(3):
        ctx.sart_iteration(x, dev(data), 0.5)
    ref = o.art(data, x0, 3, 0.5)
    exact, rel = report("art x3 64x64x16", host(x), ref)
    assert exact == 1.0


def test_fp_r_early_exit_quirk(K, O):
    """An axis-parallel ray returns 0 (Projector.cu:79-85)."""
    from kevernals_amd.geometry import Geometry

    det = np.array([[-4.0, -4.0, -30.0]], np.float32)
    src = np.array([[-3.5, 0.0, 50.0]], np.float32)
    g = Geometry(4, 4, 4, 1.0, 1.0, 1.0, 8, 8, 1, 1.0, 1.0, det, src)
    ctx = K.Context(0, g)
    out = torch.empty(g.proj_shape, device="cuda")
    ctx.forward_project(dev(np.ones(g.vol_shape, np.float32)), out, K.MODE_R)
    got = host(out)
    assert got[0, 3, 0] == 0.0
    np.testing.assert_array_equal(got, O.Oracle(g).fp_r(np.ones(g.vol_shape, np.float32)))


def test_tiny_geometries_edge_cases(K, O):
    """Odd sizes (not multiples of the forward tiles or backprojection
    bricks), off-centre detector."""
    for vol, det in (((6, 5, 4), (12, 10)), ((7, 9, 3), (13, 5)), ((1, 1, 1), (3, 3))):
        g = tiny_geometry(nviews=3, vol=vol, det=det)
        o = O.Oracle(g)
        x = (0.1 + rng(24).random(g.vol_shape)).astype(np.float32)
        y = (0.1 + rng(25).random(g.proj_shape)).astype(np.float32)
        ctx = K.Context(0, g)
        fp = torch.empty(g.proj_shape, device="cuda")
        bp = torch.empty(g.vol_shape, device="cuda")
        ctx.forward_project(dev(x), fp, K.MODE_R)
        ctx.back_project(dev(y), bp, K.MODE_R)
        np.testing.assert_array_equal(host(fp), o.fp_r(x))
        np.testing.assert_array_equal(host(bp), o.bp_r(y))
        ctx.forward_project(dev(x), fp, K.MODE_A)
        ctx.back_project(dev(y), bp, K.MODE_A)
        np.testing.assert_allclose(host(fp), o.fp_a(x), rtol=1e-5, atol=1e-6)
        np.testing.assert_allclose(host(bp), o.bp_a(y), rtol=1e-5, atol=1e-6)


def walk_fast_condition(g):
    """Host mirror of the library's choice between the forward walk with end
    thresholds and 24-bit indices and the counting walk (kv_api.hip
    walk_fast_ok)."""
    if g.nx * g.ny >= 1 << 24:
        return False
    for a, (n, sp) in enumerate(((g.nx, g.sx), (g.ny, g.sy), (g.nz, g.sz))):
        smax = float(np.max(np.abs(g.src[:, a].astype(np.float64) / sp)))
        if (n + 2.0) * (smax + n / 2.0 + 2.0) >= 4194304.0:
            return False
    return True


def test_both_forward_walks_match_oracle(K, O):
    """The threshold walk (the workloads' geometries) and the counting walk
    (a geometry outside the threshold bound: 4096 voxels along x) both give
    the reference's ray sums bit for bit."""
    g_fast = tiny_geometry(nviews=3, vol=(37, 11, 5), det=(40, 16))
    g_slow = tiny_geometry(nviews=3, vol=(4096, 3, 3), det=(48, 6), spacing=(0.05, 1.0, 2.0), pitch=(5.0, 0.6),
                           sod=300.0, sid=500.0)
    assert walk_fast_condition(g_fast) and not walk_fast_condition(g_slow)
    for g in (g_fast, g_slow):
        o = O.Oracle(g)
        x = (0.1 + rng(27).random(g.vol_shape)).astype(np.float32)
        ctx = K.Context(0, g)
        fp = torch.empty(g.proj_shape, device="cuda")
        ctx.forward_project(dev(x), fp, K.MODE_R)
        ref = o.fp_r(x)
        assert np.count_nonzero(ref) > g.nrays // 4
        np.testing.assert_array_equal(host(fp), ref)


# ------------------------------------------------------------ mode A -------
def test_mode_a_parity_and_adjoint(K, O):
    g = half_small()
    o = O.Oracle(g)
    r = rng(26)
    x = r.random(g.vol_shape).astype(np.float32)
    y = r.random(g.proj_shape).astype(np.float32)
    ctx = K.Context(0, g)
    fp = torch.empty(g.proj_shape, device="cuda")
    bp = torch.empty(g.vol_shape, device="cuda")
    ctx.forward_project(dev(x), fp, K.MODE_A)
    ctx.back_project(dev(y), bp, K.MODE_A)
    Ax, Aty = host(fp), host(bp)
    report("fp_a 64x64x16", Ax, o.fp_a(x))
    report("bp_a 64x64x16", Aty, o.bp_a(y))
    np.testing.assert_allclose(Ax, o.fp_a(x), rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(Aty, o.bp_a(y), rtol=1e-5, atol=1e-5)
    lhs = float(np.dot(Ax.ravel().astype(np.float64), y.ravel()))
    rhs = float(np.dot(x.ravel().astype(np.float64), Aty.ravel()))
    print(f"[adjoint] <Ax,y>={lhs:.9e} <x,A^T y>={rhs:.9e} rel={abs(lhs - rhs) / abs(lhs):.2e}")
    assert abs(lhs - rhs) <= 1e-6 * abs(lhs)


def test_mode_a_sequences_match_oracle(K, O):
    g = half_small()
    ctx = K.Context(0, g)
    o = O.Oracle(g)
    rays = rng(27).integers(0, g.nrays, size=1000).astype(np.int32)
    for rid, (idx, w, cnt) in zip(rays.tolist(), ctx.trace_rays(rays, cap=512, mode=K.MODE_A)):
        p, rem = divmod(rid, g.nu * g.nv)
        v, u = divmod(rem, g.nu)
        ri, rw = o.ray_trace(p, u, v, mode="A")
        np.testing.assert_array_equal(idx, ri)
        np.testing.assert_allclose(w, rw, rtol=1e-6, atol=1e-7)


def test_mlem_a_against_fp64_oracle_and_loglik(K, O):
    g = half_small()
    o = O.Oracle(g)
    from oracle.rng import poisson_data

    mu = phantom_of(g, 4)
    data = poisson_data(o.fp_a(mu), 1e4, 5)
    ctx = K.Context(0, g)
    x = dev(np.ones(g.vol_shape, np.float32))
    d = dev(data)
    lls = [ctx.log_likelihood(d, x)]
    for _ in range(5):
        ctx.mlem_iteration(x, d, 1.0, K.MODE_A)
        lls.append(ctx.log_likelihood(d, x))
    ref = o.mlem_a64(data, np.ones(g.vol_shape), 5)
    got = host(x)
    rel = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"[parity] mlem_a x5 vs fp64 oracle: max rel {rel:.2e}")
    assert rel < 1e-4
    ll_ref = o.loglik_a64(data.astype(np.float64), got.astype(np.float64))
    assert abs(lls[-1] - ll_ref) <= 1e-5 * abs(ll_ref)
    assert all(b >= a - 1e-7 * abs(a) for a, b in zip(lls, lls[1:]))


def test_mlem_a_100_iterations_fp32_reference_vs_fp64(K, O):
    """The time-to-RMSE references are 100 MLEM-A iterations in FP32 on the
    GPU; the target names a 100-iteration FP64 MLEM.  Here both run on the
    same seeded data (64x64x16, 128^2 detector, 15 views, Poisson I0 1e4):
    the FP32 result's relative RMSE against the FP64 one bounds what the
    substitution moves a 1 % threshold by."""
    g = half_small()
    o = O.Oracle(g)
    from oracle.rng import poisson_data

    data = poisson_data(o.fp_a(phantom_of(g, 41)), 1e4, 42)
    ctx = K.Context(0, g)
    d = dev(data)
    x = torch.ones(g.vol_shape, device="cuda")
    for _ in range(100):
        ctx.mlem_iteration(x, d, 1.0, K.MODE_A)
    got = host(x).astype(np.float64)
    print("[parity] oracle: 100 FP64 MLEM-A iterations ...", flush=True)
    ref = o.mlem_a64(data.astype(np.float64), np.ones(g.vol_shape), 100)
    rmse = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
    x10 = o.mlem_a64(data.astype(np.float64), np.ones(g.vol_shape), 10)
    span = float(np.linalg.norm(x10 - ref) / np.linalg.norm(ref))
    print(f"[parity] 100 MLEM-A iterations, FP32 GPU vs FP64 oracle: relative RMSE {rmse:.2e} "
          f"(iteration 10 vs 100: {span:.2e})", flush=True)
    assert rmse < 1e-3  # a tenth of the 1 % threshold
    assert span > 10 * rmse


def test_log_likelihoods_share_one_walk_bit_for_bit(K, O):
    """log_likelihoods of 1..3 volumes (one ray walk) equals log_likelihood
    of each volume bit for bit, and each matches the FP64 oracle."""
    g = half_small()
    o = O.Oracle(g)
    from oracle.rng import poisson_data

    data = poisson_data(o.fp_a(phantom_of(g, 4)), 1e4, 5)
    ctx = K.Context(0, g)
    d = dev(data)
    vols = [dev(phantom_of(g, s) + np.float32(0.001)) for s in (5, 6, 7)]
    single = [ctx.log_likelihood(d, v) for v in vols]
    for n in (1, 2, 3):
        assert ctx.log_likelihoods(d, vols[:n]) == single[:n]
    assert ctx.log_likelihoods(d, [vols[2], vols[0]]) == [single[2], single[0]]
    for v, ll in zip(vols, single):
        ref = o.loglik_a64(data.astype(np.float64), host(v).astype(np.float64))
        assert abs(ll - ref) <= 1e-5 * abs(ref)
    for bad in ([], vols + vols[:1]):
        with pytest.raises(RuntimeError):
            ctx.log_likelihoods(d, bad)


# ---------------------------------------------------------- shift-and-add --
def test_forward_launches_split_over_views(K, O):
    """A 1 x 4096 detector over 4097 identical views takes more than 2^32
    threads in the forward kernels' one-dimensional, XCD-striped launch
    order (tile columns padded to whole stripes), so the views are split over
    two launches: every view must equal view 0, view 0 must equal the
    oracle's, and the likelihood walk over several volumes must equal the
    single-volume one."""
    from kevernals_amd import geometry as G

    g1 = G.arc_geometry((8, 8, 4), (2.0, 2.0, 4.0), (1, 4096), (16.0, 0.006), 1, 0.0)
    nviews = 4097
    g = g1.scaled(nviews=nviews, det=np.repeat(g1.det, nviews, axis=0), src=np.repeat(g1.src, nviews, axis=0))
    o1 = O.Oracle(g1)
    r = rng(31)
    x = r.random(g.vol_shape).astype(np.float32) + np.float32(0.01)
    ctx = K.Context(0, g)
    out = torch.empty(g.proj_shape, device="cuda")
    for mode, ref in ((K.MODE_R, o1.fp_r(x)), (K.MODE_A, o1.fp_a(x))):
        ctx.forward_project(dev(x), out, mode)
        got = host(out)
        assert (got == got[:1]).all(), f"mode {mode}: views differ"
        if mode == K.MODE_R:
            np.testing.assert_array_equal(got[:1], ref)
        else:
            np.testing.assert_allclose(got[:1], ref, rtol=1e-5, atol=1e-6)
    assert np.count_nonzero(got[0]) > 1000
    d = dev(np.ascontiguousarray(got) * np.float32(1.1))
    vols = [dev(x * np.float32(s)) for s in (1.0, 0.9, 1.2)]
    single = [ctx.log_likelihood(d, v) for v in vols]
    for n in (1, 2, 3):
        assert ctx.log_likelihoods(d, vols[:n]) == single[:n]


def test_shift_and_add_bit_exact(K, O):
    g = small()
    proj = rng(28).random(g.proj_shape).astype(np.float32)
    ctx = K.Context(0, g)
    planes = g.plane_heights()
    sid = float(abs(g.det[0, 2] - np.float32(g.src[:, 2].mean(dtype=np.float32))))
    out = torch.empty((len(planes), g.nv, g.nu), device="cuda")
    ctx.shift_and_add(dev(proj), out, g.src[:, 1], planes, sid, 0.0, g.sy)
    rc, ref = O.saa(proj, g.src[:, 1], planes, sid, 0.0, g.sy)
    assert rc == 0
    np.testing.assert_array_equal(host(out), ref)


def test_shift_and_add_row_shifts_closed_form(K, O):
    """The hand-derived landing rows of tests/test_oracle.py (impulse at row
    20 of view p, column p, lands at row 20 - t_p; both signs of the source
    y, both signs of mZ - mF, truncation toward zero) through the kernel."""
    from test_oracle import SAA_CASES, saa_impulses

    from kevernals_amd.geometry import Geometry

    imp = saa_impulses()
    P, nv, nu = imp.shape
    for fz, z, ys, rows in SAA_CASES:
        g = Geometry(4, 4, 2, 1.0, 0.5, 1.0, nu, nv, P, 1.0, 1.0, np.zeros((P, 3)),
                     np.array([[0.0, y, 100.0] for y in ys]))
        ctx = K.Context(0, g)
        out = torch.full((2, nv, nu), -1.0, device="cuda")
        ctx.shift_and_add(dev(imp), out, np.array(ys, np.float32), [z, fz], 100.0, fz, 0.5)
        got = host(out)
        expect = np.zeros((2, nv, nu), np.float32)
        for p, r in enumerate(rows):
            expect[0, r, p] = np.float32(0.2)
        expect[1, 20, :] = np.float32(0.2)
        np.testing.assert_array_equal(got, expect)
        rc, ref = O.saa(imp, np.array(ys, np.float32), [z, fz], 100.0, fz, 0.5)
        assert rc == 0
        np.testing.assert_array_equal(got, ref)


def test_shift_and_add_preconditions(K):
    g = small()
    ctx = K.Context(0, g)
    proj = torch.zeros(g.proj_shape, device="cuda")
    out = torch.zeros((2, g.nv, g.nu), device="cuda")
    with pytest.raises(K.KvError) as e:
        ctx.shift_and_add(proj, out, g.src[:, 1], [1.0], 1000.0, 0.0, g.sy)
    assert e.value.status == 5
    with pytest.raises(K.KvError) as e:
        ctx.shift_and_add(proj, out, g.src[:, 1], [1.0, 2.0], 1000.0, 1000.0, g.sy)
    assert e.value.status == 5


# ----------------------------------------------------------- error paths --
def test_error_statuses(K):
    ctx = K.Context(0)
    with pytest.raises(K.KvError) as e:
        ctx.forward_project(1, 1, K.MODE_R)
    assert e.value.status == 4  # no geometry
    from kevernals_amd.geometry import Geometry

    bad = Geometry(0, 4, 4, 1.0, 1.0, 1.0, 8, 8, 1, 1.0, 1.0, np.zeros((1, 3)), np.ones((1, 3)))
    with pytest.raises(K.KvError) as e:
        ctx.set_geometry(bad)
    assert e.value.status == 1
    ctx.set_geometry(tiny_geometry())
    with pytest.raises(K.KvError) as e:
        ctx.forward_project(1, 1, 7)
    assert e.value.status == 1


def test_host_variants_match_device(K):
    g = half_small()
    ctx = K.Context(0, g)
    x = rng(29).random(g.vol_shape).astype(np.float32)
    fp = torch.empty(g.proj_shape, device="cuda")
    ctx.forward_project(dev(x), fp, K.MODE_R)
    np.testing.assert_array_equal(ctx.forward_project_host(x), host(fp))
    y = host(fp)
    bp = torch.empty(g.vol_shape, device="cuda")
    ctx.back_project(dev(y), bp, K.MODE_R)
    np.testing.assert_array_equal(ctx.back_project_host(y), host(bp))


def test_bp_r_skipped_pixels_ignore_non_finite_values(K, O):
    """Pixels whose weight test fails are never read by the reference
    (Projector.cu:710-729), so INF and NaN there must not reach the sums;
    those of accepted pixels must (INF / NaN voxels, as the oracle).  Run on
    a geometry with the clamped-weight backprojection (every z plane at
    alpha >= 1e-5 along every view) and on one without it (the sources
    1e-4 above the top plane), mode R bit for bit, mode A to the oracle's
    tolerance."""
    from kevernals_amd.geometry import Geometry

    g_fast = half_small()
    det = np.array([[-20.0, -20.0, -30.0]] * 3, np.float32)
    src = np.array([[0.0, -3.0, 8.0001], [0.0, 0.0, 8.0001], [0.0, 3.0, 8.0001]], np.float32)  # top plane z = 8
    g_slow = Geometry(16, 16, 8, 1.0, 1.0, 2.0, 40, 40, 3, 1.0, 1.0, det, src)
    for g in (g_fast, g_slow):
        o = O.Oracle(g)
        proj = (0.5 + rng(51).random(g.proj_shape)).astype(np.float32)
        r = rng(52)
        flat = proj.reshape(-1)
        idx = r.choice(flat.size, size=flat.size // 50, replace=False)
        flat[idx[0::3]] = np.inf
        flat[idx[1::3]] = -np.inf
        flat[idx[2::3]] = np.nan
        ctx = K.Context(0, g)
        out = torch.empty(g.vol_shape, device="cuda")
        ctx.back_project(dev(proj), out, K.MODE_R)
        got, ref = host(out), o.bp_r(proj)
        np.testing.assert_array_equal(got, ref)  # NaN equal to NaN
        assert np.isfinite(ref).mean() > 0.2 and (~np.isfinite(ref)).any()
        ctx.back_project(dev(proj), out, K.MODE_A)
        got, ref = host(out), o.bp_a(proj)
        np.testing.assert_array_equal(np.isnan(got), np.isnan(ref))
        np.testing.assert_array_equal(np.isinf(got), np.isinf(ref))
        fin = np.isfinite(ref)
        np.testing.assert_allclose(got[fin], ref[fin], rtol=1e-5, atol=1e-6)


def test_fp_non_finite_voxels_reach_only_rays_through_them(K, O):
    """INF and NaN voxels enter the ray sums of the rays whose accepted
    segments cross them (Projector.cu:356-366) and no others; mode R bit for
    bit (NaN equal to NaN), mode A with the same non-finite pattern as the
    oracle and its tolerance elsewhere."""
    g = half_small()
    o = O.Oracle(g)
    x = (0.5 + rng(53).random(g.vol_shape)).astype(np.float32)
    flat = x.reshape(-1)
    idx = rng(54).choice(flat.size, size=30, replace=False)
    flat[idx[0::3]] = np.inf
    flat[idx[1::3]] = -np.inf
    flat[idx[2::3]] = np.nan
    ctx = K.Context(0, g)
    out = torch.empty(g.proj_shape, device="cuda")
    ctx.forward_project(dev(x), out, K.MODE_R)
    got, ref = host(out), o.fp_r(x)
    np.testing.assert_array_equal(got, ref)
    assert 0.001 < (~np.isfinite(ref)).mean() < 0.5
    ctx.forward_project(dev(x), out, K.MODE_A)
    got, ref = host(out), o.fp_a(x)
    np.testing.assert_array_equal(np.isnan(got), np.isnan(ref))
    np.testing.assert_array_equal(np.isinf(got), np.isinf(ref))
    fin = np.isfinite(ref)
    np.testing.assert_allclose(got[fin], ref[fin], rtol=1e-5, atol=1e-5)
