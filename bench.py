"""Benchmark of the tomosynthesis reconstruction hot path on one MI355X.

A step is one MLEM iteration of the reference discretisation (recons::MLEM,
KevLoq/Kevernals Reconstructors.h:182-242) on the full-size workload:
1024x1024x64 voxels (0.15x0.15x1 mm), 2048^2 detector (0.15 mm), 25 views
on +-20 deg, fused into two kernels (forward projection + ratio,
backprojection + update).  Inputs are resident in HBM before timing starts.

Output: one JSON line.  value = GRays/s forward+back = views*nu*nv / t(step).
roofline: the dominant kernel's algorithmic bytes per launch (accepted
(ray, voxel) pairs * 4 B + its streamed arrays) over its average duration,
measured with HIP events on the stream the kernels run on.  cpu_baseline: the
CPU oracle (a port of the reference arithmetic, OpenMP) timed on a bounded
sample of the same workload on this host, extrapolated to one full step.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# rocprofv3 (demangled) names of the two fused MLEM kernels
FP_KERNEL = "fp_r_kernel<1>"
BP_KERNEL = "bp_tiled_kernel<0, 1, 8, false, true>"
METRIC = "time to RMSE≤1% vs converged MLEM, learned vs plain; GRays/s fwd+back"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_data(K, ctx, g, torch, seed=2):
    """Seeded phantom with 200 microcalcifications; Poisson transmission data
    at I0 = 5e3; reference-discretisation data p = FP_R(mu) * l_hat / l."""
    from kevernals_amd.phantom import phantom

    t = time.time()
    mu_h = phantom(g, seed, microcalcifications=200)
    mu = torch.from_numpy(mu_h).cuda()
    ell = torch.empty(g.proj_shape, device="cuda")
    ctx.forward_project(mu, ell, K.MODE_A)
    torch.cuda.synchronize()
    ell_hat = torch.empty_like(ell)
    ctx.poisson_data(ell, ell_hat, 5e3, seed + 1000)
    fp = torch.empty(g.proj_shape, device="cuda")
    ctx.forward_project(mu, fp, K.MODE_R)
    data = torch.where(ell > 0, fp * ell_hat / torch.clamp(ell, min=1e-30), torch.zeros_like(fp))
    torch.cuda.synchronize()
    log(f"[bench] data ready in {time.time() - t:.1f} s")
    return mu, data


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def _cpu_sample(o, g, mu_h, data_h, seconds, rng):
    """Oracle FP (whole detector rows, every view) and BP (random voxels),
    sized to about `seconds`, extrapolated to one full forward+back pass."""
    rows_per_view = 2
    t = time.time()
    for p in range(g.nviews):
        v0 = int(rng.integers(0, g.nv - rows_per_view))
        o.fp_r_rows(mu_h, p, v0, v0 + rows_per_view)
    t_fp_ray = (time.time() - t) / (g.nviews * rows_per_view * g.nu)
    vox = rng.integers(0, g.nvox, size=2000)
    t = time.time()
    o.bp_r_voxels(data_h, vox)
    t_bp_vox = (time.time() - t) / len(vox)
    budget = max(seconds, 1.0)
    est_fp = t_fp_ray * g.nrays
    est_bp = t_bp_vox * g.nvox
    share_fp = est_fp / (est_fp + est_bp)
    n_rows = max(1, int(budget * share_fp / (t_fp_ray * g.nu * g.nviews)))
    n_vox = max(1000, int(budget * (1 - share_fp) / t_bp_vox))
    t = time.time()
    for p in range(g.nviews):
        v0 = int(rng.integers(0, max(1, g.nv - n_rows)))
        o.fp_r_rows(mu_h, p, v0, min(g.nv, v0 + n_rows))
    fp_s = time.time() - t
    rays_done = g.nviews * min(n_rows, g.nv) * g.nu
    vox = rng.integers(0, g.nvox, size=n_vox)
    t = time.time()
    o.bp_r_voxels(data_h, vox)
    bp_s = time.time() - t
    full_s = fp_s / rays_done * g.nrays + bp_s / n_vox * g.nvox
    return full_s, rays_done, n_rows, n_vox, fp_s + bp_s


def cpu_small_full(nt):
    """The 128x128x32 workload run in full by the oracle (mode R, OpenMP):
    forward projection, backprojection and 10 MLEM iterations, as the
    reference's host loops compose them (Reconstructors.h:182-242)."""
    from kevernals_amd import geometry as G
    from kevernals_amd.phantom import phantom
    from oracle.oracle import Oracle

    g = G.small_plain()
    o = Oracle(g, nt)
    mu = phantom(g, 1)
    t = time.time()
    fp = o.fp_r(mu)
    fp_s = time.time() - t
    t = time.time()
    o.bp_r(fp)
    bp_s = time.time() - t
    t = time.time()
    o.mlem(fp, np.ones(g.vol_shape, np.float32), 10, 1.0)
    mlem_s = time.time() - t
    return {"workload": "128x128x32, 256^2 detector, 15 views", "threads": nt,
            "forward_s": round(fp_s, 3), "backprojection_s": round(bp_s, 3),
            "mlem_10_iterations_s": round(mlem_s, 3),
            "grays_per_s": round(g.nrays / (mlem_s / 10) / 1e9, 6)}


def cpu_baseline(g, mu_h, data_h, seconds):
    """Oracle FP+BP (mode R, OpenMP over rays / voxels, all host threads) on a
    sample spread over all views and slices, extrapolated to one full
    forward+back pass; the same on one thread with a quarter of the budget;
    and the small workload in full."""
    from oracle.oracle import Oracle

    nt = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    rng = np.random.Generator(np.random.PCG64(7))
    full_s, rays_done, n_rows, n_vox, spent = _cpu_sample(Oracle(g, nt), g, mu_h, data_h, seconds, rng)
    one_s, _, rows1, vox1, spent1 = _cpu_sample(Oracle(g, 1), g, mu_h, data_h, seconds / 4, rng)
    small = cpu_small_full(nt)
    return {
        "value": g.nrays / full_s / 1e9,
        "unit": "GRays/s",
        "cores": nt,
        "kind": "port",
        "cpu_model": cpu_model(),
        "online_cpus": os.cpu_count(),
        "sample": (f"oracle FP on {rays_done} rays ({n_rows} rows in each of {g.nviews} views) "
                   f"+ BP on {n_vox} random voxels, {spent:.1f} s, extrapolated to one "
                   f"full forward+back pass ({full_s:.0f} s)"),
        "seconds_per_step_extrapolated": full_s,
        "one_core": {"value": g.nrays / one_s / 1e9, "unit": "GRays/s",
                     "sample": f"{rows1} rows per view + {vox1} voxels on 1 thread, {spent1:.1f} s",
                     "seconds_per_step_extrapolated": one_s},
        "small_workload_full": small,
    }


def traffic_from_profiles(kernel):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    try:
        return json.load(open(path)).get(kernel, {}).get("hbm_bytes_per_launch")
    except Exception:
        return None


def bench_batched(args, torch, K, G):
    """32 uneven plain-MLEM reconstructions (seed 3; phantoms 3000+i), 15
    iterations each per step, in the persistent batched launches replayed
    from one captured hipGraph; tiles chosen by the autotuner."""
    from kevernals_amd.phantom import phantom

    gs = G.batched_uneven(3, 32)
    iters = 15
    t = time.time()
    b = K.Batch(gs)
    for i, g in enumerate(gs):
        ctx = K.Context(0, g)
        mu = torch.from_numpy(phantom(g, 3000 + i)).cuda()
        fp = torch.empty(g.proj_shape, device="cuda")
        ctx.forward_project(mu, fp, K.MODE_R)
        torch.cuda.synchronize()
        b.set_data(i, fp.cpu().numpy())
        b.set_volume(i, np.ones(g.vol_shape, np.float32))
        del ctx, mu, fp
    log(f"[bench] batch of {len(gs)} ready in {time.time() - t:.1f} s")
    t = time.time()
    tiles = b.autotune(args.tile_cache)
    log(f"[bench] autotune {time.time() - t:.1f} s: tiles {sorted(set(tiles))}")
    for _ in range(args.warmup):
        b.mlem(iters, use_graph=True)
    torch.cuda.synchronize()
    dev_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dev_ms.append(b.mlem(iters, use_graph=True))
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_step = (t1 - t0) * 1e3 / args.steps
    rays = sum(g.nrays for g in gs) * iters
    vox = sum(g.nvox for g in gs)
    value = rays / (ms_step * 1e-3) / 1e9
    log(f"[bench] batched step {ms_step:.2f} ms (device {np.mean(dev_ms):.2f} ms), {value:.3f} GRays/s")
    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "GRays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded ellipsoid phantoms, noise-free reference-discretisation projections)",
        "config": {
            "workload": "batched plain MLEM: 32 uneven reconstructions x 15 iterations per step, "
                        "persistent launches replayed from a hipGraph",
            "reconstructions": len(gs),
            "iterations": iters,
            "total_voxels": vox,
            "rays_per_iteration": sum(g.nrays for g in gs),
            "tiles": [list(x) for x in tiles],
            "parallelism": "single GPU",
        },
        "mlem_iters_per_s": round(len(gs) * iters / (ms_step * 1e-3), 2),
        "device_ms_per_step": round(float(np.mean(dev_ms)), 3),
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


CNN_FLOP_PER_VOXEL = 2.0 * 27 * (16 + 256 + 256 + 16)  # one forward application


def bench_cnn(args, torch, K, G):
    """The CNN alone: inference on a 1024x1024x64 volume and training steps
    (forward + backward + Adam) on 256x256x32, against the dense BF16 MFMA
    peak.  Algorithmic FLOPs: 2*27*(1*16 + 16*16 + 16*16 + 16*1) per voxel
    forward; a training step counts 3x (forward, input and weight
    gradients)."""
    from kevernals_amd.learned import Cnn

    ctx = K.Context(0)
    cnn = Cnn(ctx, seed=1)
    res = {}
    for name, shape, train in (("inference", (64, 1024, 1024), False), ("training", (32, 256, 256), True)):
        x = 0.02 + 0.05 * torch.rand(shape, device="cuda")
        y = torch.empty_like(x)
        tgt = x + 0.01 * torch.randn(shape, device="cuda")

        def step():
            if train:
                cnn.train_step(x, tgt, 1e-4)
            else:
                cnn.forward(x, y)

        for _ in range(max(1, args.warmup)):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / args.steps
        nvox = shape[0] * shape[1] * shape[2]
        flops = CNN_FLOP_PER_VOXEL * nvox * (3 if train else 1)
        tfs = flops / (ms * 1e-3) / 1e12
        res[name] = {"shape": list(shape), "ms": round(ms, 3), "tflops": round(tfs, 1),
                     "frac_of_bf16_peak": round(tfs / 2500.0, 4)}
        log(f"[bench] cnn {name} {shape}: {ms:.3f} ms, {tfs:.1f} TFLOP/s")
        del x, y, tgt
    out = {
        "metric": METRIC, "value": res["inference"]["tflops"], "unit": "TFLOP/s (CNN inference, dense bf16)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": res["inference"]["ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16 (f32 accumulate)",
        "data": "synthetic (uniform random volumes)",
        "config": {"workload": "3D residual CNN 1-16-16-16-1, 3x3x3", "parallelism": "single GPU"},
        "cnn": res,
        "roofline": {"bound": "mfma", "achieved": res["inference"]["tflops"], "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": res["inference"]["frac_of_bf16_peak"], "traffic": None},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def _dicom_frame_file(path, frame, acquisition_time):
    """A minimal DICOM file (PS3.10, explicit VR little endian) holding one
    16-bit frame: the input of the ingest measurement."""
    import struct

    def el(g, e, vr, v):
        if len(v) % 2:
            v += b"\x00" if vr in ("UI", "OB") else b" "
        if vr in ("OB", "OW"):
            return struct.pack("<HH", g, e) + vr.encode() + b"\x00\x00" + struct.pack("<I", len(v)) + v
        return struct.pack("<HH", g, e) + vr.encode() + struct.pack("<H", len(v)) + v

    meta = el(2, 1, "OB", b"\x00\x01") + el(2, 0x10, "UI", b"1.2.840.10008.1.2.1")
    us = lambda e, v: el(0x28, e, "US", struct.pack("<H", v))
    body = (el(8, 0x32, "TM", acquisition_time.encode()) + el(0x10, 0x20, "LO", b"BENCH") + us(2, 1) +
            el(0x28, 4, "CS", b"MONOCHROME2") + us(0x10, frame.shape[0]) + us(0x11, frame.shape[1]) +
            us(0x100, 16) + us(0x101, 14) + us(0x102, 13) + us(0x103, 0) +
            el(0x7FE0, 0x10, "OW", np.ascontiguousarray(frame, "<u2").tobytes()))
    with open(path, "wb") as f:
        f.write(b"\x00" * 128 + b"DICM" + el(2, 0, "UL", struct.pack("<I", len(meta))) + meta + body)


def bench_io(args, torch, K, G):
    """The data formats either side of the path at the full size, on the
    host: the TIFF dump of a 1024x1024x64 volume as recons:: verbose mode
    writes it (device-to-host copy, PackBits, file write), and the DICOM
    ingest of 25 projections of 2048x2048 16-bit pixels (read and decode,
    crop to the detector region, host-to-device copy, log transform on the
    GPU).  Random values: PackBits finds no runs (its worst case)."""
    import shutil
    import tempfile

    from kevernals_amd import io as KIO

    g = G.full_plain()
    ctx = K.Context(0, g)
    ctx.use_torch_stream()
    tmp = tempfile.mkdtemp(prefix="kv_io_")
    try:
        vol = torch.rand(g.vol_shape, device="cuda")
        host = torch.empty(g.vol_shape, dtype=torch.float32, pin_memory=True)
        tif = os.path.join(tmp, "reconstructionstep0.tiff")
        dump = []
        for i in range(args.warmup + args.steps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            host.copy_(vol)
            KIO.write_tiff(tif, host.numpy(), spacing=(g.sx, g.sy))
            if i >= args.warmup:
                dump.append(time.perf_counter() - t0)
        vbytes = vol.numel() * 4
        fsize = os.path.getsize(tif)
        del vol, host
        log(f"[bench] TIFF dump of {g.vol_shape}: {np.median(dump) * 1e3:.1f} ms, "
            f"{vbytes / np.median(dump) / 1e9:.2f} GB/s, file {fsize / vbytes:.4f} x the volume")

        rng = np.random.default_rng(3)
        ddir = os.path.join(tmp, "dcm")
        os.makedirs(ddir)
        for p in range(g.nviews):
            _dicom_frame_file(os.path.join(ddir, f"p{p:02d}.dcm"), rng.integers(0, 1 << 14, (g.nv, g.nu)),
                              f"{120000 + p:06d}")
        roi = (0, 0, g.nu, g.nv)
        d = torch.empty((g.nviews, g.nv, g.nu), device="cuda")
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ingest, log_ms = [], []
        for i in range(args.warmup + args.steps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            img = KIO.read_dicom_directory(ddir, roi=roi)
            d.copy_(torch.from_numpy(img.data))
            ev[0].record()
            ctx.log_scale(d, d, 1.0)
            ev[1].record()
            torch.cuda.synchronize()
            if i >= args.warmup:
                ingest.append(time.perf_counter() - t0)
                log_ms.append(ev[0].elapsed_time(ev[1]))
        pbytes = g.nviews * g.nv * g.nu * 2
        t_ingest = float(np.median(ingest))
        t_log = float(np.median(log_ms)) * 1e-3
        log_gbs = d.numel() * 8 / t_log / 1e9
        log(f"[bench] DICOM ingest of {g.nviews} x {g.nv}x{g.nu} u16: {t_ingest * 1e3:.1f} ms, "
            f"{pbytes / t_ingest / 1e9:.2f} GB/s of pixel data; log transform {t_log * 1e3:.3f} ms "
            f"({log_gbs:.0f} GB/s)")
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    out = {
        "metric": METRIC, "value": round(pbytes / t_ingest / 1e9, 3), "unit": "GB/s (DICOM ingest, host)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(t_ingest * 1e3, 1),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u16 in, f32 out",
        "data": "synthetic (uniform random pixels and voxels)",
        "config": {"workload": "DICOM ingest of 25 x 2048x2048 u16 projections; TIFF dump of a 1024x1024x64 "
                               "float volume", "parallelism": "single GPU, host I/O"},
        "io": {"tiff_dump_ms": round(float(np.median(dump)) * 1e3, 1),
               "tiff_dump_gbs": round(vbytes / float(np.median(dump)) / 1e9, 3),
               "tiff_size_ratio": round(fsize / vbytes, 5),
               "dicom_ingest_ms": round(t_ingest * 1e3, 1), "log_transform_ms": round(t_log * 1e3, 4)},
        "roofline": {"bound": "hbm", "achieved": round(log_gbs, 1), "peak": 8000.0, "unit": "GB/s",
                     "frac": round(log_gbs / 8000.0, 4), "traffic": None},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def _times(trace, thr, rmse_of):
    """(seconds, MLEM iterations run) at the first trace entry with RMSE <= thr."""
    for e in trace:
        r = rmse_of(e)
        if r is not None and r <= thr:
            return e[2], e[1]
    return None, None


def _thresholds(runs, rmse_of, label):
    out = {}
    for thr in (0.10, 0.05, 0.02, 0.01):
        row = {}
        for name, run in runs.items():
            t, k = _times(run.trace, thr, rmse_of)
            row[f"{name}_s"] = None if t is None else round(t, 4)
            row[f"{name}_iters"] = k
        tp = row["plain_s"]
        for name in runs:
            if name != "plain":
                tl = row[f"{name}_s"]
                row[f"{name}_speedup"] = round(tp / tl, 3) if (tp is not None and tl) else None
        out[f"{thr:g}"] = row
        log(f"[bench] {label} RMSE <= {thr:g}: " + ", ".join(
            f"{n} {row[n + '_s']} s ({row[n + '_iters']} it)" for n in runs))
    return out


LEARNED_WEIGHTS = "cnn_schedule_to_100.npz"


def _load_cnn_params(name):
    with np.load(_weights_path(name), allow_pickle=False) as f:
        return f["params"], float(f["scale"])


def _weights_path(name):
    return name if os.path.sep in name else os.path.join(ROOT, "kevernals_amd", "weights", name)


def _learned_schedule(name):
    """The schedule a network was trained after: its probe length (`second`,
    stored with the weights; 8 when absent)."""
    from kevernals_amd.trajectory import Schedule

    with np.load(_weights_path(name), allow_pickle=False) as f:
        second = int(f["second"]) if "second" in f.files else 8
    return Schedule(second=second)


def learned_batched_runs(torch, K, G, n, weights=LEARNED_WEIGHTS):
    """Learned plus batched at full size: n reconstructions of
    1024x1024x64 / 25 views (phantom seeds 5000+i with 200
    microcalcifications, I0 = 5e3) in one mode-A batch (persistent launches,
    hipGraph replay).  Plain batched MLEM-A against the accelerated batched
    reconstruction: the two-stretch trajectory-jump schedule (18 batched
    MLEM iterations; each reconstruction's jumps gated on its Poisson
    log-likelihood) with the likelihood-gated CNN step on its result
    (weights cnn_schedule_to_100.npz); the schedule without the network is
    timed too.  Time to every reconstruction's RMSE <= threshold against its
    own 100-iteration MLEM-A result; times exclude the RMSE evaluations."""
    from kevernals_amd.learned import batched_accelerated_mlem, batched_gated_mlem, transmission_data
    from kevernals_amd.phantom import phantom
    from kevernals_amd.trajectory import Schedule

    g = G.full_plain()
    t = time.time()
    b = K.Batch([g] * n, mode=K.MODE_A)
    for i in range(n):
        ctx = b.context(i)
        mu = torch.from_numpy(phantom(g, 5000 + i, microcalcifications=200)).cuda()
        data = transmission_data(ctx, mu, 5e3, 5000 + i)
        ctx.copy(b.buffers(i)[1], data, data.numel() * 4)
        torch.cuda.synchronize()
        del mu, data
    ones = np.ones(g.vol_shape, np.float32)

    def reset():
        for i in range(n):
            b.set_volume(i, ones)

    reset()
    log(f"[bench] {n} reconstructions ready in {time.time() - t:.1f} s")
    t = time.perf_counter()
    b.mlem(100, use_graph=True)
    torch.cuda.synchronize()
    t_ref = time.perf_counter() - t
    refs = []
    for i in range(n):
        r = torch.empty(g.vol_shape, device="cuda")
        b.context(i).copy(r, b.buffers(i)[0], r.numel() * 4)
        refs.append(r)
    torch.cuda.synchronize()
    log(f"[bench] references: 100 batched MLEM-A iterations in {t_ref:.2f} s")
    params, scale = _load_cnn_params(weights)
    sch = Schedule()
    sch_l = _learned_schedule(weights)
    runs = {}
    reset()
    runs["plain"] = batched_gated_mlem(b, 100, params=None, references=refs)
    reset()
    runs["learned"] = batched_accelerated_mlem(b, sch_l, params=params, scale=scale, references=refs)
    reset()
    runs["schedule_only"] = batched_accelerated_mlem(b, sch, params=None, references=refs)
    out_t = _thresholds(runs, lambda e: (max(e[3]) if e[3] is not None else None), f"all {n}")
    res = {
        "reconstructions": n, "volume": [g.nx, g.ny, g.nz], "views": g.nviews,
        "schedule": {"start": sch.start, "first": sch.first, "mid": sch.mid, "second": sch.second,
                     "horizon": sch.horizon, "mlem_iterations": sch.cost},
        "learned_schedule": {"second": sch_l.second, "mlem_iterations": sch_l.cost},
        "cnn": f"{os.path.basename(weights)}, gated, on the final jump",
        "time_to_rmse": out_t,
        "gate": {"accepted": runs["learned"].accepted, "rejected": runs["learned"].rejected},
        "reference": "100 batched MLEM-A iterations, FP32",
        "reference_seconds": round(t_ref, 3),
        "final_rmse": {"learned": [float(v) for v in runs["learned"].trace[-1][3]],
                       "schedule_only": [float(v) for v in runs["schedule_only"].trace[-1][3]]},
    }
    del b, refs
    return res, runs


def bench_learned_batched(args, torch, K, G):
    """The learned plus batched workload alone (learned_batched_runs) as the
    JSON line's value: plain batched time / accelerated time to RMSE <= 1%."""
    res, runs = learned_batched_runs(torch, K, G, args.recons, args.cnn_weights)
    s1 = res["time_to_rmse"]["0.01"]
    out = {
        "metric": METRIC,
        "value": s1["learned_speedup"],
        "unit": "x (plain batched time / learned batched time to RMSE <= 1%)",
        "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": round(1e3 * (s1["learned_s"] or runs["learned"].trace[-1][2]), 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 (MLEM, jumps with FP64 Gram sums), bf16/f32-accumulate (CNN)",
        "data": "synthetic (seeded phantoms + 200 microcalcifications, Poisson I0=5e3)",
        "config": {"workload": "learned plus batched MLEM-A, time to RMSE<=1% vs 100-iteration MLEM",
                   "reconstructions": res["reconstructions"], "volume": res["volume"], "views": res["views"],
                   "schedule": res["schedule"], "cnn": res["cnn"], "parallelism": "single GPU"},
        "time_to_rmse": res["time_to_rmse"],
        "gate": res["gate"],
        "reference_seconds": res["reference_seconds"],
        "final_rmse": res["final_rmse"],
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def bench_learned(args, torch, K, G):
    """Time to RMSE <= 1% against the 100-iteration MLEM-A result on the
    full-size workload (1024x1024x64, 25 views, phantom seed 5000 with 200
    microcalcifications, I0 = 5e3): plain MLEM-A against the accelerated
    reconstruction (the two-stretch trajectory-jump schedule, 18 MLEM
    iterations and two likelihood-gated jumps, with the gated CNN step on
    its result, weights cnn_schedule_to_100.npz).  Also timed: the schedule
    without the network, and the CNN interleaved with plain MLEM (candidate
    after iteration 5 and every 2 iterations, weights cnn_mlem5_to_100.npz).
    All runs are mode-A MLEM on the same device; times exclude the RMSE
    evaluations."""
    from kevernals_amd.learned import Cnn, accelerated_mlem, gated_mlem, transmission_data
    from kevernals_amd.phantom import phantom
    from kevernals_amd.trajectory import Schedule

    g = G.full_plain()
    ctx = K.Context(0, g)
    mu = torch.from_numpy(phantom(g, 5000, microcalcifications=200)).cuda()
    data = transmission_data(ctx, mu, 5e3, 5000)
    del mu
    ref = torch.ones(g.vol_shape, device="cuda")
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(100):
        ctx.mlem_iteration(ref, data, 1.0, K.MODE_A)
    torch.cuda.synchronize()
    t_ref = time.perf_counter() - t
    log(f"[bench] reference: 100 MLEM-A iterations in {t_ref:.2f} s")
    cnn = Cnn(ctx, *_load_cnn_params("cnn_schedule_to_100.npz"))
    cnn_old = Cnn(ctx, *_load_cnn_params("cnn_mlem5_to_100.npz"))
    # CNN inference alone (MFMA roofline of the dominant kernels)
    y = torch.empty_like(ref)
    cnn.forward(ref, y)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(5):
        cnn.forward(ref, y)
    e1.record()
    torch.cuda.synchronize()
    cnn_ms = e0.elapsed_time(e1) / 5
    flops = 2.0 * 27 * (16 + 256 + 256 + 16) * g.nvox
    del y
    sch = Schedule()
    runs = {
        "plain": gated_mlem(ctx, None, data, 100, reference=ref),
        "learned": accelerated_mlem(ctx, data, sch, cnn=cnn, reference=ref),
        "schedule_only": accelerated_mlem(ctx, data, sch, reference=ref),
        "interleaved_cnn": gated_mlem(ctx, cnn_old, data, 100, every=2, first=5, reference=ref, patience=2),
    }
    out_t = _thresholds(runs, lambda e: e[3], "")
    s1 = out_t["0.01"]
    lr = runs["learned"]
    out = {
        "metric": METRIC,
        "value": s1["learned_speedup"],
        "unit": "x (plain time / learned time to RMSE <= 1%)",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 0,
        "ms_per_step": round(1e3 * (s1["learned_s"] or lr.trace[-1][2]), 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (MLEM, jumps with FP64 Gram sums), bf16/f32-accumulate (CNN)",
        "data": "synthetic (seeded phantom + 200 microcalcifications, Poisson I0=5e3)",
        "config": {"workload": "learned vs plain MLEM-A, time to RMSE<=1% vs 100-iteration MLEM",
                   "volume": [g.nx, g.ny, g.nz], "views": g.nviews,
                   "schedule": {"start": sch.start, "first": sch.first, "mid": sch.mid, "second": sch.second,
                                "horizon": sch.horizon, "mlem_iterations": sch.cost},
                   "cnn": "cnn_schedule_to_100.npz, gated, on the final jump",
                   "parallelism": "single GPU"},
        "time_to_rmse": out_t,
        "steps_kept": [e[0] for e in lr.trace if e[0] not in ("mlem", "probe")],
        "cnn": {"ms_per_application": round(cnn_ms, 3),
                "roofline": {"bound": "mfma", "achieved": round(flops / (cnn_ms * 1e-3) / 1e12, 1),
                             "peak": 2500.0, "unit": "TFLOP/s",
                             "frac": round(flops / (cnn_ms * 1e-3) / 2.5e15, 4)}},
        "reference_seconds": round(t_ref, 3),
        "final_rmse": {k: v.trace[-1][3] for k, v in runs.items()},
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def nnz_fixture(name):
    """The oracle's accepted-pair counts (tests/golden/nnz/nnz.json)."""
    path = os.path.join(ROOT, "tests", "golden", "nnz", "nnz.json")
    try:
        with open(path) as f:
            return json.load(f).get(name)
    except (OSError, ValueError):
        return None


def _event_ms(torch, fn, reps):
    """Average device time of fn() over reps calls, HIP events on torch's
    current stream (the context runs on it)."""
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def bench_sart_saa(K, ctx, g, data, torch, saa_reps=5):
    """The other two loops of the full-size workload: SART with relaxation
    0.5 for 10 iterations (recons::ART, Reconstructors.h:19-134: fused
    forward + residual, backprojection + update) and shift-and-add at the
    volume's 64 plane heights (recons::ShiftAndAdd, :263-379; fulcrum at 0,
    the projections being the workload's data)."""
    x = torch.ones(g.vol_shape, device="cuda")
    ctx.sart_iteration(x, data, 0.5)  # warm-up
    x.fill_(1.0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sart_ms = _event_ms(torch, lambda: ctx.sart_iteration(x, data, 0.5), 10)
    sart_wall = (time.perf_counter() - t0) * 1e3 / 10
    del x
    planes = g.plane_heights()
    sid = float(abs(g.det[0, 2] - np.float32(g.src[:, 2].mean(dtype=np.float32))))
    out = torch.empty((len(planes), g.nv, g.nu), device="cuda")
    ctx.shift_and_add(data, out, g.src[:, 1], planes, sid, 0.0, g.sy)
    saa_ms = _event_ms(torch, lambda: ctx.shift_and_add(data, out, g.src[:, 1], planes, sid, 0.0, g.sy),
                       saa_reps)
    del out
    plane_px = g.nu * g.nv * len(planes)
    saa_bytes = 4 * g.nviews * plane_px + 4 * plane_px  # every view read per output pixel + the write
    saa_gbs = saa_bytes / (saa_ms * 1e-3) / 1e9
    log(f"[bench] SART {sart_ms:.3f} ms per iteration; shift-and-add ({len(planes)} planes) {saa_ms:.3f} ms")
    return {
        "sart": {"relaxation": 0.5, "iterations": 10, "ms_per_iteration": round(sart_ms, 4),
                 "wall_ms_per_iteration": round(sart_wall, 4),
                 "grays_per_s": round(g.nrays / (sart_ms * 1e-3) / 1e9, 4)},
        "shift_and_add": {"planes": len(planes), "ms": round(saa_ms, 4),
                          "roofline": {"bound": "hbm", "kernel": "saa_kernel",
                                       "achieved": round(saa_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                       "frac": round(saa_gbs / HBM_PEAK_GBS, 4),
                                       "traffic": traffic_from_profiles("saa_kernel"),
                                       "algorithmic_bytes": saa_bytes}},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="budget of the CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-learned", action="store_true",
                    help="full workload: skip the learned-vs-plain time-to-RMSE measurement")
    ap.add_argument("--no-loops", action="store_true", help="full workload: skip the SART and shift-and-add timing")
    ap.add_argument("--workload", choices=["full", "small", "batched", "learned", "learned-batched", "cnn", "io"],
                    default="full")
    ap.add_argument("--recons", type=int, default=8, help="reconstructions of the learned-batched workload")
    ap.add_argument("--tile-cache", default=None, help="autotuner cache file (batched workload)")
    ap.add_argument("--cnn-weights", default=LEARNED_WEIGHTS,
                    help="network of the learned runs (kevernals_amd/weights/ name or a path)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run an elementwise copy of 256 MiB (known bytes) to scale HBM counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU sharding "
              "is a design in DESIGN.md, not a measured path", flush=True)
        sys.exit(2)

    import torch

    import kevernals_amd as K
    from kevernals_amd import geometry as G

    if args.workload == "batched":
        return bench_batched(args, torch, K, G)
    if args.workload == "learned":
        return bench_learned(args, torch, K, G)
    if args.workload == "cnn":
        return bench_cnn(args, torch, K, G)
    if args.workload == "learned-batched":
        return bench_learned_batched(args, torch, K, G)
    if args.workload == "io":
        return bench_io(args, torch, K, G)
    g = G.full_plain() if args.workload == "full" else G.small_plain()
    ctx = K.Context(0, g)
    ctx.use_torch_stream()
    mu, data = make_data(K, ctx, g, torch)
    kernel_pairs = ctx.count_pairs(K.MODE_R)
    fp_pairs, bp_pairs = kernel_pairs
    log(f"[bench] accepted pairs: forward {fp_pairs}, backprojector {bp_pairs}")
    oracle_pairs = nnz_fixture("full" if args.workload == "full" else "small")
    if oracle_pairs is not None:
        # the roofline numerators are the oracle's counts; the kernels' own
        # counts are reported beside them (they are equal)
        if (fp_pairs, bp_pairs) != (oracle_pairs["forward"], oracle_pairs["backprojector"]):
            log("[bench] WARNING: kernel pair counts differ from the oracle's")
        fp_pairs, bp_pairs = oracle_pairs["forward"], oracle_pairs["backprojector"]

    if args.copy_calibration:
        # reads 256 MiB and writes 256 MiB: the FETCH_SIZE / WRITE_SIZE of this
        # kernel calibrate the counters against known bytes
        a = torch.rand(64 << 20, device="cuda")
        b = torch.empty_like(a)
        for _ in range(2):
            torch.mul(a, 1.0, out=b)
        torch.cuda.synchronize()
        del a, b

    x = torch.ones(g.vol_shape, device="cuda")
    for _ in range(args.warmup):
        ctx.mlem_forward(x, data, 1.0, K.MODE_R)
        ctx.mlem_backward(x, K.MODE_R)
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for e in ev:
        e[0].record()
        ctx.mlem_forward(x, data, 1.0, K.MODE_R)
        e[1].record()
        ctx.mlem_backward(x, K.MODE_R)
        e[2].record()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_step = (t1 - t0) * 1e3 / args.steps
    fp_ms = float(np.mean([e[0].elapsed_time(e[1]) for e in ev]))
    bp_ms = float(np.mean([e[1].elapsed_time(e[2]) for e in ev]))
    log(f"[bench] step {ms_step:.3f} ms (forward {fp_ms:.3f} ms, backward {bp_ms:.3f} ms)")

    nr, nv = g.nrays, g.nvox
    bytes_fp = 4 * fp_pairs + 8 * nr  # voxel gathers + data read + ratio write
    bytes_bp = 4 * bp_pairs + 8 * nv  # ratio gathers + volume read + write
    if bp_ms >= fp_ms:
        kname, kms, kbytes = BP_KERNEL, bp_ms, bytes_bp
    else:
        kname, kms, kbytes = FP_KERNEL, fp_ms, bytes_fp
    achieved = kbytes / (kms * 1e-3) / 1e9
    value = nr / (ms_step * 1e-3) / 1e9

    loops = None
    if args.workload == "full" and not args.no_loops:
        loops = bench_sart_saa(K, ctx, g, data, torch)

    cpu = None
    if not args.no_cpu:
        log("[bench] CPU baseline (oracle) ...")
        cpu = cpu_baseline(g, mu.cpu().numpy(), data.cpu().numpy(), args.cpu_seconds)
        log(f"[bench] CPU baseline {cpu['value']:.3e} GRays/s on {cpu['cores']} threads")

    learned = None
    if args.workload == "full" and not args.no_learned:
        del mu, data, x
        ctx = None
        torch.cuda.empty_cache()
        res, _ = learned_batched_runs(torch, K, G, args.recons, args.cnn_weights)
        s1 = res["time_to_rmse"]["0.01"]
        learned = {
            "workload": (f"learned plus batched MLEM-A: {args.recons} reconstructions 1024x1024x64 / 25 views "
                         "in one batch, time until every reconstruction is within 1% relative RMSE of its "
                         "100-iteration MLEM-A result"),
            "plain_s": s1["plain_s"], "learned_s": s1["learned_s"], "schedule_only_s": s1["schedule_only_s"],
            "learned_speedup": s1["learned_speedup"], "schedule_only_speedup": s1["schedule_only_speedup"],
            **res,
        }

    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "GRays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded ellipsoid phantom + 200 microcalcifications, Poisson I0=5e3)",
        "config": {
            "workload": ("plain MLEM iteration, reference discretisation, fused forward+ratio / "
                         "backprojection+update" if args.workload == "full" else "small plain MLEM iteration"),
            "volume": [g.nx, g.ny, g.nz],
            "voxel_mm": [g.sx, g.sy, g.sz],
            "detector": [g.nu, g.nv],
            "pixel_mm": [g.pu, g.pv],
            "views": g.nviews,
            "arc_deg": 20.0 if args.workload == "full" else 15.0,
            "parallelism": "single GPU",
        },
        "mlem_iters_per_s": round(1e3 / ms_step, 3),
        "kernels_ms": {FP_KERNEL: round(fp_ms, 4), BP_KERNEL: round(bp_ms, 4)},
        "accepted_pairs": {"forward": fp_pairs, "backprojector": bp_pairs,
                           "source": "oracle (tests/golden/nnz/nnz.json)" if oracle_pairs else "kernel",
                           "kernel_count": {"forward": kernel_pairs[0], "backprojector": kernel_pairs[1]}},
        "roofline": {
            "bound": "hbm",
            "kernel": kname,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic_from_profiles(kname),
            "algorithmic_bytes": kbytes,
        },
        "cpu_baseline": cpu,
    }
    if loops is not None:
        out.update(loops)
    if learned is not None:
        out["time_to_rmse_1pct"] = learned
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
