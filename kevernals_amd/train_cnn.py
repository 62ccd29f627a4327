"""Train the residual CNN of the learned MLEM step on pairs the method makes
itself: 5-iteration MLEM (input) and 100-iteration MLEM (target) of seeded
phantoms, 256x256x32 voxels at 0.25x0.25x2 mm, 512^2 detector at 0.3 mm,
15 views on +-15 deg, transmission data at I0 = 1e4.  L2 loss, Adam.
--stages 5,10,20,40 adds the same run's later estimates as inputs (each
paired with its 100-iteration estimate).  --input krylov pairs the
trajectory-extrapolated estimate made at iteration --at (from the last
--history log-ratios, kevernals_amd.trajectory) with the 100-iteration
estimate instead: the network then corrects what the extrapolation misses.
--input schedule pairs the final estimate of the two-stretch jump schedule
(kevernals_amd.trajectory.Schedule: 18 MLEM iterations, two jumps) with the
100-iteration estimate.

    python -m kevernals_amd.train_cnn --pairs 2000 --epochs 4 --out weights.npz

Phantoms are rasterised by CPU worker processes started before the GPU is
touched; projection, Poisson sampling, MLEM and training run on the GPU
and all pairs stay resident in device memory (16 MiB each).
"""
from __future__ import annotations

import argparse
import multiprocessing as mp
import sys
import time

import numpy as np

TRAIN_SEED0 = 4000
VAL_SEED0 = 6000


GEOMETRY = "train"  # set by main before the workers fork


def _geometry():
    from kevernals_amd import geometry as G

    return G.full_plain() if GEOMETRY == "full" else G.learned_train()


def _phantom(args):
    from kevernals_amd.phantom import phantom

    # the full-size workload's phantoms carry 200 microcalcifications
    return phantom(_geometry(), args, microcalcifications=200 if GEOMETRY == "full" else 0)


def log(msg):
    print(msg, flush=True)


def make_krylov_pairs(ctx, seeds, pool, i0, at, history, k_ref):
    """(input, target) pairs: the trajectory extrapolation made at iteration
    `at` from the last `history` log-ratios, jumping to iteration k_ref,
    paired with the k_ref-iteration estimate of the same run."""
    import torch

    from kevernals_amd.learned import transmission_data
    from kevernals_amd.trajectory import Trajectory

    g = ctx.geometry
    tr = Trajectory(ctx, g.vol_shape, history)
    xs, ts = [], []
    t0 = time.time()
    for n, mu_h in enumerate(pool.imap(_phantom, seeds, chunksize=4)):
        mu = torch.from_numpy(mu_h).cuda()
        data = transmission_data(ctx, mu, i0, seeds[n])
        x = torch.ones(g.vol_shape, device="cuda")
        tr.reset()
        if at - history <= 0:
            tr.push(x)
        for k in range(1, k_ref + 1):
            ctx.mlem_iteration(x, data, 1.0, 1)
            if at - history <= k <= at:
                tr.push(x)
            if k == at:
                y = torch.empty_like(x)
                tr.extrapolate(k_ref - at, y)
                xs.append(y)
        ts.append(x)
        if (n + 1) % 50 == 0 or n + 1 == len(seeds):
            torch.cuda.synchronize()
            log(f"[train] pairs {n + 1}/{len(seeds)} in {time.time() - t0:.1f} s")
    return xs, ts


def make_schedule_pairs(ctx, seeds, pool, i0, k_ref, second=8):
    """(input, target) pairs: the final estimate of the jump schedule (jumps
    not gated; `second` probe iterations) and the k_ref-iteration estimate of
    the same data."""
    import torch

    from kevernals_amd.learned import transmission_data
    from kevernals_amd.trajectory import Jumper, Schedule

    g = ctx.geometry
    sch = Schedule(horizon=k_ref, second=second)
    xs, ts = [], []
    t0 = time.time()
    for n, mu_h in enumerate(pool.imap(_phantom, seeds, chunksize=4)):
        mu = torch.from_numpy(mu_h).cuda()
        data = transmission_data(ctx, mu, i0, seeds[n])
        jm = Jumper(ctx, g.vol_shape, sch)
        x = torch.ones(g.vol_shape, device="cuda")
        if sch.start == 0:
            jm.record(x)
        for k in range(1, k_ref + 1):
            ctx.mlem_iteration(x, data, 1.0, 1)
            if sch.start <= k <= sch.first:
                jm.record(x)
            if k == sch.first:
                xb = torch.empty_like(x)
                jm.first_jump(xb)
                jm.record(xb)
                for _ in range(sch.second):
                    ctx.mlem_iteration(xb, data, 1.0, 1)
                    jm.record(xb)
                jm.final_jump(xb)
                xs.append(xb)
                del jm
        ts.append(x)
        if (n + 1) % 50 == 0 or n + 1 == len(seeds):
            torch.cuda.synchronize()
            log(f"[train] pairs {n + 1}/{len(seeds)} in {time.time() - t0:.1f} s")
    return xs, ts


def make_pairs(ctx, seeds, pool, i0, stages, k_ref):
    """(input, target) pairs: the MLEM-A estimate after each iteration count
    in `stages`, paired with the estimate after k_ref iterations of the same
    run (the target tensor is shared by a phantom's pairs)."""
    import torch

    from kevernals_amd.learned import transmission_data

    g = ctx.geometry
    xs, ts = [], []
    t0 = time.time()
    for n, mu_h in enumerate(pool.imap(_phantom, seeds, chunksize=4)):
        mu = torch.from_numpy(mu_h).cuda()
        data = transmission_data(ctx, mu, i0, seeds[n])
        x = torch.ones(g.vol_shape, device="cuda")
        snaps = []
        done = 0
        for k in stages:
            for _ in range(k - done):
                ctx.mlem_iteration(x, data, 1.0, 1)
            done = k
            snaps.append(x.clone())
        for _ in range(k_ref - done):
            ctx.mlem_iteration(x, data, 1.0, 1)
        for sn in snaps:
            xs.append(sn)
            ts.append(x)
        if (n + 1) % 50 == 0 or n + 1 == len(seeds):
            torch.cuda.synchronize()
            log(f"[train] pairs {n + 1}/{len(seeds)} in {time.time() - t0:.1f} s")
    return xs, ts


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--pairs", type=int, default=2000)
    ap.add_argument("--val", type=int, default=16)
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--i0", type=float, default=1e4)
    ap.add_argument("--k-in", type=int, default=5)
    ap.add_argument("--stages", default=None,
                    help="comma-separated input iteration counts (default: --k-in only)")
    ap.add_argument("--k-ref", type=int, default=100)
    ap.add_argument("--input", choices=["mlem", "krylov", "schedule"], default="mlem")
    ap.add_argument("--at", type=int, default=20, help="--input krylov: iteration of the jump")
    ap.add_argument("--history", type=int, default=18, help="--input krylov: log-ratios used")
    ap.add_argument("--workers", type=int, default=8)
    ap.add_argument("--geometry", choices=["train", "full"], default="train",
                    help="full: the full-size workload (1024x1024x64, 25 views, 200 microcalcifications)")
    ap.add_argument("--second", type=int, default=8, help="--input schedule: probe iterations")
    ap.add_argument("--crop", default=None,
                    help="train on random crops nz,ny,nx of the pairs (e.g. 32,256,256)")
    ap.add_argument("--crops-per-pair", type=int, default=4)
    ap.add_argument("--lr-decay-epoch", type=int, default=0,
                    help="from this epoch on (1-based, 0: never) the learning rate is divided by 3")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default="cnn_weights.npz")
    args = ap.parse_args(argv)

    global GEOMETRY
    GEOMETRY = args.geometry
    # CPU workers first: no process is started after the GPU is initialised
    pool = mp.get_context("fork").Pool(args.workers)

    import torch

    import kevernals_amd as K
    from kevernals_amd import geometry as G
    from kevernals_amd.learned import Cnn

    g = _geometry()
    ctx = K.Context(0, g)
    t0 = time.time()
    stages = sorted({int(v) for v in args.stages.split(",")}) if args.stages else [args.k_in]
    assert 0 < stages[0] and stages[-1] < args.k_ref
    train_seeds = list(range(TRAIN_SEED0, TRAIN_SEED0 + args.pairs))
    val_seeds = list(range(VAL_SEED0, VAL_SEED0 + args.val))
    if args.input == "krylov":
        xs, ts = make_krylov_pairs(ctx, train_seeds, pool, args.i0, args.at, args.history, args.k_ref)
        vx, vt = make_krylov_pairs(ctx, val_seeds, pool, args.i0, args.at, args.history, args.k_ref)
        stages = [args.at]
    elif args.input == "schedule":
        xs, ts = make_schedule_pairs(ctx, train_seeds, pool, args.i0, args.k_ref, args.second)
        vx, vt = make_schedule_pairs(ctx, val_seeds, pool, args.i0, args.k_ref, args.second)
        stages = [0]
    else:
        xs, ts = make_pairs(ctx, train_seeds, pool, args.i0, stages, args.k_ref)
        vx, vt = make_pairs(ctx, val_seeds, pool, args.i0, [args.k_in], args.k_ref)
    pool.close()
    pool.join()
    t_data = time.time() - t0
    log(f"[train] {len(xs)} training pairs (stages {stages}) + {len(vx)} validation pairs in {t_data:.1f} s")

    cnn = Cnn(ctx, seed=args.seed)

    def val_loss():
        return float(np.mean([cnn.loss_and_grad(a, b) for a, b in zip(vx, vt)]))

    def identity_loss():
        s = cnn.scale
        return float(np.mean([float(((a - b).double() * s).pow(2).mean()) for a, b in zip(vx, vt)]))

    base = identity_loss()
    log(f"[train] validation loss of the input itself: {base:.5e}; untrained CNN {val_loss():.5e}")
    rng = np.random.Generator(np.random.PCG64(args.seed))
    history = []
    t1 = time.time()
    crop = [int(v) for v in args.crop.split(",")] if args.crop else None
    for ep in range(args.epochs):
        order = rng.permutation(len(xs))
        tl = []
        lr = args.lr / 3 if args.lr_decay_epoch and ep + 1 >= args.lr_decay_epoch else args.lr
        for i in order:
            if crop is None:
                tl.append(cnn.train_step(xs[i], ts[i], lr))
                continue
            for _ in range(args.crops_per_pair):
                z0, y0, x0 = (int(rng.integers(0, n - c + 1)) for n, c in zip(g.vol_shape, crop))
                sl = (slice(z0, z0 + crop[0]), slice(y0, y0 + crop[1]), slice(x0, x0 + crop[2]))
                tl.append(cnn.train_step(xs[i][sl].contiguous(), ts[i][sl].contiguous(), lr))
        vl = val_loss()
        history.append((ep + 1, float(np.mean(tl)), vl))
        log(f"[train] epoch {ep + 1}: train {np.mean(tl):.5e} val {vl:.5e} ({time.time() - t1:.1f} s)")
    steps_per_s = cnn.step / (time.time() - t1)
    cnn.save(args.out, history=np.array(history), identity_val=base, pairs=len(xs), phantoms=args.pairs,
             k_in=args.k_in, stages=np.array(stages), k_ref=args.k_ref, i0=args.i0, data_seconds=t_data,
             input_kind=args.input, at=args.at, krylov_history=args.history,
             steps_per_s=steps_per_s, geometry=args.geometry, second=args.second,
             crop=np.array(crop if crop else [0, 0, 0]), crops_per_pair=args.crops_per_pair)
    log(f"[train] saved {args.out}: {cnn.step} steps, {steps_per_s:.1f} steps/s")
    return 0


if __name__ == "__main__":
    sys.exit(main())
