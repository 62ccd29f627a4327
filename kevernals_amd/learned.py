"""Learned-accelerated MLEM: training pairs, CNN training and the
likelihood-gated interleaving of CNN steps with MLEM iterations; the
trajectory-jump schedule (kevernals_amd.trajectory) with the CNN applied to
its result.

Everything numerical runs on the GPU through the C ABI (kv_mlem_iteration
in mode A, kv_log_likelihood, kv_cnn_*); this module only sequences calls.
The reconstruction is standard MLEM on line-integral data,
x <- x / A^T 1 * A^T(data / Ax), the mode-A form of recons::MLEM
(KevLoq/Kevernals Reconstructors.h:182-242).  The CNN maps an early MLEM
estimate towards the 100-iteration one; a CNN step is kept only if it does
not lower the Poisson log-likelihood sum(data*log(Ax) - Ax), so plain MLEM
stays both the fallback and the arbiter.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import CNN_NPARAMS, MODE_A, Context, cnn_workspace_size

DEFAULT_SCALE = 20.0  # attenuation values (1/mm) are O(0.02-0.1): s*x is O(1)


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal weights (fan-in 27*cin), zero biases, last layer x0.1."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for li, (co, ci) in enumerate([(16, 1), (16, 16), (16, 16), (1, 16)]):
        w = rng.standard_normal((co, ci, 27)) * np.sqrt(2.0 / (27 * ci))
        if li == 3:
            w *= 0.1
        out += [w.ravel(), np.zeros(co)]
    p = np.concatenate(out).astype(np.float32)
    assert p.size == CNN_NPARAMS
    return p


class Cnn:
    """Parameters, Adam state and workspaces of the residual CNN on one
    device; all calls are enqueued on the context's (torch's current)
    stream."""

    def __init__(self, ctx: Context, params: Optional[np.ndarray] = None, scale: float = DEFAULT_SCALE,
                 seed: int = 0):
        import torch

        self.torch = torch
        self.ctx = ctx
        self.scale = float(scale)
        p = init_params(seed) if params is None else np.ascontiguousarray(params, dtype=np.float32)
        if p.size != CNN_NPARAMS:
            raise ValueError(f"expected {CNN_NPARAMS} parameters, got {p.size}")
        self.params = torch.from_numpy(p.copy()).cuda()
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.grad = torch.zeros_like(self.params)
        self.step = 0
        self._work = {}

    def _workspace(self, shape, training):
        key = (tuple(shape), bool(training))
        w = self._work.get(key)
        if w is None:
            w = self.torch.empty(cnn_workspace_size(shape, training), dtype=self.torch.uint8, device="cuda")
            self._work[key] = w
        return w

    def forward(self, x, y):
        """y = x + f(s*x)/s (device tensors of shape (nz, ny, nx))."""
        self.ctx.cnn_forward(self.params, self.scale, x, y, self._workspace(x.shape, False))

    def loss_and_grad(self, x, target) -> float:
        return self.ctx.cnn_gradient(self.params, self.scale, x, target, self._workspace(x.shape, True), self.grad)

    def train_step(self, x, target, lr: float = 1e-3) -> float:
        loss = self.loss_and_grad(x, target)
        self.step += 1
        self.ctx.cnn_adam(self.params, self.grad, self.m, self.v, self.step, lr)
        return loss

    def params_host(self) -> np.ndarray:
        return self.params.cpu().numpy()

    def save(self, path: str, **meta):
        np.savez(path, params=self.params_host(), scale=np.float32(self.scale), steps=np.int64(self.step),
                 **{k: np.asarray(v) for k, v in meta.items()})

    @classmethod
    def load(cls, ctx: Context, path: str) -> "Cnn":
        with np.load(path, allow_pickle=False) as f:
            c = cls(ctx, f["params"], float(f["scale"]))
            c.step = int(f["steps"])
        return c


def mlem_a(ctx: Context, x, data, iters: int):
    for _ in range(iters):
        ctx.mlem_iteration(x, data, 1.0, MODE_A)


def transmission_data(ctx: Context, mu, i0: float, seed: int):
    """Noisy line integrals -ln(clamp(I, 1, I0)/I0), I ~ Poisson(I0 exp(-A mu)),
    sampled on the device by the library's seeded generator
    (kv_poisson_data: Philox4x64-10 streams, inversion / PTRS).  Counts
    above I0 are clamped: the line integrals stay non-negative, as the
    multiplicative MLEM update requires (negative data drive estimates
    negative and the iteration unstable)."""
    torch = __import__("torch")
    g = ctx.geometry
    ell = torch.empty(g.proj_shape, device="cuda")
    ctx.forward_project(mu, ell, MODE_A)
    ctx.poisson_data(ell, ell, i0, seed)
    return ell


def rmse(x, ref) -> float:
    """||x - ref|| / ||ref|| over the volume (FP64 accumulation)."""
    d = (x.double() - ref.double()).norm()
    return float(d / ref.double().norm())


@dataclass
class GatedRun:
    x: object
    trace: List[tuple] = field(default_factory=list)  # (kind, iteration, seconds, rmse or None, loglik or None)
    accepted: int = 0
    rejected: int = 0


def gated_mlem(ctx: Context, cnn: Optional[Cnn], data, iters: int, every: int = 2, first: int = 2,
               reference=None, target_rmse: Optional[float] = None, x0=None,
               patience: Optional[int] = None) -> GatedRun:
    """MLEM-A from a uniform (or given) start with a CNN candidate after
    iteration `first` and every `every` iterations after it; a candidate
    replaces the estimate only if its log-likelihood is not lower.  After
    `patience` consecutive rejections no further candidates are made (None:
    never stop).  With a reference volume the RMSE after every step is
    traced; with target_rmse the run stops as soon as it is reached.  Wall
    times include the likelihood gate and are taken with the device
    synchronised (the RMSE evaluation itself is excluded)."""
    torch = __import__("torch")
    g = ctx.geometry
    x = torch.ones(g.vol_shape, device="cuda") if x0 is None else x0.clone()
    cand = torch.empty_like(x) if cnn is not None else None
    run = GatedRun(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    paused = 0.0

    def note(kind, k, ll=None):
        nonlocal paused
        torch.cuda.synchronize()
        t = time.perf_counter() - t0 - paused
        r = None
        if reference is not None:
            tp = time.perf_counter()
            r = rmse(x, reference)
            paused += time.perf_counter() - tp
        run.trace.append((kind, k, t, r, ll))
        return r

    misses = 0
    for k in range(1, iters + 1):
        ctx.mlem_iteration(x, data, 1.0, MODE_A)
        r = note("mlem", k)
        if target_rmse is not None and r is not None and r <= target_rmse:
            break
        active = cnn is not None and (patience is None or misses < patience)
        if active and k >= first and (k - first) % every == 0:
            cnn.forward(x, cand)
            ll_x, ll_c = ctx.log_likelihoods(data, [x, cand])
            if ll_c >= ll_x:
                x.copy_(cand)
                run.accepted += 1
                misses = 0
                r = note("cnn+", k, ll_c)
            else:
                run.rejected += 1
                misses += 1
                r = note("cnn-", k, ll_x)
            if target_rmse is not None and r is not None and r <= target_rmse:
                break
    run.x = x
    return run


def time_to_rmse(trace, target: float) -> Optional[float]:
    for kind, k, t, r, ll in trace:
        if r is not None and r <= target:
            return t
    return None


@dataclass
class BatchedRun:
    trace: List[tuple] = field(default_factory=list)  # (kind, iteration, seconds, [rmse per reconstruction])
    accepted: List[int] = field(default_factory=list)
    rejected: List[int] = field(default_factory=list)


def batched_gated_mlem(batch, iters: int, params: Optional[np.ndarray] = None, scale: float = DEFAULT_SCALE,
                       every: int = 2, first: int = 5, patience: Optional[int] = 2, references=None) -> BatchedRun:
    """MLEM-A on every reconstruction of a mode-A batch (kv_batch: one
    persistent forward and one persistent backprojection launch per
    iteration, replayed from a captured hipGraph), with a CNN candidate per
    reconstruction after iteration `first` and every `every` iterations after
    it, kept only if that reconstruction's Poisson log-likelihood does not
    decrease (params None: plain batched MLEM).  Volumes start from the
    batch's current contents.  With references (one device tensor per
    reconstruction) the RMSE of every reconstruction is traced after each
    batch step; wall times exclude those evaluations."""
    torch = __import__("torch")
    n = len(batch)
    ctxs = [batch.context(i) for i in range(n)]
    bufs = [batch.buffers(i) for i in range(n)]
    shapes = [g.vol_shape for g in batch.geometries]
    cnn = Cnn(ctxs[0], params, scale) if params is not None else None
    run = BatchedRun(accepted=[0] * n, rejected=[0] * n)
    misses = [0] * n
    views = [torch.empty(s, device="cuda") for s in shapes] if (references is not None or cnn is not None) else None
    cand = [torch.empty(s, device="cuda") for s in shapes] if cnn is not None else None

    def fetch(i):
        ctxs[i].copy(views[i], bufs[i][0], views[i].numel() * 4)

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    paused = 0.0

    def note(kind, k):
        nonlocal paused
        torch.cuda.synchronize()
        t = time.perf_counter() - t0 - paused
        r = None
        if references is not None:
            tp = time.perf_counter()
            r = []
            for i in range(n):
                fetch(i)
                r.append(rmse(views[i], references[i]))
            torch.cuda.synchronize()
            paused += time.perf_counter() - tp
        run.trace.append((kind, k, t, r))

    k = 0
    while k < iters:
        step = min(first if (cnn is not None and k < first) else every, iters - k) if cnn is not None else 1
        batch.mlem(step, use_graph=True)
        k += step
        note("mlem", k)
        if cnn is None or k < first:
            continue
        any_active = False
        for i in range(n):
            if patience is not None and misses[i] >= patience:
                continue
            any_active = True
            fetch(i)
            cnn.forward(views[i], cand[i])
            data_addr = bufs[i][1]
            ll_x, ll_c = ctxs[i].log_likelihoods(data_addr, [views[i], cand[i]])
            if ll_c >= ll_x:
                ctxs[i].copy(bufs[i][0], cand[i], cand[i].numel() * 4)
                run.accepted[i] += 1
                misses[i] = 0
            else:
                run.rejected[i] += 1
                misses[i] += 1
        torch.cuda.synchronize()  # the batch stream must see the accepted volumes
        if any_active:
            note("cnn", k)
    return run


def _chain_gate(ctx, data, x, y, y2=None):
    """The likelihood gate of the final steps: the jump x -> y is kept if
    log L(y) >= log L(x); then the CNN step y -> y2 if log L(y2) >= log L(y).
    Returns (steps kept: 0, 1 or 2, log-likelihood of the kept volume)."""
    # one ray walk for every candidate (values as separate evaluations)
    lls = ctx.log_likelihoods(data, [x, y] if y2 is None else [x, y, y2])
    ll_x, ll_y = lls[0], lls[1]
    if ll_y < ll_x:
        return 0, ll_x
    if y2 is None:
        return 1, ll_y
    ll_y2 = lls[2]
    return (2, ll_y2) if ll_y2 >= ll_y else (1, ll_y)


def accelerated_mlem(ctx: Context, data, schedule=None, cnn: Optional[Cnn] = None, iters: Optional[int] = None,
                     reference=None, target_rmse: Optional[float] = None) -> GatedRun:
    """MLEM-A from a uniform start with the two-stretch trajectory jump
    (trajectory.Schedule) and, with a network, a CNN step on the jump's
    result.  After `first` iterations the estimate x_first is held while a
    probe (first jump, `second` MLEM iterations) collects stretch B; the
    final jump from x_first then replaces the estimate only if its Poisson
    log-likelihood is not lower than x_first's, and the CNN step only if it
    does not lower the jump's.  A rejected jump leaves x_first, and plain
    MLEM continues from it.  `iters` MLEM iterations run in total (probe
    included; default: the schedule's cost, so the run ends with the final
    steps).  Trace kinds: mlem (an iteration of the estimate), probe (an
    iteration of the probe; the traced RMSE is the estimate's), final+ /
    final- (jump kept / rejected), cnn+ (the CNN step kept as well).  With a
    reference the RMSE of the estimate after every step is traced (excluded
    from the times); with target_rmse the run stops once it is reached."""
    from .trajectory import Jumper, Schedule

    torch = __import__("torch")
    sch = schedule or Schedule()
    iters = sch.cost if iters is None else iters
    g = ctx.geometry
    x = torch.ones(g.vol_shape, device="cuda")
    jm = Jumper(ctx, g.vol_shape, sch)
    cand = torch.empty_like(x)
    cand2 = torch.empty_like(x) if cnn is not None else None
    run = GatedRun(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    paused = 0.0

    def note(kind, k, est, ll=None):
        nonlocal paused
        torch.cuda.synchronize()
        t = time.perf_counter() - t0 - paused
        r = None
        if reference is not None:
            tp = time.perf_counter()
            r = rmse(est, reference)
            paused += time.perf_counter() - tp
        run.trace.append((kind, k, t, r, ll))
        return r is not None and target_rmse is not None and r <= target_rmse

    if sch.start == 0:
        jm.record(x)
    probing = False
    k = 0
    while k < iters:
        ctx.mlem_iteration(x, data, 1.0, MODE_A)
        k += 1
        if not probing and sch.start <= k <= sch.first and k < sch.cost:
            jm.record(x)
        if probing:
            jm.record(x)
            if k < sch.cost:
                note("probe", k, jm.estimate)
                continue
            # final steps: the estimate is x_first
            probing = False
            jm.final_jump(cand)
            if cnn is not None:
                cnn.forward(cand, cand2)
            kept, ll = _chain_gate(ctx, data, jm.estimate, cand, cand2)
            x.copy_((jm.estimate, cand, cand2)[kept])
            run.accepted += kept
            run.rejected += (2 if cnn is not None else 1) - kept
            if note(("final-", "final+", "cnn+")[kept], k, x, ll):
                break
            continue
        if note("mlem", k, x):
            break
        if k == sch.first and k < sch.cost and k < iters:
            # the probe: x becomes the jump to iteration `mid`
            jm.first_jump(cand)
            x.copy_(cand)
            jm.record(x)
            probing = True
    if probing:
        # stopped inside the probe: the estimate is still iteration `first`
        x.copy_(jm.estimate)
        note("mlem", k, x)
    run.x = x
    return run


def batched_accelerated_mlem(batch, schedule=None, params: Optional[np.ndarray] = None,
                             scale: float = DEFAULT_SCALE, iters: Optional[int] = None,
                             references=None) -> BatchedRun:
    """The two-stretch jump schedule on every reconstruction of a mode-A batch
    (as accelerated_mlem, reconstruction by reconstruction): the MLEM
    iterations of all reconstructions (estimates and probes alike) run as
    one persistent forward and one persistent backprojection launch per
    iteration (captured hipGraph); each reconstruction records its own
    stretches straight from the batch's volume buffers and makes its own
    jumps and gated final steps through its batch context.  Volumes start
    from the batch's current contents.  Trace entries: (kind, iteration,
    seconds, [RMSE of each reconstruction's estimate] or None)."""
    from .trajectory import Jumper, Schedule

    torch = __import__("torch")
    sch = schedule or Schedule()
    iters = sch.cost if iters is None else iters
    n = len(batch)
    ctxs = [batch.context(i) for i in range(n)]
    bufs = [batch.buffers(i) for i in range(n)]
    shapes = [g.vol_shape for g in batch.geometries]
    cnn = Cnn(ctxs[0], params, scale) if params is not None else None
    views = [torch.empty(s, device="cuda") for s in shapes]
    cand = [torch.empty(s, device="cuda") for s in shapes]
    cand2 = [torch.empty(s, device="cuda") for s in shapes] if cnn is not None else None
    jms = [Jumper(ctxs[i], shapes[i], sch) for i in range(n)]
    run = BatchedRun(accepted=[0] * n, rejected=[0] * n)

    def store(i, src):
        ctxs[i].copy(bufs[i][0], src, src.numel() * 4)

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    paused = 0.0

    def note(kind, k, probing):
        nonlocal paused
        torch.cuda.synchronize()
        t = time.perf_counter() - t0 - paused
        r = None
        if references is not None:
            tp = time.perf_counter()
            r = []
            for i in range(n):
                if probing:
                    r.append(rmse(jms[i].estimate, references[i]))
                else:
                    ctxs[i].copy(views[i], bufs[i][0], views[i].numel() * 4)
                    r.append(rmse(views[i], references[i]))
            torch.cuda.synchronize()
            paused += time.perf_counter() - tp
        run.trace.append((kind, k, t, r))

    if sch.start == 0:
        for i in range(n):
            jms[i].record(bufs[i][0])
    probing = False
    k = 0
    while k < iters:
        batch.mlem(1, use_graph=True)
        k += 1
        if probing or (sch.start <= k <= sch.first and k < sch.cost):
            # copies on the contexts' stream; note() synchronises the device
            # before the next batched launch writes these buffers
            for i in range(n):
                jms[i].record(bufs[i][0])
        if probing and k == sch.cost:
            probing = False
            for i in range(n):
                jms[i].final_jump(cand[i])
                if cnn is not None:
                    cnn.forward(cand[i], cand2[i])
                kept, _ = _chain_gate(ctxs[i], bufs[i][1], jms[i].estimate, cand[i],
                                      cand2[i] if cnn is not None else None)
                store(i, (jms[i].estimate, cand[i], cand2[i] if cnn is not None else None)[kept])
                run.accepted[i] += kept
                run.rejected[i] += (2 if cnn is not None else 1) - kept
            torch.cuda.synchronize()  # the batch stream must see the new volumes
            note("final", k, False)
            continue
        note("probe" if probing else "mlem", k, probing)
        if k == sch.first and k < sch.cost and k < iters:
            for i in range(n):
                jms[i].first_jump(cand[i])
                store(i, cand[i])
                jms[i].record(cand[i])
            torch.cuda.synchronize()
            probing = True
    if probing:
        # stopped inside the probe: every estimate is still iteration `first`
        for i in range(n):
            store(i, jms[i].estimate)
        torch.cuda.synchronize()
        note("mlem", k, False)
    return run
