// Accelerated mode-A MLEM: the Rayleigh-Ritz coefficient solve of the
// trajectory jump (host, FP64) and the two-stretch jump schedule with the
// likelihood-gated CNN step, sequenced over the library's own entry points.
//
// No counterpart in the reference: this accelerates recons::MLEM in its
// mode-A form (KevLoq/Kevernals source/modules/reconstruction/
// Reconstructors.h:182-242, x <- x / A^T 1 * A^T(data / Ax)).
//
// Jump.  Past the first iterations MLEM-A is close to a linear iteration of
// the log-ratios d_i = log(x_{i+1}/x_i), d_{i+1} = J d_i, with J self-adjoint
// in the inner product weighted by w = sens * x and its spectrum in [0, 1).
// From the weighted Gram matrix of recorded log-ratios (kv_traj_gram), the
// Rayleigh-Ritz projection T of J on the span B of the pairs' first members
// gives the remaining path p(J) J d_target ~ B Q p(T) T (coordinates of
// d_target), p(t) = sum_{i=1..N} t^i, in the W-orthonormal basis B Q.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "kv_internal.h"

namespace {

// Symmetric eigenproblem of a small FP64 matrix (n <= 32) by cyclic Jacobi
// rotations: a (row-major, n x n) is destroyed; ev receives the eigenvalues in
// ascending order, v (row-major) the eigenvectors as columns in that order.
// Jacobi converges to full relative accuracy on the small symmetric matrices
// here (the threshold sweeps stop when every off-diagonal entry is below
// 1e-300 or 2^-60 of its diagonal pair's geometric mean).
void jacobi_eigh(int n, std::vector<double> &a, std::vector<double> &ev, std::vector<double> &v)
{
    v.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i)
        v[(size_t)i * n + i] = 1.0;
    auto A = [&](int i, int j) -> double & { return a[(size_t)i * n + j]; };
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0;
        for (int i = 0; i < n; ++i)
            for (int j = i + 1; j < n; ++j)
                off = std::max(off, std::fabs(A(i, j)) / std::max(1e-300, std::sqrt(std::fabs(A(i, i) * A(j, j)))));
        if (off < 0x1p-60)
            break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = A(p, q);
                if (std::fabs(apq) < 1e-300)
                    continue;
                // rotation that zeroes A(p, q) (Golub & Van Loan, sym.schur2)
                const double tau = (A(q, q) - A(p, p)) / (2.0 * apq);
                const double t = (tau >= 0.0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
                const double c = 1.0 / std::sqrt(1.0 + t * t), s = t * c;
                for (int k = 0; k < n; ++k) {
                    const double akp = A(k, p), akq = A(k, q);
                    A(k, p) = c * akp - s * akq;
                    A(k, q) = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {
                    const double apk = A(p, k), aqk = A(q, k);
                    A(p, k) = c * apk - s * aqk;
                    A(q, k) = s * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    const double vkp = v[(size_t)k * n + p], vkq = v[(size_t)k * n + q];
                    v[(size_t)k * n + p] = c * vkp - s * vkq;
                    v[(size_t)k * n + q] = s * vkp + c * vkq;
                }
            }
    }
    // ascending order (stable on ties: index order)
    std::vector<int> ord(n);
    for (int i = 0; i < n; ++i)
        ord[i] = i;
    std::stable_sort(ord.begin(), ord.end(), [&](int i, int j) { return A(i, i) < A(j, j); });
    ev.resize(n);
    std::vector<double> vs((size_t)n * n);
    for (int c = 0; c < n; ++c) {
        ev[c] = A(ord[c], ord[c]);
        for (int k = 0; k < n; ++k)
            vs[(size_t)k * n + c] = v[(size_t)k * n + ord[c]];
    }
    v.swap(vs);
}

}  // namespace

kv_status kv_ritz_coefficients(const double *gram, int h, const int32_t *pairs, int npairs, int target, int horizon,
                               double rel_tol, double ritz_max, double *coef, double *ritz, int *nritz)
{
    if (!gram || !pairs || !coef || h < 2 || h > KV_TRAJ_MAX || npairs < 1 || npairs > h || horizon < 0) {
        kv_set_error("ritz coefficients: need 2 <= h <= %d, 1 <= npairs <= h, horizon >= 0", KV_TRAJ_MAX);
        return KV_ERR_INVALID;
    }
    std::vector<int> bi(npairs), bj(npairs);
    int ti = -1;
    for (int k = 0; k < npairs; ++k) {
        bi[k] = pairs[2 * k];
        bj[k] = pairs[2 * k + 1];
        if (bi[k] < 0 || bi[k] >= h || bj[k] < 0 || bj[k] >= h) {
            kv_set_error("ritz coefficients: pair %d indexes outside the %d x %d Gram matrix", k, h, h);
            return KV_ERR_INVALID;
        }
        if (bi[k] == target && ti < 0)
            ti = k;
    }
    if (ti < 0) {
        kv_set_error("ritz coefficients: the target must be the first member of a pair");
        return KV_ERR_INVALID;
    }
    const int m = npairs;
    std::fill(coef, coef + h, 0.0);
    if (nritz)
        *nritz = 0;
    // B^T W B and B^T W JB
    std::vector<double> gb((size_t)m * m), hb((size_t)m * m);
    for (int r = 0; r < m; ++r)
        for (int c = 0; c < m; ++c) {
            gb[(size_t)r * m + c] = gram[(size_t)bi[r] * h + bi[c]];
            hb[(size_t)r * m + c] = gram[(size_t)bi[r] * h + bj[c]];
        }
    std::vector<double> a = gb, ev, U;
    jacobi_eigh(m, a, ev, U);
    if (ev[m - 1] <= 0.0 || horizon == 0)
        return KV_OK;
    // W-orthonormal basis of the span: Q = U[:, keep] / sqrt(ev[keep])
    std::vector<int> keep;
    for (int c = 0; c < m; ++c)
        if (ev[c] > ev[m - 1] * rel_tol)
            keep.push_back(c);
    const int r = (int)keep.size();
    std::vector<double> Q((size_t)m * r);
    for (int k = 0; k < m; ++k)
        for (int c = 0; c < r; ++c)
            Q[(size_t)k * r + c] = U[(size_t)k * m + keep[c]] / std::sqrt(ev[keep[c]]);
    // T = Q^T H Q, symmetrised
    std::vector<double> HQ((size_t)m * r, 0.0), T((size_t)r * r, 0.0);
    for (int i = 0; i < m; ++i)
        for (int c = 0; c < r; ++c) {
            double s = 0.0;
            for (int k = 0; k < m; ++k)
                s += hb[(size_t)i * m + k] * Q[(size_t)k * r + c];
            HQ[(size_t)i * r + c] = s;
        }
    for (int i = 0; i < r; ++i)
        for (int j = 0; j < r; ++j) {
            double s = 0.0;
            for (int k = 0; k < m; ++k)
                s += Q[(size_t)k * r + i] * HQ[(size_t)k * r + j];
            T[(size_t)i * r + j] = s;
        }
    for (int i = 0; i < r; ++i)
        for (int j = i + 1; j < r; ++j) {
            const double s = 0.5 * (T[(size_t)i * r + j] + T[(size_t)j * r + i]);
            T[(size_t)i * r + j] = T[(size_t)j * r + i] = s;
        }
    std::vector<double> th, Y;
    jacobi_eigh(r, T, th, Y);
    std::vector<double> f(r);
    for (int i = 0; i < r; ++i) {
        const double t = std::min(std::max(th[i], 0.0), ritz_max);
        th[i] = t;
        const double p = t * (1.0 - std::pow(t, (double)horizon)) / (1.0 - t);
        f[i] = p * t;
    }
    // c = Q^T Gb[:, ti];  cb = Q (Y (f * (Y^T c)))
    std::vector<double> cq(r, 0.0), yc(r, 0.0), z(r, 0.0);
    for (int i = 0; i < r; ++i)
        for (int k = 0; k < m; ++k)
            cq[i] += Q[(size_t)k * r + i] * gb[(size_t)k * m + ti];
    for (int i = 0; i < r; ++i) {
        for (int k = 0; k < r; ++k)
            yc[i] += Y[(size_t)k * r + i] * cq[k];
        yc[i] *= f[i];
    }
    for (int i = 0; i < r; ++i)
        for (int k = 0; k < r; ++k)
            z[i] += Y[(size_t)i * r + k] * yc[k];
    for (int k = 0; k < m; ++k) {
        double s = 0.0;
        for (int i = 0; i < r; ++i)
            s += Q[(size_t)k * r + i] * z[i];
        coef[bi[k]] += s;
    }
    if (ritz)
        std::copy(th.begin(), th.end(), ritz);
    if (nritz)
        *nritz = r;
    return KV_OK;
}

namespace {

kv_status check_schedule(const kv_accel_schedule *s)
{
    if (!s) {
        kv_set_error("null schedule");
        return KV_ERR_INVALID;
    }
    const int na = s->first - s->start + 1;
    if (s->start < 0 || na < 4 || s->second < 2 || na + s->second > KV_TRAJ_MAX + 1 ||
        !(s->first < s->mid && s->mid < s->horizon)) {
        kv_set_error("schedule: need start >= 0, first - start >= 3, second >= 2, "
                     "first - start + 1 + second <= %d, first < mid < horizon",
                     KV_TRAJ_MAX + 1);
        return KV_ERR_INVALID;
    }
    return KV_OK;
}

// Coefficients of the probe jump (one stretch of na iterates, to `mid`).
kv_status first_jump_coef(const std::vector<double> &G, int h, int horizon, std::vector<double> &coef)
{
    std::vector<int32_t> pairs;
    for (int i = 0; i < h - 1; ++i) {
        pairs.push_back(i);
        pairs.push_back(i + 1);
    }
    coef.assign(h, 0.0);
    return kv_ritz_coefficients(G.data(), h, pairs.data(), h - 1, h - 2, horizon, 1e-11, 1.0 - 1e-5, coef.data(),
                                nullptr, nullptr);
}

// Coefficients of the final jump: stretch A (log-ratios 0..na-2), the probe
// jump (na-1, no J-pair), stretch B (na..h-1); the combination is rebased
// from the newest iterate to x_{na-1} (stretch A's last) by subtracting 1
// from the coefficients of the jump and of stretch B.
kv_status final_jump_coef(const std::vector<double> &G, int h, int na, int horizon, std::vector<double> &coef)
{
    std::vector<int32_t> pairs;
    for (int i = 0; i < na - 2; ++i) {
        pairs.push_back(i);
        pairs.push_back(i + 1);
    }
    for (int i = na; i < h - 1; ++i) {
        pairs.push_back(i);
        pairs.push_back(i + 1);
    }
    coef.assign(h, 0.0);
    if (kv_status st = kv_ritz_coefficients(G.data(), h, pairs.data(), (int)pairs.size() / 2, na - 3, horizon, 1e-11,
                                            1.0 - 1e-5, coef.data(), nullptr, nullptr))
        return st;
    for (int i = na - 1; i < h; ++i)
        coef[i] -= 1.0;
    return KV_OK;
}

struct DevBuf {
    void *p = nullptr;
    ~DevBuf()
    {
        if (p)
            (void)hipFree(p);
    }
};

// One reconstruction's recorded stretches, candidates and CNN workspace.
struct AccelState {
    int na = 0, nslots = 0, count = 0;
    size_t vox_bytes = 0;
    DevBuf slots, cands, work;
    size_t work_bytes = 0;
    float *slot(int i) const { return (float *)slots.p + (size_t)i * (vox_bytes / 4); }
    float *cand(int i) const { return (float *)cands.p + (size_t)i * (vox_bytes / 4); }
};

kv_status accel_alloc(kv_context *ctx, const kv_accel_schedule *s, bool cnn, AccelState &st)
{
    const KvGeom &g = ctx->g;
    st.na = s->first - s->start + 1;
    st.nslots = st.na + s->second + 1;
    st.vox_bytes = (size_t)kv_nvox(g) * 4;
    if (hipMalloc(&st.slots.p, st.vox_bytes * st.nslots) != hipSuccess ||
        hipMalloc(&st.cands.p, st.vox_bytes * 2) != hipSuccess) {
        kv_set_error("accelerated MLEM: cannot allocate %d volumes", st.nslots + 2);
        return KV_ERR_NOMEM;
    }
    if (cnn) {
        if (kv_status e = kv_cnn_workspace_size(g.nx, g.ny, g.nz, 0, &st.work_bytes))
            return e;
        if (hipMalloc(&st.work.p, st.work_bytes) != hipSuccess) {
            kv_set_error("accelerated MLEM: cannot allocate the CNN workspace");
            return KV_ERR_NOMEM;
        }
    }
    return KV_OK;
}

kv_status accel_record(kv_context *ctx, AccelState &st, const float *x)
{
    if (st.count >= st.nslots) {
        kv_set_error("accelerated MLEM: both stretches are recorded already");
        return KV_ERR_STATE;
    }
    return kv_memcpy_d2d(ctx, st.slot(st.count++), x, st.vox_bytes);
}

// x <- the probe's start: the jump from stretch A to an estimate of `mid`.
kv_status accel_first_jump(kv_context *ctx, const kv_accel_schedule *s, AccelState &st, float *out)
{
    std::vector<const float *> xs(st.na);
    for (int i = 0; i < st.na; ++i)
        xs[i] = st.slot(i);
    const int h = st.na - 1;
    std::vector<double> G((size_t)h * h), coef;
    if (kv_status e = kv_traj_gram(ctx, xs.data(), h, G.data()))
        return e;
    if (kv_status e = first_jump_coef(G, h, s->mid - s->first, coef))
        return e;
    return kv_traj_combine(ctx, xs.data(), h, coef.data(), out);
}

kv_status accel_final_jump(kv_context *ctx, const kv_accel_schedule *s, AccelState &st, float *out)
{
    std::vector<const float *> xs(st.nslots);
    for (int i = 0; i < st.nslots; ++i)
        xs[i] = st.slot(i);
    const int h = st.nslots - 1;
    std::vector<double> G((size_t)h * h), coef;
    if (kv_status e = kv_traj_gram(ctx, xs.data(), h, G.data()))
        return e;
    if (kv_status e = final_jump_coef(G, h, st.na, s->horizon - s->first, coef))
        return e;
    return kv_traj_combine(ctx, xs.data(), h, coef.data(), out);
}

// The final gate: the jump x_first -> y is kept if log L(y) >= log L(x_first),
// then the CNN step y -> y2 if log L(y2) >= log L(y); the kept volume is
// copied to d_vol.  *kept = 0, 1 or 2 steps.
kv_status accel_final_steps(kv_context *ctx, const kv_accel_schedule *s, AccelState &st, const float *d_data,
                            const float *d_params, float scale, float *d_vol, int *kept)
{
    const KvGeom &g = ctx->g;
    float *y = st.cand(0), *y2 = st.cand(1);
    if (kv_status e = accel_final_jump(ctx, s, st, y))
        return e;
    if (d_params) {
        if (kv_status e = kv_cnn_forward(ctx, d_params, scale, y, y2, g.nx, g.ny, g.nz, st.work.p, st.work_bytes))
            return e;
    }
    const float *est = st.slot(st.na - 1);
    const float *vols[3] = {est, y, y2};
    double ll[3] = {0, 0, 0};
    if (kv_status e = kv_log_likelihoods(ctx, d_data, vols, d_params ? 3 : 2, ll))
        return e;
    int k = 0;
    if (ll[1] >= ll[0]) {
        k = 1;
        if (d_params && ll[2] >= ll[1])
            k = 2;
    }
    *kept = k;
    return kv_memcpy_d2d(ctx, d_vol, vols[k], st.vox_bytes);
}

}  // namespace

kv_status kv_mlem_accelerated(kv_context *ctx, float *d_vol, const float *d_data, const kv_accel_schedule *schedule,
                              const float *d_params, float scale, int *kept)
{
    if (!ctx || !ctx->has_geom) {
        kv_set_error(ctx ? "no geometry set" : "null context");
        return ctx ? KV_ERR_STATE : KV_ERR_INVALID;
    }
    if (kv_status e = check_schedule(schedule))
        return e;
    if (!d_vol || !d_data) {
        kv_set_error("null buffer");
        return KV_ERR_INVALID;
    }
    if (d_params && !(scale > 0.f)) {
        kv_set_error("CNN scale must be positive");
        return KV_ERR_INVALID;
    }
    const kv_accel_schedule *s = schedule;
    AccelState st;
    if (kv_status e = accel_alloc(ctx, s, d_params != nullptr, st))
        return e;
    const int cost = s->first + s->second;
    int steps = 0;
    if (s->start == 0)
        if (kv_status e = accel_record(ctx, st, d_vol))
            return e;
    bool probing = false;
    for (int k = 1; k <= cost; ++k) {
        if (kv_status e = kv_mlem_iteration(ctx, d_vol, d_data, 1.0f, KV_MODE_A))
            return e;
        if (!probing && s->start <= k && k <= s->first && k < cost)
            if (kv_status e = accel_record(ctx, st, d_vol))
                return e;
        if (probing) {
            if (kv_status e = accel_record(ctx, st, d_vol))
                return e;
            if (k < cost)
                continue;
            probing = false;
            if (kv_status e = accel_final_steps(ctx, s, st, d_data, d_params, scale, d_vol, &steps))
                return e;
            continue;
        }
        if (k == s->first && k < cost) {
            // the probe: the volume becomes the jump to iteration `mid`
            if (kv_status e = accel_first_jump(ctx, s, st, d_vol))
                return e;
            if (kv_status e = accel_record(ctx, st, d_vol))
                return e;
            probing = true;
        }
    }
    if (kept)
        *kept = steps;
    return kv_synchronize(ctx);
}

kv_status kv_batch_mlem_accelerated(kv_batch *b, const kv_accel_schedule *schedule, const float *d_params,
                                    float scale, int32_t *kept)
{
    if (!b) {
        kv_set_error("null batch");
        return KV_ERR_INVALID;
    }
    if (kv_status e = check_schedule(schedule))
        return e;
    if (d_params && !(scale > 0.f)) {
        kv_set_error("CNN scale must be positive");
        return KV_ERR_INVALID;
    }
    const kv_accel_schedule *s = schedule;
    const int n = kv_internal_batch_size(b);
    if (n < 1 || kv_internal_batch_mode(b) != KV_MODE_A) {
        kv_set_error("accelerated MLEM needs a mode-A batch");
        return KV_ERR_STATE;
    }
    std::vector<kv_context *> ctxs(n);
    std::vector<float *> vol(n), data(n);
    std::vector<AccelState> st(n);
    for (int i = 0; i < n; ++i) {
        if (kv_status e = kv_batch_context(b, i, &ctxs[i]))
            return e;
        if (kv_status e = kv_batch_buffers(b, i, &vol[i], &data[i]))
            return e;
        if (kv_status e = accel_alloc(ctxs[i], s, d_params != nullptr, st[i]))
            return e;
    }
    auto sync_all = [&]() -> kv_status {
        for (int i = 0; i < n; ++i)
            if (kv_status e = kv_synchronize(ctxs[i]))
                return e;
        return KV_OK;
    };
    const int cost = s->first + s->second;
    const std::vector<int32_t> one(n, 1);
    if (s->start == 0)
        for (int i = 0; i < n; ++i)
            if (kv_status e = accel_record(ctxs[i], st[i], vol[i]))
                return e;
    bool probing = false;
    for (int k = 1; k <= cost; ++k) {
        if (kv_status e = kv_batch_mlem(b, one.data(), 1.0f, 1, nullptr))
            return e;
        if (probing || (s->start <= k && k <= s->first && k < cost))
            for (int i = 0; i < n; ++i)
                if (kv_status e = accel_record(ctxs[i], st[i], vol[i]))
                    return e;
        if (probing && k == cost) {
            probing = false;
            for (int i = 0; i < n; ++i) {
                int steps = 0;
                if (kv_status e =
                        accel_final_steps(ctxs[i], s, st[i], data[i], d_params, scale, vol[i], &steps))
                    return e;
                if (kept)
                    kept[i] = steps;
            }
        } else if (!probing && k == s->first && k < cost) {
            for (int i = 0; i < n; ++i) {
                float *c = st[i].cand(0);
                if (kv_status e = accel_first_jump(ctxs[i], s, st[i], c))
                    return e;
                if (kv_status e = kv_memcpy_d2d(ctxs[i], vol[i], c, st[i].vox_bytes))
                    return e;
                if (kv_status e = accel_record(ctxs[i], st[i], c))
                    return e;
            }
            probing = true;
        }
        // the next batched launch reads the volumes the contexts wrote
        if (kv_status e = sync_all())
            return e;
    }
    return KV_OK;
}
