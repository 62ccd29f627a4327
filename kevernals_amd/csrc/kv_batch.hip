// Batched reconstructions: persistent launches over uneven problems,
// hipGraph-captured iterations and the tile-shape autotuner.
//
// Each reconstruction keeps its own context (geometry, crossing tables,
// ratio buffer) and volume/data buffers; a device array of KvProblem
// descriptors tells the persistent kernels where every reconstruction's
// forward tiles and backprojection bricks sit in the batch-wide work list.
// One iteration of the whole batch is two launches: the forward pass (ratio
// epilogue) for every reconstruction, then the backprojection (update
// epilogue).  Work items are claimed from counters (one per XCD list for the
// forward pass, one for the backprojection); the counters
// are cleared by a small kernel (not hipMemsetAsync, which misbehaves inside
// captured graphs on replay), so the same sequence runs eagerly or replayed.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "kv_internal.h"

struct kv_batch {
    int device = 0;
    int mode = KV_MODE_R;
    hipStream_t stream = nullptr;
    int n = 0;
    std::vector<kv_context *> ctx;
    std::vector<float *> vol, data;
    std::vector<KvProblem> probs;
    KvProblem *d_probs = nullptr;
    unsigned long long *d_counters = nullptr;
    int counters_cap = 0;
    long long fp_total = 0, bp_total = 0;
    int fp_grid = 0, bp_grid = 0;
    // captured iteration sequence
    hipGraphExec_t exec = nullptr;
    int exec_iters = -1;
    float exec_relax = 0.f;
    long long exec_fp_total = -1, exec_bp_total = -1;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

namespace {

kv_status fail_hip(hipError_t e, const char *what)
{
    kv_set_error("%s failed: %s", what, hipGetErrorString(e));
    return KV_ERR_DEVICE;
}

// Recomputes tile counts and the batch-wide item ranges.
void layout_items(kv_batch *b)
{
    long long fp = 0, bp = 0;
    for (auto &P : b->probs) {
        const int L = P.fp_layout, wu = 64 / L, span = KV_XCDS * std::max(1, KV_FP_STRIPE / wu);
        P.fp_chunk = span / KV_XCDS;
        P.fp_tx = ((P.g.nu + wu - 1) / wu + span - 1) / span * span;
        P.fp_ty = (P.g.nv + 4 * L - 1) / (4 * L);
        P.fp_nch = P.fp_tx / span;
        P.fp_dchunk = kv_udiv((unsigned)P.fp_chunk);
        P.fp_dnch = kv_udiv((unsigned)P.fp_nch);
        P.fp_dty = kv_udiv((unsigned)P.fp_ty);
        P.fp_first = fp;
        fp += (long long)P.g.np * P.fp_ty * (P.fp_tx / KV_XCDS);
        P.bp_bx = (P.g.nx + 63) / 64;
        P.bp_by = (P.g.ny + 3) / 4;
        P.bp_bz = (P.g.nz + P.bp_zt - 1) / P.bp_zt;
        P.bp_first = bp;
        bp += (long long)P.bp_bx * P.bp_by * P.bp_bz;
    }
    b->fp_total = fp;
    b->bp_total = bp;
}

kv_status upload_probs(kv_batch *b)
{
    const hipError_t e = hipMemcpyAsync(b->d_probs, b->probs.data(), b->probs.size() * sizeof(KvProblem),
                                        hipMemcpyHostToDevice, b->stream);
    return e == hipSuccess ? KV_OK : fail_hip(e, "upload of batch descriptors");
}

// per iteration: KV_XCDS forward-list counters, one backprojection counter
constexpr int kCountersPerIter = KV_XCDS + 1;

kv_status ensure_counters(kv_batch *b, int iters)
{
    const int need = kCountersPerIter * std::max(iters, 1);
    if (need <= b->counters_cap)
        return KV_OK;
    if (b->d_counters)
        (void)hipFree(b->d_counters);
    b->d_counters = nullptr;
    b->counters_cap = 0;
    const hipError_t e = hipMalloc((void **)&b->d_counters, (size_t)need * sizeof(unsigned long long));
    if (e != hipSuccess)
        return e == hipErrorOutOfMemory ? (kv_set_error("out of device memory"), KV_ERR_NOMEM)
                                        : fail_hip(e, "counter allocation");
    b->counters_cap = need;
    // the graph refers to the old counters
    if (b->exec) {
        (void)hipGraphExecDestroy(b->exec);
        b->exec = nullptr;
    }
    return KV_OK;
}

// The iteration sequence on the batch stream (eagerly or under capture).
hipError_t enqueue_iterations(kv_batch *b, int iters, float relax)
{
    hipError_t e = kv_launch_zero_u64(b->d_counters, kCountersPerIter * iters, b->stream);
    for (int k = 0; k < iters && e == hipSuccess; ++k) {
        unsigned long long *c = b->d_counters + (size_t)kCountersPerIter * k;
        e = kv_launch_fp_batch(b->mode, b->d_probs, b->n, b->fp_total, c, k, relax, b->fp_grid, b->stream);
        if (e == hipSuccess)
            e = kv_launch_bp_batch(b->mode, b->d_probs, b->n, b->bp_total, c + KV_XCDS, k, b->bp_grid, b->stream);
    }
    return e;
}

kv_status run(kv_batch *b, int iters, float relax, bool use_graph, float *ms)
{
    if (kv_status st = ensure_counters(b, iters))
        return st;
    if (kv_status st = upload_probs(b))
        return st;
    hipError_t e = hipEventRecord(b->ev0, b->stream);
    if (e != hipSuccess)
        return fail_hip(e, "event record");
    if (use_graph) {
        if (!b->exec || b->exec_iters != iters || b->exec_relax != relax || b->exec_fp_total != b->fp_total ||
            b->exec_bp_total != b->bp_total) {
            if (b->exec)
                (void)hipGraphExecDestroy(b->exec);
            b->exec = nullptr;
            hipGraph_t graph = nullptr;
            // descriptors must be on the device before capture begins
            if ((e = hipStreamSynchronize(b->stream)) != hipSuccess)
                return fail_hip(e, "stream synchronize");
            if ((e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal)) != hipSuccess)
                return fail_hip(e, "hipStreamBeginCapture");
            const hipError_t ee = enqueue_iterations(b, iters, relax);
            e = hipStreamEndCapture(b->stream, &graph);
            if (ee != hipSuccess || e != hipSuccess) {
                if (graph)
                    (void)hipGraphDestroy(graph);
                return fail_hip(ee != hipSuccess ? ee : e, "graph capture");
            }
            e = hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (e != hipSuccess) {
                b->exec = nullptr;
                return fail_hip(e, "hipGraphInstantiate");
            }
            b->exec_iters = iters;
            b->exec_relax = relax;
            b->exec_fp_total = b->fp_total;
            b->exec_bp_total = b->bp_total;
            if ((e = hipEventRecord(b->ev0, b->stream)) != hipSuccess)
                return fail_hip(e, "event record");
        }
        e = hipGraphLaunch(b->exec, b->stream);
    } else {
        e = enqueue_iterations(b, iters, relax);
    }
    if (e != hipSuccess)
        return fail_hip(e, "batch launch");
    if ((e = hipEventRecord(b->ev1, b->stream)) != hipSuccess || (e = hipStreamSynchronize(b->stream)) != hipSuccess)
        return fail_hip(e, "batch synchronize");
    if (ms) {
        float t = 0.f;
        (void)hipEventElapsedTime(&t, b->ev0, b->ev1);
        *ms = t;
    }
    return KV_OK;
}

kv_status check_index(kv_batch *b, int i)
{
    if (!b) {
        kv_set_error("null batch");
        return KV_ERR_INVALID;
    }
    if (i < 0 || i >= b->n) {
        kv_set_error("reconstruction index %d out of range (%d)", i, b->n);
        return KV_ERR_INVALID;
    }
    if (hipSetDevice(b->device) != hipSuccess) {
        kv_set_error("hipSetDevice failed");
        return KV_ERR_DEVICE;
    }
    return KV_OK;
}

int size_class(const KvProblem &P)
{
    const long long v = (long long)P.g.nx * P.g.ny * P.g.nz;
    return v < (16LL << 20) ? 0 : (v < (40LL << 20) ? 1 : 2);
}

void release(kv_batch *b)
{
    if (b->exec)
        (void)hipGraphExecDestroy(b->exec);
    for (float *p : b->vol)
        if (p)
            (void)hipFree(p);
    for (float *p : b->data)
        if (p)
            (void)hipFree(p);
    for (kv_context *c : b->ctx)
        if (c)
            kv_destroy(c);
    if (b->d_probs)
        (void)hipFree(b->d_probs);
    if (b->d_counters)
        (void)hipFree(b->d_counters);
    if (b->ev0)
        (void)hipEventDestroy(b->ev0);
    if (b->ev1)
        (void)hipEventDestroy(b->ev1);
    if (b->stream)
        (void)hipStreamDestroy(b->stream);
}

}  // namespace

int kv_internal_batch_size(const kv_batch *b) { return b ? b->n : 0; }
int kv_internal_batch_mode(const kv_batch *b) { return b ? b->mode : -1; }

extern "C" {

kv_status kv_batch_create(int device, int n, const kv_geometry *geoms, int mode, kv_batch **out)
{
    if (!out || !geoms || n <= 0 || (mode != KV_MODE_R && mode != KV_MODE_A)) {
        kv_set_error("bad batch arguments");
        return KV_ERR_INVALID;
    }
    *out = nullptr;
    auto *b = new kv_batch();
    b->device = device;
    b->n = n;
    b->mode = mode;
    kv_status st = KV_OK;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess)
        e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e == hipSuccess)
        e = hipEventCreate(&b->ev0);
    if (e == hipSuccess)
        e = hipEventCreate(&b->ev1);
    if (e != hipSuccess) {
        st = fail_hip(e, "batch stream/event creation");
        release(b);
        delete b;
        return st;
    }
    b->ctx.assign((size_t)n, nullptr);
    b->vol.assign((size_t)n, nullptr);
    b->data.assign((size_t)n, nullptr);
    b->probs.resize((size_t)n);
    for (int i = 0; i < n && st == KV_OK; ++i) {
        if ((st = kv_create(device, &b->ctx[(size_t)i])))
            break;
        if ((st = kv_set_geometry(b->ctx[(size_t)i], &geoms[i])))
            break;
        kv_context *c = b->ctx[(size_t)i];
        if ((st = kv_internal_ensure_tables(c, mode)))
            break;
        if (mode == KV_MODE_A && (st = kv_internal_ensure_sens(c)))
            break;
        const size_t nv = (size_t)kv_nvox(c->g), nr = (size_t)kv_nray(c->g);
        if ((e = hipMalloc((void **)&b->vol[(size_t)i], nv * 4)) != hipSuccess ||
            (e = hipMalloc((void **)&b->data[(size_t)i], nr * 4)) != hipSuccess) {
            st = e == hipErrorOutOfMemory ? (kv_set_error("out of device memory"), KV_ERR_NOMEM)
                                          : fail_hip(e, "batch buffers");
            break;
        }
        if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) {
            st = fail_hip(e, "table construction");
            break;
        }
        KvProblem &P = b->probs[(size_t)i];
        P.g = c->g;
        P.t = mode == KV_MODE_A ? c->tab_a : c->tab_r;
        P.vol = b->vol[(size_t)i];
        P.data = b->data[(size_t)i];
        P.ratio = c->d_ratio;
        P.sens = mode == KV_MODE_A ? c->d_sens : nullptr;
        P.iters = 0;
        P.fp_layout = mode == KV_MODE_A ? KV_FP_BATCH_LAYOUT_A : KV_FP_BATCH_LAYOUT_R;
        P.bp_zt = 8;
    }
    if (st == KV_OK) {
        layout_items(b);
        e = hipMalloc((void **)&b->d_probs, (size_t)n * sizeof(KvProblem));
        if (e != hipSuccess)
            st = fail_hip(e, "descriptor allocation");
    }
    if (st == KV_OK) {
        int cus = 0;
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
        cus = cus > 0 ? cus : 1;
        b->fp_grid = cus * kv_fp_batch_blocks_per_cu(mode);
        b->bp_grid = cus * kv_bp_batch_blocks_per_cu(mode);
    }
    if (st != KV_OK) {
        release(b);
        delete b;
        return st;
    }
    *out = b;
    return KV_OK;
}

kv_status kv_batch_destroy(kv_batch *b)
{
    if (!b)
        return KV_OK;
    (void)hipSetDevice(b->device);
    if (b->stream)
        (void)hipStreamSynchronize(b->stream);
    release(b);
    delete b;
    return KV_OK;
}

kv_status kv_batch_set_volume(kv_batch *b, int i, const float *h_vol)
{
    if (kv_status st = check_index(b, i))
        return st;
    const size_t bytes = (size_t)kv_nvox(b->probs[(size_t)i].g) * 4;
    hipError_t e = hipMemcpyAsync(b->vol[(size_t)i], h_vol, bytes, hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(b->stream);
    return e == hipSuccess ? KV_OK : fail_hip(e, "volume upload");
}

kv_status kv_batch_set_data(kv_batch *b, int i, const float *h_data)
{
    if (kv_status st = check_index(b, i))
        return st;
    const size_t bytes = (size_t)kv_nray(b->probs[(size_t)i].g) * 4;
    hipError_t e = hipMemcpyAsync(b->data[(size_t)i], h_data, bytes, hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(b->stream);
    return e == hipSuccess ? KV_OK : fail_hip(e, "data upload");
}

kv_status kv_batch_get_volume(kv_batch *b, int i, float *h_vol)
{
    if (kv_status st = check_index(b, i))
        return st;
    const size_t bytes = (size_t)kv_nvox(b->probs[(size_t)i].g) * 4;
    hipError_t e = hipMemcpyAsync(h_vol, b->vol[(size_t)i], bytes, hipMemcpyDeviceToHost, b->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(b->stream);
    return e == hipSuccess ? KV_OK : fail_hip(e, "volume download");
}

kv_status kv_batch_mlem(kv_batch *b, const int32_t *iters, float relax, int use_graph, float *ms)
{
    if (kv_status st = check_index(b, 0))
        return st;
    if (!iters) {
        kv_set_error("null iteration counts");
        return KV_ERR_INVALID;
    }
    int k = 0;
    for (int i = 0; i < b->n; ++i) {
        if (iters[i] < 0 || iters[i] > 100000) {
            kv_set_error("iteration count %d out of range", iters[i]);
            return KV_ERR_INVALID;
        }
        b->probs[(size_t)i].iters = iters[i];
        k = std::max(k, (int)iters[i]);
    }
    if (k == 0) {
        if (ms)
            *ms = 0.f;
        return KV_OK;
    }
    return run(b, k, relax, use_graph != 0, ms);
}

kv_status kv_batch_context(kv_batch *b, int i, kv_context **ctx)
{
    if (kv_status st = check_index(b, i))
        return st;
    if (!ctx) {
        kv_set_error("null output");
        return KV_ERR_INVALID;
    }
    *ctx = b->ctx[(size_t)i];
    return KV_OK;
}

kv_status kv_batch_buffers(kv_batch *b, int i, float **d_vol, float **d_data)
{
    if (kv_status st = check_index(b, i))
        return st;
    if (d_vol)
        *d_vol = b->vol[(size_t)i];
    if (d_data)
        *d_data = b->data[(size_t)i];
    return KV_OK;
}

kv_status kv_batch_set_tiles(kv_batch *b, int i, int fp_layout, int bp_slices)
{
    if (kv_status st = check_index(b, i))
        return st;
    if (!(fp_layout == 1 || fp_layout == 2 || fp_layout == 4 || fp_layout == 8) ||
        !(bp_slices == 4 || bp_slices == 8)) {
        kv_set_error("unsupported tile shape (layout %d, slices %d)", fp_layout, bp_slices);
        return KV_ERR_INVALID;
    }
    b->probs[(size_t)i].fp_layout = fp_layout;
    b->probs[(size_t)i].bp_zt = bp_slices;
    layout_items(b);
    return KV_OK;
}

kv_status kv_batch_get_tiles(kv_batch *b, int i, int *fp_layout, int *bp_slices)
{
    if (kv_status st = check_index(b, i))
        return st;
    if (fp_layout)
        *fp_layout = b->probs[(size_t)i].fp_layout;
    if (bp_slices)
        *bp_slices = b->probs[(size_t)i].bp_zt;
    return KV_OK;
}

kv_status kv_batch_autotune(kv_batch *b, const char *cache_path)
{
    if (kv_status st = check_index(b, 0))
        return st;
    hipDeviceProp_t prop{};
    (void)hipGetDeviceProperties(&prop, b->device);
    const std::string dev = prop.gcnArchName[0] ? prop.gcnArchName : "gpu";
    const std::string devkey = dev.substr(0, dev.find(':'));
    // cached choices: "<device> <class> <layout> <slices> <ms>" per line
    int best_l[3] = {0, 0, 0}, best_z[3] = {0, 0, 0};
    float best_ms[3] = {0.f, 0.f, 0.f};
    std::vector<std::string> kept;
    if (cache_path && *cache_path) {
        std::ifstream in(cache_path);
        std::string line;
        while (std::getline(in, line)) {
            std::istringstream ls(line);
            std::string d;
            int c, l, z;
            float t;
            if (!(ls >> d >> c >> l >> z >> t))
                continue;
            if (d == devkey && c >= 0 && c < 3 && (l == 1 || l == 2 || l == 4 || l == 8) && (z == 4 || z == 8)) {
                best_l[c] = l;
                best_z[c] = z;
                best_ms[c] = t;
            } else {
                kept.push_back(line);
            }
        }
    }
    bool present[3] = {false, false, false};
    for (const auto &P : b->probs)
        present[size_class(P)] = true;
    bool need = false;
    for (int c = 0; c < 3; ++c)
        need = need || (present[c] && best_l[c] == 0);
    hipError_t e = hipSuccess;
    std::vector<float *> backup((size_t)b->n, nullptr);
    kv_status st = KV_OK;
    if (need) {
        // keep the volumes: tuning runs real iterations
        for (int i = 0; i < b->n && e == hipSuccess; ++i) {
            const size_t bytes = (size_t)kv_nvox(b->probs[(size_t)i].g) * 4;
            e = hipMalloc((void **)&backup[(size_t)i], bytes);
            if (e == hipSuccess)
                e = hipMemcpyAsync(backup[(size_t)i], b->vol[(size_t)i], bytes, hipMemcpyDeviceToDevice, b->stream);
        }
        const int layouts[4] = {1, 2, 4, 8}, slices[2] = {4, 8};
        std::vector<int32_t> it((size_t)b->n);
        for (int c = 0; c < 3 && e == hipSuccess && st == KV_OK; ++c) {
            if (!present[c] || best_l[c] != 0)
                continue;
            for (int i = 0; i < b->n; ++i)
                it[(size_t)i] = size_class(b->probs[(size_t)i]) == c ? 1 : 0;
            float best = 1e30f;
            for (int l : layouts)
                for (int z : slices) {
                    for (auto &P : b->probs)
                        if (size_class(P) == c) {
                            P.fp_layout = l;
                            P.bp_zt = z;
                        }
                    layout_items(b);
                    float t = 1e30f;
                    for (int rep = 0; rep < 2 && st == KV_OK; ++rep) {
                        float ms = 0.f;
                        st = kv_batch_mlem(b, it.data(), 1.0f, 0, &ms);
                        t = std::min(t, ms);
                    }
                    if (st == KV_OK && t < best) {
                        best = t;
                        best_l[c] = l;
                        best_z[c] = z;
                        best_ms[c] = t;
                    }
                }
        }
        for (int i = 0; i < b->n && e == hipSuccess; ++i)
            if (backup[(size_t)i])
                e = hipMemcpyAsync(b->vol[(size_t)i], backup[(size_t)i],
                                   (size_t)kv_nvox(b->probs[(size_t)i].g) * 4, hipMemcpyDeviceToDevice, b->stream);
        if (e == hipSuccess)
            e = hipStreamSynchronize(b->stream);
        for (float *p : backup)
            if (p)
                (void)hipFree(p);
    }
    if (st != KV_OK)
        return st;
    if (e != hipSuccess)
        return fail_hip(e, "autotune");
    for (auto &P : b->probs) {
        const int c = size_class(P);
        if (best_l[c]) {
            P.fp_layout = best_l[c];
            P.bp_zt = best_z[c];
        }
    }
    layout_items(b);
    if (need && cache_path && *cache_path) {
        std::ofstream outf(cache_path);
        for (const auto &l : kept)
            outf << l << "\n";
        for (int c = 0; c < 3; ++c)
            if (best_l[c])
                outf << devkey << " " << c << " " << best_l[c] << " " << best_z[c] << " " << best_ms[c] << "\n";
    }
    return KV_OK;
}

}  // extern "C"
