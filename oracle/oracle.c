/*
 * CPU oracle -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp).  Contraction
 * must stay off: the mode-R voxel sequences depend on every FP32 rounding.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------- FP32 -- */
#define REAL float
#define SFX _f32
#define RABS fabsf
#define RMIN fminf
#define RMAX fmaxf
#define RSQRT sqrtf
#include "oracle_r.inc"
#include "oracle_a.inc"
#undef REAL
#undef SFX
#undef RABS
#undef RMIN
#undef RMAX
#undef RSQRT

/* ---------------------------------------------------------------- FP64 -- */
#define REAL double
#define SFX _f64
#define RABS fabs
#define RMIN fmin
#define RMAX fmax
#define RSQRT sqrt
#include "oracle_r.inc"
#include "oracle_a.inc"
#undef REAL
#undef SFX
#undef RABS
#undef RMIN
#undef RMAX
#undef RSQRT

static int nthr(int n) { return n > 0 ? n : 1; }

/* ------------------------------------------------------- mode R, one ray */
int or_r_ray_alphas(const or_geom *g, int p, int u, int v, float *alphas, int cap)
{
    const int nmax = g->nx + g->ny + g->nz + 3;
    float *a = (float *)malloc(sizeof(float) * (size_t)nmax);
    r_ray_t_f32 r;
    r_ray_setup_f32(g, p, u, v, &r);
    const int n = r_find_alphas_f32(&r, g->nx, g->ny, g->nz, a);
    const int m = n < cap ? n : cap;
    memcpy(alphas, a, sizeof(float) * (size_t)(m > 0 ? m : 0));
    free(a);
    return n <= cap ? n : -n;
}

int or_r_ray_trace(const or_geom *g, int p, int u, int v, int64_t *idx, float *w, int cap)
{
    const int nmax = g->nx + g->ny + g->nz + 3;
    float *a = (float *)malloc(sizeof(float) * (size_t)nmax);
    r_ray_t_f32 r;
    r_ray_setup_f32(g, p, u, v, &r);
    const int n = r_find_alphas_f32(&r, g->nx, g->ny, g->nz, a);
    int count = 0;
    r_ray_value_f32(g, &r, NULL, a, n, idx, w, cap, &count);
    free(a);
    return count <= cap ? count : -count;
}

/* ------------------------------------------------------ mode R operators */
void or_r_fp(const or_geom *g, const float *vol, float *proj, int nthreads)
{
    r_fp_range_f32(g, vol, proj, 0, g->np, 0, g->nv, nthr(nthreads));
}

void or_r_fp_rows(const or_geom *g, const float *vol, float *proj, int p, int v0, int v1, int nthreads)
{
    r_fp_range_f32(g, vol, proj, p, p + 1, v0, v1, nthr(nthreads));
}

void or_r_bp(const or_geom *g, const float *proj, float *vol, int nthreads)
{
    r_bp_range_f32(g, proj, vol, 0, g->nz, nthr(nthreads));
}

void or_r_bp_slices(const or_geom *g, const float *proj, float *vol, int z0, int z1, int nthreads)
{
    r_bp_range_f32(g, proj, vol, z0, z1, nthr(nthreads));
}

void or_r_fp_f64(const or_geom *g, const double *vol, double *proj, int nthreads)
{
    r_fp_range_f64(g, vol, proj, 0, g->np, 0, g->nv, nthr(nthreads));
}

void or_r_bp_f64(const or_geom *g, const double *proj, double *vol, int nthreads)
{
    r_bp_range_f64(g, proj, vol, 0, g->nz, nthr(nthreads));
}

int64_t or_r_fp_nnz(const or_geom *g, int nthreads)
{
    const int nmax = g->nx + g->ny + g->nz + 3;
    int64_t total = 0;
#pragma omp parallel num_threads(nthr(nthreads)) reduction(+ : total)
    {
        float *a = (float *)malloc(sizeof(float) * (size_t)nmax);
#pragma omp for schedule(dynamic, 4) collapse(2)
        for (int p = 0; p < g->np; p++)
            for (int v = 0; v < g->nv; v++)
                for (int u = 0; u < g->nu; u++) {
                    r_ray_t_f32 r;
                    r_ray_setup_f32(g, p, u, v, &r);
                    const int n = r_find_alphas_f32(&r, g->nx, g->ny, g->nz, a);
                    int c = 0;
                    r_ray_value_f32(g, &r, NULL, a, n, NULL, NULL, 0, &c);
                    total += c;
                }
        free(a);
    }
    return total;
}

int64_t or_r_bp_nnz(const or_geom *g, int nthreads)
{
    int64_t total = 0;
#pragma omp parallel for num_threads(nthr(nthreads)) schedule(dynamic, 1) collapse(2) reduction(+ : total)
    for (int z = 0; z < g->nz; z++)
        for (int y = 0; y < g->ny; y++)
            for (int x = 0; x < g->nx; x++) {
                int64_t c = 0;
                r_bp_voxel_f32(g, NULL, x, y, z, &c, NULL, NULL, 0);
                total += c;
            }
    return total;
}

int or_r_bp_voxel_trace(const or_geom *g, int x, int y, int z, int64_t *pix, float *w, int cap)
{
    int64_t c = 0;
    r_bp_voxel_f32(g, NULL, x, y, z, &c, pix, w, cap);
    return c <= cap ? (int)c : -(int)c;
}

void or_r_fp_rays(const or_geom *g, const float *vol, int n, const int64_t *rays, float *out, int nthreads)
{
    const int nmax = g->nx + g->ny + g->nz + 3;
    const int64_t per = (int64_t)g->nu * g->nv;
#pragma omp parallel num_threads(nthr(nthreads))
    {
        float *a = (float *)malloc(sizeof(float) * (size_t)nmax);
#pragma omp for schedule(dynamic, 4)
        for (int i = 0; i < n; i++) {
            const int p = (int)(rays[i] / per);
            const int64_t rem = rays[i] - p * per;
            const int v = (int)(rem / g->nu), u = (int)(rem - (int64_t)v * g->nu);
            r_ray_t_f32 r;
            r_ray_setup_f32(g, p, u, v, &r);
            const int m = r_find_alphas_f32(&r, g->nx, g->ny, g->nz, a);
            out[i] = r_ray_value_f32(g, &r, vol, a, m, NULL, NULL, 0, NULL);
        }
        free(a);
    }
}

void or_r_bp_voxels(const or_geom *g, const float *proj, int n, const int64_t *vox, float *out, int nthreads)
{
    const int64_t per = (int64_t)g->nx * g->ny;
#pragma omp parallel for num_threads(nthr(nthreads)) schedule(dynamic, 4)
    for (int i = 0; i < n; i++) {
        const int z = (int)(vox[i] / per);
        const int64_t rem = vox[i] - z * per;
        const int y = (int)(rem / g->nx), x = (int)(rem - (int64_t)y * g->nx);
        out[i] = r_bp_voxel_f32(g, proj, x, y, z, NULL, NULL, NULL, 0);
    }
}

/* ------------------------------------------------- Maths.h:25-34 ------- */
int or_almost_equal_relative(float a, float b, float max_rel_diff)
{
    const float diff = fabsf(a - b);
    a = fabsf(a);
    b = fabsf(b);
    const float scaled = max_rel_diff * ((a < b) ? b : a); /* std::max(a, b) */
    return diff <= scaled;
}

/* ------------------------------------- Reconstructors.h:19-134 (ART) --- */
void or_r_art(const or_geom *g, const float *data, float *vol, int iters, float relax, int nthreads)
{
    const int64_t nr = (int64_t)g->np * g->nv * g->nu;
    const int64_t nv = (int64_t)g->nz * g->ny * g->nx;
    float *cur = (float *)malloc(sizeof(float) * (size_t)nr);
    float *back = (float *)malloc(sizeof(float) * (size_t)nv);
    for (int it = 0; it < iters; it++) {
        or_r_fp(g, vol, cur, nthreads);
        for (int64_t i = 0; i < nr; i++)
            cur[i] = relax * (data[i] - cur[i]);
        or_r_bp(g, cur, back, nthreads);
        for (int64_t i = 0; i < nv; i++)
            vol[i] += back[i];
    }
    free(cur);
    free(back);
}

/* ------------------------------------ Reconstructors.h:136-253 (MLEM) --- */
void or_r_mlem(const or_geom *g, const float *data, float *vol, int iters, float relax, int nthreads)
{
    const int64_t nr = (int64_t)g->np * g->nv * g->nu;
    const int64_t nv = (int64_t)g->nz * g->ny * g->nx;
    float *cur = (float *)malloc(sizeof(float) * (size_t)nr);
    float *back = (float *)malloc(sizeof(float) * (size_t)nv);
    for (int it = 0; it < iters; it++) {
        or_r_fp(g, vol, cur, nthreads);
        for (int64_t i = 0; i < nr; i++)
            if (!or_almost_equal_relative(cur[i], 0.f, FLT_EPSILON))
                cur[i] = relax * (data[i] / cur[i]);
        or_r_bp(g, cur, back, nthreads);
        for (int64_t i = 0; i < nv; i++)
            vol[i] *= back[i];
    }
    free(cur);
    free(back);
}

/* ------------------------------ Reconstructors.h:263-379 (ShiftAndAdd) --- */
int or_saa(int nu, int nv, int np, const float *proj, const float *src_y, int nplanes, const float *plane_z,
           float sid, float fulcrum_z, float voxel_dy, float *out)
{
    if (nplanes <= 1)
        return -1;
    if (np <= 1)
        return -2;
    if (or_almost_equal_relative(fulcrum_z, sid, FLT_EPSILON))
        return -3;
    const float mF = sid / (sid - fulcrum_z);
    int *t = (int *)malloc(sizeof(int) * (size_t)np);
    const double wgt = 1. / (double)np;
    for (int k = 0; k < nplanes; k++) {
        const float z = plane_z[k];
        if (or_almost_equal_relative(z, sid, FLT_EPSILON)) {
            free(t);
            return -4;
        }
        const float mZ = sid / (sid - z);
        const float coeff = mZ - mF;
        for (int p = 0; p < np; p++) {
            const float world = src_y[p] * coeff;
            t[p] = (int)(world / voxel_dy);
        }
        /* zero-padded copy, shifted VOI per view, weighted sum with 1/P:
         * out(u,v) = sum_p (1/P) * proj_p(u, v + t_p), zero outside. */
        for (int v = 0; v < nv; v++)
            for (int u = 0; u < nu; u++) {
                double acc = 0;
                for (int p = 0; p < np; p++) {
                    const int vv = v + t[p];
                    const double x = (vv >= 0 && vv < nv) ? (double)proj[((int64_t)p * nv + vv) * nu + u] : 0.0;
                    acc += x * wgt;
                }
                out[((int64_t)k * nv + v) * nu + u] = (float)acc;
            }
    }
    free(t);
    return 0;
}

/* ------------------------------------------------------------ mode A --- */
void or_a_fp(const or_geom *g, const float *vol, float *proj, int nthreads)
{
    a_fp_all_f32(g, vol, proj, nthr(nthreads), 0);
}

void or_a_fp_brute(const or_geom *g, const float *vol, float *proj, int nthreads)
{
    a_fp_all_f32(g, vol, proj, nthr(nthreads), 1);
}

void or_a_bp(const or_geom *g, const float *proj, float *vol, int nthreads)
{
    a_bp_all_f32(g, proj, vol, nthr(nthreads));
}

void or_a_fp_f64(const or_geom *g, const double *vol, double *proj, int nthreads)
{
    a_fp_all_f64(g, vol, proj, nthr(nthreads), 0);
}

void or_a_bp_f64(const or_geom *g, const double *proj, double *vol, int nthreads)
{
    a_bp_all_f64(g, proj, vol, nthr(nthreads));
}

int or_a_ray_trace(const or_geom *g, int p, int u, int v, int64_t *idx, float *len, int cap)
{
    a_ray_t_f32 r;
    a_ray_setup_f32(g, p, u, v, &r);
    const int c = a_walk_f32(&r, NULL, NULL, idx, len, cap);
    return c <= cap ? c : -c;
}

void or_a_mlem_f64(const or_geom *g, const double *data, double *vol, int iters, int nthreads)
{
    const int64_t nr = (int64_t)g->np * g->nv * g->nu;
    const int64_t nv = (int64_t)g->nz * g->ny * g->nx;
    double *ones = (double *)malloc(sizeof(double) * (size_t)nr);
    double *sens = (double *)malloc(sizeof(double) * (size_t)nv);
    double *ax = (double *)malloc(sizeof(double) * (size_t)nr);
    double *back = (double *)malloc(sizeof(double) * (size_t)nv);
    for (int64_t i = 0; i < nr; i++)
        ones[i] = 1.0;
    or_a_bp_f64(g, ones, sens, nthreads);
    for (int it = 0; it < iters; it++) {
        or_a_fp_f64(g, vol, ax, nthreads);
        for (int64_t i = 0; i < nr; i++)
            ax[i] = ax[i] > 0 ? data[i] / ax[i] : 0.0;
        or_a_bp_f64(g, ax, back, nthreads);
        for (int64_t i = 0; i < nv; i++)
            if (sens[i] > 0)
                vol[i] = vol[i] * back[i] / sens[i];
    }
    free(ones);
    free(sens);
    free(ax);
    free(back);
}

double or_a_loglik_f64(const or_geom *g, const double *data, const double *vol, int nthreads)
{
    const int64_t nr = (int64_t)g->np * g->nv * g->nu;
    double *ax = (double *)malloc(sizeof(double) * (size_t)nr);
    or_a_fp_f64(g, vol, ax, nthreads);
    double ll = 0;
    for (int64_t i = 0; i < nr; i++)
        if (ax[i] > 0)
            ll += data[i] * log(ax[i]) - ax[i];
    free(ax);
    return ll;
}
