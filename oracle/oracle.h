/*
 * CPU oracle for the tomosynthesis projection / reconstruction path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: the parity tests,
 * the smoke check and the CPU-baseline leg of the benchmark call it.  The
 * product path (libkevernals_amd.so) never links, loads or calls it.
 *
 * Two operator families are restated here:
 *
 *  mode R  ("reference parity") follows KevLoq/Kevernals
 *          source/modules/reconstruction/Projector.cu literally:
 *          FP32 arithmetic, no floating-point contraction, the crossing list
 *          built by fill-with-1.0 + three arithmetic runs + selection sort
 *          (Projector.cu:45-223), the FP64 midpoint voxel lookup
 *          (Projector.cu:348-371) and the voxel-driven +-5 pixel window
 *          backprojector (Projector.cu:376-785).  Reconstruction loops follow
 *          source/modules/reconstruction/Reconstructors.h:19-379.
 *
 *  mode A  ("adjoint") is this project's own matched pair: un-normalised
 *          Siddon line integrals in mm and their exact transpose, used by the
 *          physics loops (log-likelihood, learned steps).  It has no
 *          reference counterpart.
 *
 * Every mode-R function has an FP64 twin (same algorithm, double arithmetic)
 * used only to state the FP32 error.
 *
 * Layouts: volume float[nz][ny][nx] (x fastest), projections
 * float[np][nv][nu] (u fastest), as Projector.cu:264,361-363,401.
 */
#ifndef KV_ORACLE_H
#define KV_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Kernel argument list of CudaPerformProjection / CudaPerformBackProjection
 * (Projector.cu:227-236, 376-385).  det[3p..3p+2] is the per-view detector
 * bottom-left corner (x, y) and plane z; src[3p..3p+2] the source position,
 * both in world mm with the volume centred on the origin. */
typedef struct or_geom {
    int32_t nx, ny, nz;
    float sx, sy, sz;
    int32_t nu, nv, np;
    float pu, pv;
    const float *det;
    const float *src;
} or_geom;

/* ---- mode R, single ray (A1/A2) ---------------------------------------- */
/* Sorted crossing list of ray (p,u,v) exactly as FindProjectionAlphas_d
 * returns it.  Returns n (0 when the reference exits early), or -n when n
 * exceeds cap (nothing written beyond cap). */
int or_r_ray_alphas(const or_geom *g, int p, int u, int v, float *alphas, int cap);
/* Accepted (voxel index, weight) pairs of ray (p,u,v) in walk order.
 * Returns the count, or -count when it exceeds cap. */
int or_r_ray_trace(const or_geom *g, int p, int u, int v, int64_t *idx, float *w, int cap);

/* ---- mode R operators ---------------------------------------------------- */
void or_r_fp(const or_geom *g, const float *vol, float *proj, int nthreads);
void or_r_bp(const or_geom *g, const float *proj, float *vol, int nthreads);
void or_r_fp_f64(const or_geom *g, const double *vol, double *proj, int nthreads);
void or_r_bp_f64(const or_geom *g, const double *proj, double *vol, int nthreads);
/* Sub-ranges for bounded CPU-baseline timing: detector rows [v0,v1) of view
 * p, and volume slices [z0,z1). */
void or_r_fp_rows(const or_geom *g, const float *vol, float *proj, int p, int v0, int v1, int nthreads);
void or_r_bp_slices(const or_geom *g, const float *proj, float *vol, int z0, int z1, int nthreads);
/* Mode-R values of sampled rays (id = (p*nv + v)*nu + u) and voxels
 * (id = (z*ny + y)*nx + x), for bounded checks at full size. */
void or_r_fp_rays(const or_geom *g, const float *vol, int n, const int64_t *rays, float *out, int nthreads);
void or_r_bp_voxels(const or_geom *g, const float *proj, int n, const int64_t *vox, float *out, int nthreads);
/* Accepted (pixel, weight) pairs of voxel (x,y,z) in the backprojector's
 * accumulation order; pixel = (p*nv + v)*nu + u.  Returns the count, or
 * -count when it exceeds cap. */
int or_r_bp_voxel_trace(const or_geom *g, int x, int y, int z, int64_t *pix, float *w, int cap);
/* Accepted (ray, voxel) pair counts of the two mode-R operators. */
int64_t or_r_fp_nnz(const or_geom *g, int nthreads);
int64_t or_r_bp_nnz(const or_geom *g, int nthreads);

/* ---- mode R reconstruction loops (Reconstructors.h) ---------------------- */
/* recons::ART (Reconstructors.h:19-134): vol is the initial volume, updated
 * in place.  recons::MLEM (Reconstructors.h:136-253) likewise. */
void or_r_art(const or_geom *g, const float *data, float *vol, int iters, float relax, int nthreads);
void or_r_mlem(const or_geom *g, const float *data, float *vol, int iters, float relax, int nthreads);
/* recons::ShiftAndAdd (Reconstructors.h:263-379).  plane_z[nplanes] are the
 * slice heights, src_y[np] the source y positions.  out is
 * float[nplanes][nv][nu].  Returns 0, or a negative code for the reference's
 * error exits (-1: planes<=1, -2: views<=1, -3: fulcrum at sid,
 * -4: a plane at sid). */
int or_saa(int nu, int nv, int np, const float *proj, const float *src_y,
           int nplanes, const float *plane_z, float sid, float fulcrum_z,
           float voxel_dy, float *out);

/* ---- AlmostEqualRelative (commons/Maths.h:25-34) ------------------------ */
int or_almost_equal_relative(float a, float b, float max_rel_diff);

/* ---- mode A (matched adjoint pair) --------------------------------------- */
void or_a_fp(const or_geom *g, const float *vol, float *proj, int nthreads);
void or_a_bp(const or_geom *g, const float *proj, float *vol, int nthreads);
void or_a_fp_f64(const or_geom *g, const double *vol, double *proj, int nthreads);
void or_a_bp_f64(const or_geom *g, const double *proj, double *vol, int nthreads);
/* Brute-force mode-A forward projection: every voxel of every ray through
 * the box formula.  Small volumes only. */
void or_a_fp_brute(const or_geom *g, const float *vol, float *proj, int nthreads);
/* mode-A voxels and alpha-lengths of one ray, in walk order. */
int or_a_ray_trace(const or_geom *g, int p, int u, int v, int64_t *idx, float *len, int cap);
/* Plain MLEM on line-integral data in FP64: x <- x / s * A^T(y / Ax). */
void or_a_mlem_f64(const or_geom *g, const double *data, double *vol, int iters, int nthreads);
/* Poisson log-likelihood sum(y log(Ax) - Ax) over rays with Ax > 0, FP64. */
double or_a_loglik_f64(const or_geom *g, const double *data, const double *vol, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
