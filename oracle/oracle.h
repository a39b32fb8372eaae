/*
 * oracle.h - CPU restatement of TomoGPI's cone-beam operator path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code; only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg use it, and only as the checker or the timed CPU baseline.
 *
 * Reference files cited below are relative to the reference repository root,
 * with C = src/TomoGPI_lib/src/CPP and K = src/TomoGPI_lib/src/CUDA.
 *
 * Two arithmetic variants of every operator:
 *   *_f32  the FP32 recipe: per-ray / per-column geometry in FP64, inner
 *          sample / tap / chord arithmetic and accumulation in FP32, written
 *          with explicit fmaf so that a GPU kernel following the same recipe
 *          reproduces it bit for bit (compile with -ffp-contract=off);
 *   *_f64  everything in FP64 (the accuracy reference for the mixed-precision
 *          forms).
 * Transposes (*_T_f64) are FP64 only.
 *
 * Pinning status of each function is stated in tests/test_oracle.py and in
 * the design notes; every function here is pinned by at least one closed
 * form, invariant or independent brute force.
 */
#ifndef TOMOGPI_ORACLE_H
#define TOMOGPI_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Raw configCT values (C/Config/ConfigCT.cpp:110-164). */
typedef struct og_desc {
    double  fod, fdd;
    int32_t nphi;
    double  num_start, den_start, num_stop, den_stop;
    float   su, sv;            /* detector sizes (mm)   */
    int32_t nu, nv;
    float   sx, sy, sz;        /* volume sizes (mm)     */
    int32_t nx, ny, nz;
} og_desc;

/* Derived geometry, typed as the reference stores it. */
typedef struct og_geom {
    int32_t nx, ny, nz, nu, nv, nphi;
    double  fod, fdd;
    float   du, dv, cu, cv;        /* Detector.cpp:9-14            */
    float   delta, cx, cy, cz;     /* Volume.cu:17-25 (dz := dx)   */
    double  start, stop, step;     /* ConfigCT.cpp:119             */
    double *phi, *cphi, *sphi;     /* Acquisition.cpp:90-97, FP64  */
    float  *alpha_f, *beta_f;      /* Projector.cu:15-38 (float)   */
} og_geom;

int  og_geom_init(const og_desc* d, og_geom* g);
void og_geom_free(og_geom* g);

/* VI backprojector: vol <- delta * sum_p bilinear(g_p, u, v)
 * (K/BackProjector/backprojection3D_kernel_16reg_UM.cuh:23-151).           */
void og_vi_f32(const og_geom* g, const float* sino, float* vol);
void og_vi_f64(const og_geom* g, const double* sino, double* vol);
void og_vi_T_f64(const og_geom* g, const double* vol, double* sino);
/* FDK-weighted VI (back_fdk, K/BackProjector/FDK3D_kernel_16reg.cuh). */
void og_vi_fdk_f32(const og_geom* g, const float* sino, float* vol);

void og_vi_fdk_f64(const og_geom* g, const double* sino, double* vol);

/* RS projector; cube = 0: inscribed cylinder (K/Projector/
 * projection_ER_kernel_UM.cuh:25-153), cube = 1: circumscribed circle +
 * square clip (K/Projector/projection_ER_kernel.cuh:56-312).               */
void     og_rs_f32(const og_geom* g, int cube, const float* vol, float* sino);
void     og_rs_f64(const og_geom* g, int cube, const double* vol, double* sino);
void     og_rs_T_f64(const og_geom* g, int cube, const double* sino, double* vol);
uint64_t og_rs_count(const og_geom* g, int cube);       /* N_samp = sum L_r */

/* Clean Siddon (exact chords, FOV = inscribed cylinder cut by the z extent,
 * source on the RS/VI side).                                               */
void     og_siddon_f32(const og_geom* g, const float* vol, float* sino);
void     og_siddon_f64(const og_geom* g, const double* vol, double* sino);
void     og_siddon_T_f64(const og_geom* g, const double* sino, double* vol);
uint64_t og_siddon_count(const og_geom* g);   /* N_seg: chords > 0, FP32 recipe */

/* Reference-compatible Siddon (C/SiddonProjector_compute_C_mem_CPU.cpp:
 * 20-276): float traversal, quirks kept.  Requires nx=ny=nz=256 and
 * nu,nv <= 256; returns nonzero otherwise.  nthreads <= 0: all cores.      */
int og_siddon_compat(const og_geom* g, const float* vol, float* sino, int nthreads);

/* 7-point Laplacian, zero padding (C/Convolution3D_CPU_GPU.cu:146-148). */
void og_laplacian_f32(int nx, int ny, int nz, const float* x, float* y);
void og_laplacian_f64(int nx, int ny, int nz, const double* x, double* y);

/* Phantoms. */
void og_shepp_logan(int nx, int ny, int nz, float* vol);
void og_ellipsoids(int nx, int ny, int nz, int count, uint64_t seed,
                   double amin, double amax, float* vol);
/* 10 doubles per ellipsoid: centre(3) semi-axes(3) density, then the 3x3
 * rotation follows in a second array of 9 doubles per ellipsoid.           */
void og_ellipsoid_table(int nx, int ny, int nz, int count, uint64_t seed,
                        double amin, double amax, double* params7, double* rot9);

/* Simple-gradient loop (C/Iter3D_Algo_Gradient.cu:25-223). */
#define OG_PAIR_RSVI     0
#define OG_PAIR_SIDDONVI 1
typedef struct og_gd_params {
    int32_t pair, cube, compat;
    int32_t iters, kopt;
    double  lambda;
    int32_t positivity;
    double  alpha_min_init;
    int32_t matched;      /* 1: use the exact transpose of P as B (tests) */
} og_gd_params;
/* j_reg = sum (L x)^2 of the iterate before the update when lambda != 0
 * (the GG beta=2 criterion, C/GeneralizedGaussianRegularizer_CPU.cu:178),
 * beta = the conjugate gradient's |g|^2/|g_prev|^2 (0 otherwise).          */
typedef struct og_gd_stat {
    double residual_sq, norm_dj, norm_dj_proj, alpha, j_reg, beta;
} og_gd_stat;
int og_gd_run(const og_geom* g, const og_gd_params* p, const float* sino,
              float* vol, og_gd_stat* stats);

/* Weighted gradient (C/Iter3D_Algo_Gradient.cu:227-371): the simple loop with
 * the residual multiplied by the weight sinogram before the backprojection;
 * J_MC and the step's |P dJ|^2 stay unweighted, as the reference writes
 * them.  The reference starts alpha_min at -999999999999 here.              */
int og_wgd_run(const og_geom* g, const og_gd_params* p, const float* sino, const float* weight,
               float* vol, og_gd_stat* stats);

/* Conjugate gradient on the criterion of og_gd_run (defined by this build:
 * the reference's doConjugateGradient, C/Iter3D_Algo_Gradient.cu:375-559,
 * starts its direction at 0 and divides by its zero norm at iteration 2):
 *   g = -2 B (y - P x) + 2 lambda L(L x)       (the loop's dJ)
 *   d = g (k = 1),  d = g + beta d,  beta = |g|^2 / |g_prev|^2 (Fletcher-Reeves)
 *   alpha = -(g . d) / (2 |P d|^2 + lambda 2 |L d|^2);  x += alpha d
 * stats: residual_sq, norm_dj = |g|^2, norm_dj_proj = the denominator, alpha.
 * kopt and alpha_min_init are not used.                                      */
int og_cg_run(const og_geom* g, const og_gd_params* p, const float* sino, float* vol, og_gd_stat* stats);

/* SIRT, defined by this build (the reference has none):
 *   R = 1 / (P 1) on the sinogram, C = relax / (B 1) on the volume (0 where
 *   the denominator is 0), x0 = 0,
 *   r = (y - P x) * R;  x = x + C * (B r)   (clamped at 0 with positivity)
 * with P, B the projector and backprojector of the pair.  stats (may be
 * NULL) receives sum (y - P x)^2 before each update.                         */
typedef struct og_sirt_params {
    int32_t pair, cube;
    int32_t iters;
    double  relax;
    int32_t positivity;
} og_sirt_params;
int og_sirt_weights(const og_geom* g, const og_sirt_params* p, float* rw, float* cw);
int og_sirt_run(const og_geom* g, const og_sirt_params* p, const float* sino, float* vol,
                double* residual_sq);

/* Thread count used by the OpenMP loops (<= 0: OpenMP default). */
void og_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
