/*
 * oracle_geom.c - geometry derivation, phantoms, Laplacian and the
 * simple-gradient loop of the CPU oracle.  TEST INFRASTRUCTURE ONLY (see
 * oracle.h).  Build with -ffp-contract=off: every fused multiply-add here is
 * an explicit fma()/fmaf() call so that the GPU kernels can follow the same
 * arithmetic.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static int g_threads = 0;
void og_set_threads(int n) {
    g_threads = n;
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#endif
}

/* Geometry exactly as the reference constructors type it:
 *  ConfigCT::createUniformAcquisition (C/Config/ConfigCT.cpp:110-123):
 *      step = M_PI*(stop - start)/Nphi in double, start = num/den;
 *  UniformAcquisition (C/Acquisition.cpp:90-97): phi_p = (p + start)*step;
 *  Detector (C/Detector.cpp:9-14): du = Su/Nu (float), cu = Nu/2.0 - 0.5;
 *  Volume (C/Volume.cu:17-25): delta = Sx/Nx (float), dz := dx, c = N/2.0-0.5;
 *  Projector (C/Projector.cu:15-38): alpha_p = (float)cos(phi_p).         */
int og_geom_init(const og_desc* d, og_geom* g) {
    memset(g, 0, sizeof(*g));
    if (d->nx <= 0 || d->ny <= 0 || d->nz <= 0 || d->nu <= 0 || d->nv <= 0 || d->nphi <= 0)
        return 1;
    g->nx = d->nx; g->ny = d->ny; g->nz = d->nz;
    g->nu = d->nu; g->nv = d->nv; g->nphi = d->nphi;
    g->fod = d->fod; g->fdd = d->fdd;
    g->du = d->su / (float)d->nu;
    g->dv = d->sv / (float)d->nv;
    g->cu = (float)(d->nu / 2.0 - 0.5);
    g->cv = (float)(d->nv / 2.0 - 0.5);
    g->delta = d->sx / (float)d->nx;
    g->cx = (float)(d->nx / 2.0 - 0.5);
    g->cy = (float)(d->ny / 2.0 - 0.5);
    g->cz = (float)(d->nz / 2.0 - 0.5);
    g->start = d->num_start / d->den_start;
    g->stop = d->num_stop / d->den_stop;
    g->step = M_PI * (g->stop - g->start) / (double)d->nphi;
    g->phi = (double*)malloc(sizeof(double) * d->nphi);
    g->cphi = (double*)malloc(sizeof(double) * d->nphi);
    g->sphi = (double*)malloc(sizeof(double) * d->nphi);
    g->alpha_f = (float*)malloc(sizeof(float) * d->nphi);
    g->beta_f = (float*)malloc(sizeof(float) * d->nphi);
    for (int p = 0; p < d->nphi; ++p) {
        g->phi[p] = ((double)p + g->start) * g->step;
        g->cphi[p] = cos(g->phi[p]);
        g->sphi[p] = sin(g->phi[p]);
        g->alpha_f[p] = (float)g->cphi[p];
        g->beta_f[p] = (float)g->sphi[p];
    }
    return 0;
}

void og_geom_free(og_geom* g) {
    free(g->phi); free(g->cphi); free(g->sphi); free(g->alpha_f); free(g->beta_f);
    memset(g, 0, sizeof(*g));
}

/* ---------------------------------------------------------------- phantoms */

/* Modified Shepp-Logan ellipsoid geometry (Data3D_0256/phantom3d.m:186-200)
 * with the Kak-Slaney densities 2,-.98,-.02,-.02,.01 x6 (the "0-2" range of
 * the benchmark definition).  Columns: A a b c x0 y0 z0 phi theta psi.      */
static const double kSheppLogan[10][10] = {
    { 2.00, .6900, .920, .810,  0.00,  0.0000,  0.00,   0, 0,  0},
    {-0.98, .6624, .874, .780,  0.00, -0.0184,  0.00,   0, 0,  0},
    {-0.02, .1100, .310, .220,  0.22,  0.0000,  0.00, -18, 0, 10},
    {-0.02, .1600, .410, .280, -0.22,  0.0000,  0.00,  18, 0, 10},
    { 0.01, .2100, .250, .410,  0.00,  0.3500, -0.15,   0, 0,  0},
    { 0.01, .0460, .046, .050,  0.00,  0.1000,  0.25,   0, 0,  0},
    { 0.01, .0460, .046, .050,  0.00, -0.1000,  0.25,   0, 0,  0},
    { 0.01, .0460, .023, .050, -0.08, -0.6050,  0.00,   0, 0,  0},
    { 0.01, .0230, .023, .020,  0.00, -0.6060,  0.00,   0, 0,  0},
    { 0.01, .0230, .046, .020,  0.06, -0.6050,  0.00,   0, 0,  0},
};

/* Euler rotation used by phantom3d.m:96-103 (z-x-z convention). */
static void euler_matrix(double phi_deg, double theta_deg, double psi_deg, double m[9]) {
    double ph = phi_deg * M_PI / 180.0, th = theta_deg * M_PI / 180.0, ps = psi_deg * M_PI / 180.0;
    double cph = cos(ph), sph = sin(ph), cth = cos(th), sth = sin(th), cps = cos(ps), sps = sin(ps);
    m[0] = cps * cph - cth * sph * sps;  m[1] = cps * sph + cth * cph * sps;  m[2] = sps * sth;
    m[3] = -sps * cph - cth * sph * cps; m[4] = -sps * sph + cth * cph * cps; m[5] = cps * sth;
    m[6] = sth * sph;                    m[7] = -sth * cph;                   m[8] = cth;
}

/* Voxel (i,j,k) has normalised coordinates ((i-(n-1)/2)/((n-1)/2), ...) with
 * x the file-fastest axis (phantom3d.m:76).  A voxel belongs to an ellipsoid
 * when sum(((M.r)_a - c_a)^2 / s_a^2) <= 1 (phantom3d.m:106-107).          */
void og_shepp_logan(int nx, int ny, int nz, float* vol) {
    double m[10][9];
    for (int e = 0; e < 10; ++e)
        euler_matrix(kSheppLogan[e][7], kSheppLogan[e][8], kSheppLogan[e][9], m[e]);
    const double hx = (nx - 1) / 2.0, hy = (ny - 1) / 2.0, hz = (nz - 1) / 2.0;
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const double x = ((double)i - hx) / hx, y = ((double)j - hy) / hy, z = ((double)k - hz) / hz;
                double acc = 0.0;
                for (int e = 0; e < 10; ++e) {
                    const double* r = m[e];
                    const double px = (r[0] * x + r[1] * y) + r[2] * z;
                    const double py = (r[3] * x + r[4] * y) + r[5] * z;
                    const double pz = (r[6] * x + r[7] * y) + r[8] * z;
                    const double tx = px - kSheppLogan[e][4], ty = py - kSheppLogan[e][5], tz = pz - kSheppLogan[e][6];
                    const double a2 = kSheppLogan[e][1] * kSheppLogan[e][1];
                    const double b2 = kSheppLogan[e][2] * kSheppLogan[e][2];
                    const double c2 = kSheppLogan[e][3] * kSheppLogan[e][3];
                    const double q = ((tx * tx) / a2 + (ty * ty) / b2) + (tz * tz) / c2;
                    if (q <= 1.0) acc += kSheppLogan[e][0];
                }
                vol[(size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k)] = (float)acc;
            }
}

/* splitmix64 (Steele, Lea, Flood 2014) and 53-bit uniforms in [0,1). */
static uint64_t splitmix64(uint64_t* s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static double u01(uint64_t* s) { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

/* Random ellipsoid table.  Per ellipsoid, 10 draws in this order: centre
 * x,y,z (uniform over the inner 80% of each axis, in voxels from the volume
 * centre), semi-axes a,b,c in [amin,amax] voxels, three uniforms giving a
 * uniformly random unit quaternion (Shoemake 1992), and the density U[0,1).
 * params7 = cx cy cz 1/a^2 1/b^2 1/c^2 density; rot9 = row-major rotation
 * R taking volume offsets to ellipsoid axes.                               */
void og_ellipsoid_table(int nx, int ny, int nz, int count, uint64_t seed,
                        double amin, double amax, double* params7, double* rot9) {
    uint64_t s = seed;
    const int n[3] = {nx, ny, nz};
    for (int e = 0; e < count; ++e) {
        double* P = params7 + 7 * e;
        double* R = rot9 + 9 * e;
        for (int a = 0; a < 3; ++a) P[a] = (u01(&s) - 0.5) * 0.8 * (double)n[a];
        for (int a = 0; a < 3; ++a) {
            const double semi = amin + (amax - amin) * u01(&s);
            P[3 + a] = 1.0 / (semi * semi);
        }
        const double u1 = u01(&s), u2 = u01(&s), u3 = u01(&s);
        const double r1 = sqrt(1.0 - u1), r2 = sqrt(u1);
        const double qx = r1 * sin(2.0 * M_PI * u2), qy = r1 * cos(2.0 * M_PI * u2);
        const double qz = r2 * sin(2.0 * M_PI * u3), qw = r2 * cos(2.0 * M_PI * u3);
        R[0] = 1.0 - 2.0 * (qy * qy + qz * qz); R[1] = 2.0 * (qx * qy - qz * qw);       R[2] = 2.0 * (qx * qz + qy * qw);
        R[3] = 2.0 * (qx * qy + qz * qw);       R[4] = 1.0 - 2.0 * (qx * qx + qz * qz); R[5] = 2.0 * (qy * qz - qx * qw);
        R[6] = 2.0 * (qx * qz - qy * qw);       R[7] = 2.0 * (qy * qz + qx * qw);       R[8] = 1.0 - 2.0 * (qx * qx + qy * qy);
        P[6] = u01(&s);
    }
}

void og_ellipsoids(int nx, int ny, int nz, int count, uint64_t seed, double amin, double amax, float* vol) {
    double* P = (double*)malloc(sizeof(double) * 7 * (count > 0 ? count : 1));
    double* R = (double*)malloc(sizeof(double) * 9 * (count > 0 ? count : 1));
    og_ellipsoid_table(nx, ny, nz, count, seed, amin, amax, P, R);
    const double hx = (nx - 1) / 2.0, hy = (ny - 1) / 2.0, hz = (nz - 1) / 2.0;
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const double ox = (double)i - hx, oy = (double)j - hy, oz = (double)k - hz;
                double acc = 0.0;
                for (int e = 0; e < count; ++e) {
                    const double* p = P + 7 * e;
                    const double* r = R + 9 * e;
                    const double dx = ox - p[0], dy = oy - p[1], dz = oz - p[2];
                    const double l0 = (r[0] * dx + r[1] * dy) + r[2] * dz;
                    const double l1 = (r[3] * dx + r[4] * dy) + r[5] * dz;
                    const double l2 = (r[6] * dx + r[7] * dy) + r[8] * dz;
                    const double q = ((l0 * l0) * p[3] + (l1 * l1) * p[4]) + (l2 * l2) * p[5];
                    if (q <= 1.0) acc += p[6];
                }
                vol[(size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k)] = (float)acc;
            }
    free(P); free(R);
}

/* --------------------------------------------------------------- Laplacian */

/* L x = 6x - sum of the 6 neighbours, zero outside the volume: the {-1,2,-1}
 * row pass written, column and plane passes accumulated
 * (C/Convolution3D_CPU_GPU.cu:146-148, C/GeneralizedGaussianRegularizer_CPU
 * .cu:163-165, zero halos K/Convolution/Convolution3D_sep_shared_kernel_Acc
 * .cuh:34-61).  Summation order: ((row + column) + plane), each pass
 * (2x - previous) - next.                                                  */
#define LAPLACIAN_BODY(T)                                                                  \
    const size_t sx = 1, sy = (size_t)nx, sz = (size_t)nx * ny;                            \
    _Pragma("omp parallel for collapse(2) schedule(static)")                               \
    for (int k = 0; k < nz; ++k)                                                           \
        for (int j = 0; j < ny; ++j)                                                       \
            for (int i = 0; i < nx; ++i) {                                                 \
                const size_t o = (size_t)i + sy * j + sz * k;                              \
                const T c = x[o], c2 = c + c;                                              \
                const T xm = i > 0 ? x[o - sx] : (T)0, xp = i < nx - 1 ? x[o + sx] : (T)0; \
                const T ym = j > 0 ? x[o - sy] : (T)0, yp = j < ny - 1 ? x[o + sy] : (T)0; \
                const T zm = k > 0 ? x[o - sz] : (T)0, zp = k < nz - 1 ? x[o + sz] : (T)0; \
                const T h = (c2 - xm) - xp, v = (c2 - ym) - yp, p = (c2 - zm) - zp;        \
                y[o] = (h + v) + p;                                                        \
            }

void og_laplacian_f32(int nx, int ny, int nz, const float* x, float* y) { LAPLACIAN_BODY(float) }
void og_laplacian_f64(int nx, int ny, int nz, const double* x, double* y) { LAPLACIAN_BODY(double) }

/* ------------------------------------------------------------------- loop */

static double sq_norm(const float* x, size_t n) {
    double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (size_t i = 0; i < n; ++i) s += (double)x[i] * (double)x[i];
    return s;
}

static void project(const og_geom* g, const og_gd_params* p, const float* vol, float* sino) {
    if (p->pair == OG_PAIR_SIDDONVI) {
        if (p->compat) og_siddon_compat(g, vol, sino, 0);
        else og_siddon_f32(g, vol, sino);
    } else {
        og_rs_f32(g, p->cube, vol, sino);
    }
}

/* Backprojector of the loop: VI, or (matched = 1, tests only) the FP64
 * transpose of the projector rounded to float.                             */
static void backproject(const og_geom* g, const og_gd_params* p, const float* sino, float* vol) {
    if (!p->matched) { og_vi_f32(g, sino, vol); return; }
    const size_t nv = (size_t)g->nx * g->ny * g->nz, ns = (size_t)g->nu * g->nv * g->nphi;
    double* s64 = (double*)malloc(sizeof(double) * ns);
    double* v64 = (double*)malloc(sizeof(double) * nv);
    for (size_t i = 0; i < ns; ++i) s64[i] = sino[i];
    if (p->pair == OG_PAIR_SIDDONVI) og_siddon_T_f64(g, s64, v64);
    else og_rs_T_f64(g, p->cube, s64, v64);
    for (size_t i = 0; i < nv; ++i) vol[i] = (float)v64[i];
    free(s64); free(v64);
}

/* Iter3D::doSimpleGradient (C/Iter3D_Algo_Gradient.cu:25-223) with the
 * generalised-Gaussian beta=2 regulariser
 * (C/GeneralizedGaussianRegularizer_CPU.cu:158-226):
 *   est = P x; r = y - est; dJ = B r; dJ *= -2;
 *   if lambda != 0: dJ += (2 lambda) L(L x);
 *                   if k < Kopt: normdJProjReg = 2 sum (L dJ)^2;
 *   if k < Kopt: est = P dJ; normdJ = sum dJ^2;
 *                normdJProj = 2 sum est^2 + lambda normdJProjReg;
 *                alpha = -normdJ/normdJProj; alpha_min = max(alpha, alpha_min)
 *   else alpha = alpha_min;
 *   x += (float)alpha * dJ, clamped at 0 if positivity (C/Projector.cu:109-124,
 *   C/Image3D_CPU.cu:151-205).                                             */
/* The simple gradient, and with a weight sinogram the weighted one
 * (C/Iter3D_Algo_Gradient.cu:227-371): r = y - P x, J_MC = sum r^2, then
 * r = r * W before the backprojection; the step keeps the unweighted
 * |P dJ|^2 of the simple loop, as the reference does.                      */
static int gd_loop(const og_geom* g, const og_gd_params* p, const float* sino, const float* weight, float* vol,
                   og_gd_stat* stats) {
    const size_t nv = (size_t)g->nx * g->ny * g->nz, ns = (size_t)g->nu * g->nv * g->nphi;
    float* est = (float*)malloc(sizeof(float) * ns);
    float* dj = (float*)malloc(sizeof(float) * nv);
    float* t1 = (float*)malloc(sizeof(float) * nv);
    float* t2 = (float*)malloc(sizeof(float) * nv);
    if (!est || !dj || !t1 || !t2) { free(est); free(dj); free(t1); free(t2); return 1; }
    double alpha_min = p->alpha_min_init;
    for (int k = 1; k <= p->iters; ++k) {
        og_gd_stat st = {0, 0, 0, 0, 0, 0};
        project(g, p, vol, est);
        for (size_t i = 0; i < ns; ++i) est[i] = sino[i] - est[i];
        st.residual_sq = sq_norm(est, ns);
        if (weight)
            for (size_t i = 0; i < ns; ++i) est[i] = est[i] * weight[i];
        backproject(g, p, est, dj);
        for (size_t i = 0; i < nv; ++i) dj[i] = dj[i] * -2.0f;
        double norm_reg = 0.0;
        if (p->lambda != 0.0) {
            og_laplacian_f32(g->nx, g->ny, g->nz, vol, t1);
            st.j_reg = sq_norm(t1, nv);
            og_laplacian_f32(g->nx, g->ny, g->nz, t1, t2);
            const float w = (float)(2.0 * (double)(float)p->lambda);
            for (size_t i = 0; i < nv; ++i) dj[i] = dj[i] + w * t2[i];
            if (k < p->kopt) {
                og_laplacian_f32(g->nx, g->ny, g->nz, dj, t1);
                norm_reg = 2.0 * sq_norm(t1, nv);
            }
        }
        double alpha;
        if (k < p->kopt) {
            project(g, p, dj, est);
            st.norm_dj = sq_norm(dj, nv);
            st.norm_dj_proj = 2.0 * sq_norm(est, ns) + p->lambda * norm_reg;
            alpha = -st.norm_dj / st.norm_dj_proj;
            alpha_min = fmax(alpha, alpha_min);
        } else {
            alpha = alpha_min;
        }
        st.alpha = alpha;
        const float af = (float)alpha;
        for (size_t i = 0; i < nv; ++i) {
            float v = vol[i] + af * dj[i];
            if (p->positivity && v < 0.0f) v = 0.0f;
            vol[i] = v;
        }
        if (stats) stats[k - 1] = st;
    }
    free(est); free(dj); free(t1); free(t2);
    return 0;
}

int og_gd_run(const og_geom* g, const og_gd_params* p, const float* sino, float* vol, og_gd_stat* stats) {
    return gd_loop(g, p, sino, NULL, vol, stats);
}

int og_wgd_run(const og_geom* g, const og_gd_params* p, const float* sino, const float* weight, float* vol,
               og_gd_stat* stats) {
    return gd_loop(g, p, sino, weight, vol, stats);
}

static double dot_f32(const float* a, const float* b, size_t n) {
    double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (size_t i = 0; i < n; ++i) s += (double)a[i] * (double)b[i];
    return s;
}

int og_cg_run(const og_geom* g, const og_gd_params* p, const float* sino, float* vol, og_gd_stat* stats) {
    const size_t nv = (size_t)g->nx * g->ny * g->nz, ns = (size_t)g->nu * g->nv * g->nphi;
    float* est = (float*)malloc(sizeof(float) * ns);
    float* gr = (float*)malloc(sizeof(float) * nv);
    float* d = (float*)malloc(sizeof(float) * nv);
    float* t1 = (float*)malloc(sizeof(float) * nv);
    float* t2 = (float*)malloc(sizeof(float) * nv);
    if (!est || !gr || !d || !t1 || !t2) { free(est); free(gr); free(d); free(t1); free(t2); return 1; }
    double gg_prev = 0.0;
    for (int k = 1; k <= p->iters; ++k) {
        og_gd_stat st = {0, 0, 0, 0, 0, 0};
        project(g, p, vol, est);
        for (size_t i = 0; i < ns; ++i) est[i] = sino[i] - est[i];
        st.residual_sq = sq_norm(est, ns);
        backproject(g, p, est, gr);
        for (size_t i = 0; i < nv; ++i) gr[i] = gr[i] * -2.0f;
        if (p->lambda != 0.0) {
            og_laplacian_f32(g->nx, g->ny, g->nz, vol, t1);
            st.j_reg = sq_norm(t1, nv);
            og_laplacian_f32(g->nx, g->ny, g->nz, t1, t2);
            const float w = (float)(2.0 * (double)(float)p->lambda);
            for (size_t i = 0; i < nv; ++i) gr[i] = gr[i] + w * t2[i];
        }
        const double gg = sq_norm(gr, nv);
        if (k == 1) {
            for (size_t i = 0; i < nv; ++i) d[i] = gr[i];
        } else {
            st.beta = gg / gg_prev;
            const float bf = (float)st.beta;
            for (size_t i = 0; i < nv; ++i) d[i] = bf * d[i] + gr[i];
        }
        gg_prev = gg;
        const double gd = dot_f32(gr, d, nv);
        project(g, p, d, est);
        double reg = 0.0;
        if (p->lambda != 0.0) {
            og_laplacian_f32(g->nx, g->ny, g->nz, d, t1);
            reg = 2.0 * sq_norm(t1, nv);
        }
        st.norm_dj = gg;
        st.norm_dj_proj = 2.0 * sq_norm(est, ns) + p->lambda * reg;
        st.alpha = -gd / st.norm_dj_proj;
        const float af = (float)st.alpha;
        for (size_t i = 0; i < nv; ++i) {
            float v = vol[i] + af * d[i];
            if (p->positivity && v < 0.0f) v = 0.0f;
            vol[i] = v;
        }
        if (stats) stats[k - 1] = st;
    }
    free(est); free(gr); free(d); free(t1); free(t2);
    return 0;
}

/* SIRT weights: rw = 1/(P 1) (sinogram), cw = (float)relax / (B 1) (volume),
 * 0 where the sum is not positive.  FP32 arithmetic.                        */
int og_sirt_weights(const og_geom* g, const og_sirt_params* p, float* rw, float* cw) {
    const size_t nv = (size_t)g->nx * g->ny * g->nz, ns = (size_t)g->nu * g->nv * g->nphi;
    float* ones_v = (float*)malloc(sizeof(float) * nv);
    float* ones_s = (float*)malloc(sizeof(float) * ns);
    if (!ones_v || !ones_s) { free(ones_v); free(ones_s); return 1; }
    og_gd_params gp;
    memset(&gp, 0, sizeof(gp));
    gp.pair = p->pair;
    gp.cube = p->cube;
    for (size_t i = 0; i < nv; ++i) ones_v[i] = 1.0f;
    for (size_t i = 0; i < ns; ++i) ones_s[i] = 1.0f;
    project(g, &gp, ones_v, rw);
    backproject(g, &gp, ones_s, cw);
    const float relax = (float)p->relax;
    for (size_t i = 0; i < ns; ++i) rw[i] = rw[i] > 0.0f ? 1.0f / rw[i] : 0.0f;
    for (size_t i = 0; i < nv; ++i) cw[i] = cw[i] > 0.0f ? relax / cw[i] : 0.0f;
    free(ones_v); free(ones_s);
    return 0;
}

int og_sirt_run(const og_geom* g, const og_sirt_params* p, const float* sino, float* vol, double* residual_sq) {
    const size_t nv = (size_t)g->nx * g->ny * g->nz, ns = (size_t)g->nu * g->nv * g->nphi;
    float* rw = (float*)malloc(sizeof(float) * ns);
    float* cw = (float*)malloc(sizeof(float) * nv);
    float* est = (float*)malloc(sizeof(float) * ns);
    float* b = (float*)malloc(sizeof(float) * nv);
    if (!rw || !cw || !est || !b || og_sirt_weights(g, p, rw, cw)) {
        free(rw); free(cw); free(est); free(b);
        return 1;
    }
    og_gd_params gp;
    memset(&gp, 0, sizeof(gp));
    gp.pair = p->pair;
    gp.cube = p->cube;
    for (size_t i = 0; i < nv; ++i) vol[i] = 0.0f;
    for (int k = 0; k < p->iters; ++k) {
        project(g, &gp, vol, est);
        double rs = 0.0;
        for (size_t i = 0; i < ns; ++i) {
            const float d = sino[i] - est[i];
            rs += (double)d * (double)d;
            est[i] = d * rw[i];
        }
        if (residual_sq) residual_sq[k] = rs;
        backproject(g, &gp, est, b);
        for (size_t i = 0; i < nv; ++i) {
            float v = vol[i] + cw[i] * b[i];
            if (p->positivity && v < 0.0f) v = 0.0f;
            vol[i] = v;
        }
    }
    free(rw); free(cw); free(est); free(b);
    return 0;
}
