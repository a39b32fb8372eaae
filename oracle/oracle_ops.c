/*
 * oracle_ops.c - instantiates the operators for the FP32 recipe and for
 * FP64, and adds the FP64 exact transposes and the work counters.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define R float
#define SFX f32
#define RFMA fmaf
#define RFLOOR floorf
#define RMIN fminf
#define RSQRT sqrtf
#include "oracle_ops.inc"
#undef R
#undef SFX
#undef RFMA
#undef RFLOOR
#undef RMIN
#undef RSQRT

#define R double
#define SFX f64
#define RFMA fma
#define RFLOOR floor
#define RMIN fmin
#define RSQRT sqrt
#include "oracle_ops.inc"
#undef R
#undef SFX
#undef RFMA
#undef RFLOOR
#undef RMIN
#undef RSQRT

static int nthreads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
static int thread_id(void) {
#ifdef _OPENMP
    return omp_get_thread_num();
#else
    return 0;
#endif
}

/* Per-thread private volumes, summed at the end (scatter transposes). */
static double* private_volumes(size_t nvox, int* nt) {
    *nt = nthreads();
    return (double*)calloc((size_t)*nt * nvox, sizeof(double));
}
static void reduce_volumes(const double* priv, int nt, size_t nvox, double* out) {
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < nvox; ++i) {
        double s = 0.0;
        for (int t = 0; t < nt; ++t) s += priv[(size_t)t * nvox + i];
        out[i] = s;
    }
}

/* VI^T: g_p[tap] += delta * w_tap(i,j,k,p) * f(i,j,k) with the FP64
 * bilinear weights of og_vi_f64.  Views are independent, so threads split
 * views and write disjoint projections.                                   */
void og_vi_T_f64(const og_geom* g, const double* vol, double* sino) {
    const int nx = g->nx, ny = g->ny, nz = g->nz, nu = g->nu, nv = g->nv, np = g->nphi;
    const double dl = (double)g->delta, ku = g->fdd / (double)g->du, kv = g->fdd / (double)g->dv;
#pragma omp parallel for schedule(dynamic, 1)
    for (int p = 0; p < np; ++p) {
        double* gp = sino + (size_t)p * nu * nv;
        memset(gp, 0, sizeof(double) * (size_t)nu * nv);
        const double c0 = dl * g->cphi[p], c1 = dl * g->sphi[p];
        const double c2 = -(dl * g->sphi[p]) * ku, c3 = (dl * g->cphi[p]) * ku, c4 = dl * kv;
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const double xp = (double)i - (double)g->cx, yp = (double)j - (double)g->cy;
                const double w = fma(c0, xp, fma(c1, yp, g->fod));
                const double inv = 1.0 / w;
                const double u = fma(fma(c2, xp, c3 * yp), inv, (double)g->cu);
                const double A = c4 * inv;
                const double uf = floor(u);
                if (!(uf >= -1.0 && uf <= (double)nu)) continue;
                const int iu = (int)uf;
                const double fu = u - uf;
                for (int k = 0; k < nz; ++k) {
                    const double v = fma((double)k - (double)g->cz, A, (double)g->cv);
                    const double vf = floor(v);
                    if (!(vf >= -1.0 && vf <= (double)nv)) continue;
                    const int iv = (int)vf;
                    const double fv = v - vf;
                    const double val = dl * vol[(size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k)];
                    const double w00 = (1.0 - fu) * (1.0 - fv), w01 = fu * (1.0 - fv);
                    const double w10 = (1.0 - fu) * fv, w11 = fu * fv;
                    if (iv >= 0 && iv < nv) {
                        if (iu >= 0 && iu < nu) gp[(size_t)iv * nu + iu] += w00 * val;
                        if (iu + 1 >= 0 && iu + 1 < nu) gp[(size_t)iv * nu + iu + 1] += w01 * val;
                    }
                    if (iv + 1 >= 0 && iv + 1 < nv) {
                        if (iu >= 0 && iu < nu) gp[(size_t)(iv + 1) * nu + iu] += w10 * val;
                        if (iu + 1 >= 0 && iu + 1 < nu) gp[(size_t)(iv + 1) * nu + iu + 1] += w11 * val;
                    }
                }
            }
    }
}

/* RS^T: f[tap] += delta * w_tap * g(ray) for every sample of every ray. */
void og_rs_T_f64(const og_geom* g, int cube, const double* sino, double* vol) {
    const int nx = g->nx, ny = g->ny, nz = g->nz, nu = g->nu, nv = g->nv, np = g->nphi;
    const size_t nvox = (size_t)nx * ny * nz, sy = (size_t)nx, sz = (size_t)nx * ny;
    int nt;
    double* priv = private_volumes(nvox, &nt);
#pragma omp parallel for collapse(2) schedule(dynamic, 4)
    for (int p = 0; p < np; ++p)
        for (int n = 0; n < nv; ++n) {
            double* f = priv + (size_t)thread_id() * nvox;
            for (int m = 0; m < nu; ++m) {
                const double gv = sino[(size_t)m + (size_t)nu * ((size_t)n + (size_t)nv * p)];
                double e[3], d[3];
                const int64_t Lr = og_rs_setup(g, cube, p, m, n, e, d);
                if (Lr <= 0 || gv == 0.0) continue;
                const double val = (double)g->delta * gv;
                for (int64_t s = 0; s < Lr; ++s) {
                    const double x = fma((double)s, d[0], e[0]), y = fma((double)s, d[1], e[1]), z = fma((double)s, d[2], e[2]);
                    const double xf = floor(x), yf = floor(y), zf = floor(z);
                    if (!(xf >= -1.0 && xf <= nx && yf >= -1.0 && yf <= ny && zf >= -1.0 && zf <= nz)) continue;
                    const int ix = (int)xf, iy = (int)yf, iz = (int)zf;
                    const double fx = x - xf, fy = y - yf, fz = z - zf;
                    for (int c = 0; c < 8; ++c) {
                        const int ox = c & 1, oy = (c >> 1) & 1, oz = (c >> 2) & 1;
                        const int X = ix + ox, Y = iy + oy, Z = iz + oz;
                        if (X < 0 || X >= nx || Y < 0 || Y >= ny || Z < 0 || Z >= nz) continue;
                        const double w = (ox ? fx : 1.0 - fx) * (oy ? fy : 1.0 - fy) * (oz ? fz : 1.0 - fz);
                        f[(size_t)X + sy * Y + sz * Z] += w * val;
                    }
                }
            }
        }
    reduce_volumes(priv, nt, nvox, vol);
    free(priv);
}

uint64_t og_rs_count(const og_geom* g, int cube) {
    uint64_t total = 0;
#pragma omp parallel for collapse(2) reduction(+ : total) schedule(dynamic, 4)
    for (int p = 0; p < g->nphi; ++p)
        for (int n = 0; n < g->nv; ++n)
            for (int m = 0; m < g->nu; ++m) {
                double e[3], d[3];
                const int64_t Lr = og_rs_setup(g, cube, p, m, n, e, d);
                if (Lr > 0) total += (uint64_t)Lr;
            }
    return total;
}

/* Siddon^T: f[voxel] += chord * g(ray), FP64 traversal. */
void og_siddon_T_f64(const og_geom* g, const double* sino, double* vol) {
    const int nu = g->nu, nv = g->nv, np = g->nphi;
    const size_t nvox = (size_t)g->nx * g->ny * g->nz;
    int nt;
    double* priv = private_volumes(nvox, &nt);
#pragma omp parallel for collapse(2) schedule(dynamic, 4)
    for (int p = 0; p < np; ++p)
        for (int n = 0; n < nv; ++n) {
            double* f = priv + (size_t)thread_id() * nvox;
            for (int m = 0; m < nu; ++m) {
                const double gv = sino[(size_t)m + (size_t)nu * ((size_t)n + (size_t)nv * p)];
                og_sid_ray r;
                if (gv == 0.0 || !og_sid_setup(g, p, m, n, &r)) continue;
                sid_trace_f64(g, &r, NULL, f, gv, NULL);
            }
        }
    reduce_volumes(priv, nt, nvox, vol);
    free(priv);
}

/* Positive chords of the FP32-recipe traversal (the traversal the kernels run). */
uint64_t og_siddon_count(const og_geom* g) {
    uint64_t total = 0;
#pragma omp parallel for collapse(2) reduction(+ : total) schedule(dynamic, 4)
    for (int p = 0; p < g->nphi; ++p)
        for (int n = 0; n < g->nv; ++n)
            for (int m = 0; m < g->nu; ++m) {
                og_sid_ray r;
                uint64_t c = 0;
                if (og_sid_setup(g, p, m, n, &r)) sid_trace_f32(g, &r, NULL, NULL, 0.0f, &c);
                total += c;
            }
    return total;
}
