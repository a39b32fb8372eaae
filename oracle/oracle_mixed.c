/*
 * oracle_mixed.c - the FP16-arithmetic form of the VI backprojector.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The mixed-precision VI on FP16-stored sinograms defines its own arithmetic
 * (the reference's half backprojector, K/BackProjector/
 * backprojection3D_kernel_16reg_half_UM.cuh:23-237, reads FP16 textures as
 * float): per voxel-view term, with the geometry of the FP32 recipe
 * (K/BackProjector/backprojection3D_kernel_16reg_UM.cuh:74-78,103-107),
 *   fu_h = half(fu)                                  (round to nearest even)
 *   a    = half(fu_h * half(g(u+1,v)   - g(u,v))   + g(u,v))      fused
 *   b    = half(fu_h * half(g(u+1,v+1) - g(u,v+1)) + g(u,v+1))    fused
 *   r    = fmaf(fv, (float)half(b - a), (float)a)    (FP32)
 *   acc += r (FP32);  f = acc * delta
 * with taps off the detector 0.  half() rounds a real number to binary16
 * (11 significant bits, subnormals from 2^-24, overflow to infinity at
 * 65520), which og_round_half restates exactly from the IEEE 754 definition.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>

/* Round x to the nearest binary16 value, ties to even.  Every product and sum
 * passed here is exact in double (operands are binary16 values), so one
 * rounding of the exact result is the fused / correctly rounded operation. */
double og_round_half(double x) {
    if (x == 0.0 || !isfinite(x)) return x;
    int e;
    (void)frexp(x, &e);              /* |x| in [2^(e-1), 2^e) */
    int qe = e - 11;                 /* quantum of an 11-bit significand */
    if (qe < -24) qe = -24;          /* subnormal spacing */
    const double q = ldexp(1.0, qe);
    const double r = nearbyint(x / q) * q;   /* default mode: ties to even */
    if (fabs(r) >= 65520.0) return copysign(INFINITY, x);
    return r;
}

static inline double tap_h(const float* gp, int nu, int nv, int iu, int iv) {
    return (iu >= 0 && iu < nu && iv >= 0 && iv < nv) ? (double)gp[(size_t)iv * nu + iu] : 0.0;
}

void og_vi_h16(const og_geom* g, const float* sino, float* vol) {
    const int nx = g->nx, ny = g->ny, nz = g->nz, nu = g->nu, nv = g->nv, np = g->nphi;
    const double dl = (double)g->delta, ku = g->fdd / (double)g->du, kv = g->fdd / (double)g->dv;
    const float cv = g->cv;
#pragma omp parallel
    {
        float* acc = (float*)malloc(sizeof(float) * nz);
#pragma omp for collapse(2) schedule(static)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const double xp = (double)i - (double)g->cx, yp = (double)j - (double)g->cy;
                for (int k = 0; k < nz; ++k) acc[k] = 0.0f;
                for (int p = 0; p < np; ++p) {
                    const double c0 = dl * g->cphi[p], c1 = dl * g->sphi[p];
                    const double w = fma(c0, xp, fma(c1, yp, g->fod));
                    const double inv = 1.0 / w;
                    const double un = fma(-(dl * g->sphi[p]) * ku, xp, ((dl * g->cphi[p]) * ku) * yp);
                    const float u = (float)fma(un, inv, (double)g->cu);
                    const float A = (float)((dl * kv) * inv);
                    if (!isfinite(u) || !isfinite(A)) { /* reads off the detector: adds +0 */
                        for (int k = 0; k < nz; ++k) acc[k] += 0.0f;
                        continue;
                    }
                    const float uf = floorf(u);
                    const float fu = u - uf;
                    const double fuh = og_round_half((double)fu);
                    const float* gp = sino + (size_t)p * nu * nv;
                    for (int k = 0; k < nz; ++k) {
                        const float zp = (float)((double)k - (double)g->cz);
                        const float v = fmaf(zp, A, cv);
                        const float vf = floorf(v);
                        const float fv = v - vf;
                        /* corners outside [-1, n-1] read four zeros (the result is then +0) */
                        if (!(uf >= -1.0f && uf <= (float)(nu - 1) && vf >= -1.0f && vf <= (float)(nv - 1))) {
                            acc[k] += 0.0f;
                            continue;
                        }
                        const int iu = (int)uf, iv = (int)vf;
                        const double t00 = tap_h(gp, nu, nv, iu, iv), t10 = tap_h(gp, nu, nv, iu + 1, iv);
                        const double t01 = tap_h(gp, nu, nv, iu, iv + 1), t11 = tap_h(gp, nu, nv, iu + 1, iv + 1);
                        const double a = og_round_half(fuh * og_round_half(t10 - t00) + t00);
                        const double b = og_round_half(fuh * og_round_half(t11 - t01) + t01);
                        acc[k] += fmaf(fv, (float)og_round_half(b - a), (float)a);
                    }
                }
                for (int k = 0; k < nz; ++k)
                    vol[(size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k)] = acc[k] * g->delta;
            }
        free(acc);
    }
}
