/*
 * oracle_compat.c - the reference C path of the Siddon projector, restated
 * with its arithmetic types and quirks.  TEST INFRASTRUCTURE ONLY.
 *
 * Follows SiddonProjector_compute_C_mem_CPU::doProjection
 * (src/TomoGPI_lib/src/CPP/SiddonProjector_compute_C_mem_CPU.cpp:20-276),
 * compiled there by g++ for x86-64 without FMA contraction:
 *   - geometry scalars are float copies of the reference's getters (:31-45);
 *   - source at O + FOD(a,b,0), detector at O + (-a(FDD-FOD) - s b,
 *     -b(FDD-FOD) + s a, t) (:76-96), i.e. mirrored with respect to RS/VI;
 *   - cylinder entry/exit lambda_min/max, C evaluated in double (:99-111);
 *   - first voxel floor(P-0.5) / ceil(P+0.5) (half a voxel behind the
 *     entry) and first crossing evaluated in double (:120-206);
 *   - while (lambda <= lambda_max): chord = min - lambda, last chord not
 *     clipped, indices taken % 256 with unsigned wrap, no z clip (:213-258);
 *   - output index (u%256) + (v%256) Nu + phi Nu Nv (:260-265).
 * Only defined for a 256^3 volume and a detector of at most 256x256 cells.
 * Build with -ffp-contract=off.  An iteration cap of 4(nx+ny+nz)+16 bounds
 * every ray (never reached on a valid geometry).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int og_siddon_compat(const og_geom* g, const float* vol, float* sino, int nthr) {
    if (g->nx != 256 || g->ny != 256 || g->nz != 256 || g->nu > 256 || g->nv > 256) return 1;
    const float fdd = (float)g->fdd, fod = (float)g->fod;
    const float dl = g->delta;
    const float xlr = g->cx * dl, ylr = g->cy * dl, zlr = g->cz * dl;
    const unsigned long nx = (unsigned long)g->nx, ny = (unsigned long)g->ny, nz = (unsigned long)g->nz;
    const unsigned long nu = (unsigned long)g->nu, nv = (unsigned long)g->nv;
    const int cap = 4 * (g->nx + g->ny + g->nz) + 16;
    const int np = g->nphi;
#ifdef _OPENMP
    const int nt = nthr > 0 ? nthr : omp_get_max_threads();
#else
    const int nt = 1;
    (void)nthr;
#endif
#pragma omp parallel for collapse(2) schedule(dynamic, 1) num_threads(nt)
    for (int phi = 0; phi < np; ++phi)
        for (unsigned long vn = 0; vn < nv; ++vn) {
            const float al = g->alpha_f[phi], be = g->beta_f[phi];
            const float xs = xlr + fod * al, ys = ylr + fod * be, zs = zlr;
            for (unsigned long un = 0; un < nu; ++un) {
                const float s = ((float)un - g->cu) * g->du;
                const float t = ((float)vn - g->cv) * g->dv;
                const float xd = (xlr - al * (fdd - fod)) - s * be;
                const float yd = (ylr - be * (fdd - fod)) + s * al;
                const float zd = zlr + t;
                const float L = sqrtf((fdd * fdd + s * s) + t * t);
                const float A = (fdd * fdd + s * s) / (L * L);
                const float B = (-fod * fdd) / L;
                const float C = (float)((double)(fod * fod) -
                                        0.25 * (double)(float)nx * (double)(float)nx * (double)dl * (double)dl);
                const float dlam = B * B - A * C;
                float ray = 0.0f;
                if (dlam > 0.0f) {
                    const float lmin = (-B - sqrtf(dlam)) / A;
                    const float lmax = (-B + sqrtf(dlam)) / A;
                    const float P[3] = {(xs + (lmin * (xd - xs) / L)) / dl,
                                        (ys + (lmin * (yd - ys) / L)) / dl,
                                        (zs + (lmin * (zd - zs) / L)) / dl};
                    const float S[3] = {xs, ys, zs}, D[3] = {xd, yd, zd};
                    const unsigned long N[3] = {nx, ny, nz};
                    unsigned long idx[3];
                    float lam[3], inc[3];
                    int up[3];
                    for (int a = 0; a < 3; ++a) {
                        up[a] = S[a] <= D[a];
                        float pm = up[a] ? (float)((double)P[a] - 0.5) : (float)((double)P[a] + 0.5);
                        if (pm < 0.0f) idx[a] = 0;
                        else if (pm > (float)(N[a] - 1)) idx[a] = N[a] - 1;
                        else idx[a] = up[a] ? (unsigned long)floorf(pm) : (unsigned long)ceilf(pm);
                        const float span = D[a] - S[a];
                        if (fabsf(span) > dl) {
                            const double off = up[a] ? 0.5 : -0.5;
                            lam[a] = (float)((double)L * (((double)(float)idx[a] + off) * (double)dl - (double)S[a]) /
                                             (double)span);
                        } else {
                            lam[a] = L;
                        }
                        inc[a] = (L / fabsf(span)) * dl;
                    }
                    float lambda = lmin;
                    int it = 0;
                    while (lambda <= lmax && it++ < cap) {
                        float lk = lam[0];
                        if (lam[1] < lk) lk = lam[1];
                        if (lam[2] < lk) lk = lam[2];
                        const float chord = lk - lambda;
                        const unsigned long o = (idx[0] % 256) + (idx[1] % 256) * nx + (idx[2] % 256) * nx * ny;
                        const float prod = chord * vol[o];
                        ray = ray + prod;
                        lambda = lk;
                        for (int a = 0; a < 3; ++a) {
                            if (lam[a] <= lk) {
                                lam[a] = lam[a] + inc[a];
                                idx[a] = up[a] ? idx[a] + 1 : idx[a] - 1;
                            }
                        }
                    }
                }
                sino[(un % 256) + (vn % 256) * nu + (unsigned long)phi * nu * nv] = ray;
            }
        }
    return 0;
}
