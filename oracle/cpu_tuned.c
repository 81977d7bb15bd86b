/*
 * Tuned CPU form of the grid-density workload.  MEASUREMENT AND TEST
 * INFRASTRUCTURE ONLY: the "tuned" CPU baseline that bench.py times beside
 * the restatement of orbkit's own loops (orc_rho_slices in
 * orbkit_oracle.c); never part of the product.
 *
 * Same results as orbkit's slice_rho (orbkit/core.py:179-308) within FP64
 * rounding, with the redundant work of orbkit's loops removed:
 *   - per contraction and point, each primitive's exponential is computed
 *     once and shared by every Cartesian component and every AO set (orbkit's
 *     c_lcreator recomputes it per component and per derivative code,
 *     orbkit/c_grid-based.c:9-79): the radial sums
 *       S0 = sum_p w_p e_p,  S1 = sum_p -2 a_p w_p e_p,  S2 = sum_p 4 a_p^2 w_p e_p
 *     (e_p = exp(-a_p r^2), w_p the contraction coefficient times the
 *     alpha-dependent part of orbkit's norm, orbkit/c_support.c:177-195)
 *     give every component's value and first / diagonal second derivatives
 *     from its monomials;
 *   - the 'N' row factor and the Cartesian-to-spherical map are folded into
 *     the MO coefficients by the caller (TC = C T), so each AO set costs one
 *     contraction over the Cartesian rows;
 *   - each AO set is evaluated once per slice (orbkit evaluates the first
 *     derivatives a second time for the Laplacian's product terms);
 *   - the MO contraction is a cache-blocked loop vectorised over points.
 * Slices run in parallel (OpenMP dynamic schedule) as orbkit's Pool does
 * over slices (orbkit/core.py:524-536).
 *
 * Derivative codes 1-6 (x, y, z, xx, yy, zz), as orc_rho_slices.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PB 256   /* points per block of the MO contraction */

static int dfact(int n)
{
    int r = 1;
    while (n > 0) { r *= n; n -= 2; }
    return r;
}

/* AO sets `sets` (bit c: derivative code c, 0..6) of every Cartesian row for
 * the n points of one slice: A[c][row][p], row stride n. */
static void slice_sets(double *A, int sets, const int *lxlylz, const int *ncart_per_cont, const double *prim_coeffs,
                       const int *nprim_per_cont, const double *geo, const int *cont_atom, int ncont, int nao_cart,
                       const double *x, const double *y, const double *z, int64_t n, int is_normalized,
                       double *X, double *Y, double *Z, double *s0, double *s1, double *s2)
{
    const double pi = 3.14159265358979323846;
    int row = 0, prim = 0;
    for (int c = 0; c < ncont; c++) {
        const int nc = ncart_per_cont[c], np = nprim_per_cont[c];
        const int *lc = lxlylz + 3 * row;
        const int l = lc[0] + lc[1] + lc[2];
        const double *R = geo + 3 * cont_atom[c];
        for (int64_t p = 0; p < n; p++) {
            X[p] = x[p] - R[0];
            Y[p] = y[p] - R[1];
            Z[p] = z[p] - R[2];
            s0[p] = s1[p] = s2[p] = 0.0;
        }
        for (int q = 0; q < np; q++) {
            const double a = prim_coeffs[2 * (prim + q)];
            double w = prim_coeffs[2 * (prim + q) + 1];
            if (is_normalized <= 0) w *= pow(2.0 / pi, 0.75) * pow(2.0, l) * pow(a, (2.0 * l + 3.0) / 4.0);
            const double w1 = -2.0 * a * w, w2 = 4.0 * a * a * w;
            for (int64_t p = 0; p < n; p++) {
                const double e = exp(-a * (X[p] * X[p] + Y[p] * Y[p] + Z[p] * Z[p]));
                s0[p] += w * e;
                s1[p] += w1 * e;
                s2[p] += w2 * e;
            }
        }
        for (int k = 0; k < nc; k++) {
            const int *e = lc + 3 * k;
            const double f = is_normalized > 0
                                 ? 1.0
                                 : 1.0 / sqrt((double)(dfact(2 * e[0] - 1) * dfact(2 * e[1] - 1) * dfact(2 * e[2] - 1)));
            double *out[7];
            for (int code = 0; code <= 6; code++) out[code] = A + ((int64_t)code * nao_cart + row + k) * n;
            for (int64_t p = 0; p < n; p++) {
                const double q[3] = {X[p], Y[p], Z[p]};
                double m[3], lo1[3], lo2[3];   /* q^e, q^(e-1), q^(e-2) (0 below zero) */
                for (int d = 0; d < 3; d++) {
                    double v = 1.0, v1 = 0.0, v2 = 0.0;
                    for (int i = 0; i < e[d]; i++) { v2 = v1; v1 = v; v *= q[d]; }
                    m[d] = v; lo1[d] = v1; lo2[d] = v2;
                }
                const double mm = m[0] * m[1] * m[2];
                out[0][p] = f * mm * s0[p];
                for (int ax = 0; ax < 3; ++ax) {
                    const double rest = ax == 0 ? m[1] * m[2] : ax == 1 ? m[0] * m[2] : m[0] * m[1];
                    const int la = e[ax];
                    if (sets >> (1 + ax) & 1)
                        out[1 + ax][p] = f * rest * (la * lo1[ax] * s0[p] + m[ax] * q[ax] * s1[p]);
                    if (sets >> (4 + ax) & 1)
                        out[4 + ax][p] = f * rest * (la * (la - 1) * lo2[ax] * s0[p] + (2 * la + 1) * m[ax] * s1[p] +
                                                     m[ax] * q[ax] * q[ax] * s2[p]);
                }
            }
        }
        row += nc;
        prim += np;
    }
}

/* Phi[m][p] = sum_k TC[m][k] A[k][p], blocks of 4 MOs x PB points */
static void contract(double *Phi, const double *A, const double *TC, int nao, int64_t n, int nmo)
{
    double acc[4][PB];
    for (int64_t p0 = 0; p0 < n; p0 += PB) {
        const int nb = (int)(n - p0 < PB ? n - p0 : PB);
        for (int m0 = 0; m0 < nmo; m0 += 4) {
            const int mb = nmo - m0 < 4 ? nmo - m0 : 4;
            for (int i = 0; i < 4; i++)
                for (int j = 0; j < nb; j++) acc[i][j] = 0.0;
            for (int k = 0; k < nao; k++) {
                const double *a = A + (int64_t)k * n + p0;
                const double c0 = TC[(int64_t)m0 * nao + k];
                const double c1 = mb > 1 ? TC[(int64_t)(m0 + 1) * nao + k] : 0.0;
                const double c2 = mb > 2 ? TC[(int64_t)(m0 + 2) * nao + k] : 0.0;
                const double c3 = mb > 3 ? TC[(int64_t)(m0 + 3) * nao + k] : 0.0;
                for (int j = 0; j < nb; j++) {
                    const double v = a[j];
                    acc[0][j] += c0 * v;
                    acc[1][j] += c1 * v;
                    acc[2][j] += c2 * v;
                    acc[3][j] += c3 * v;
                }
            }
            for (int i = 0; i < mb; i++)
                for (int j = 0; j < nb; j++) Phi[(int64_t)(m0 + i) * n + p0 + j] = acc[i][j];
        }
    }
}

/*
 * rho[p] and delta[r][p] (code drv[r] of 1..6) for the npts points, slices of
 * slice_len points in parallel.  TC: nmo x nao_cart, the MO coefficients
 * times the row map (orbkit's 'N' factor and cartesian2spherical).  Returns 0
 * on success.
 */
int orc_rho_slices_tuned(double *rho, double *delta, const int *drv, int ndrv, const int *lxlylz,
                         const int *ncart_per_cont, const double *prim_coeffs, const int *nprim_per_cont,
                         const double *geo, const int *cont_atom, int ncont, int nao_cart, const double *TC,
                         const double *occ, int nmo, const double *x, const double *y, const double *z,
                         int64_t npts, int64_t slice_len, int is_normalized, int nthreads)
{
    if (slice_len <= 0) slice_len = npts;
    const int64_t nslices = (npts + slice_len - 1) / slice_len;
    int sets = 1;
    for (int r = 0; r < ndrv; r++) {
        if (drv[r] < 1 || drv[r] > 6) return 2;
        sets |= 1 << drv[r];
        if (drv[r] >= 4) sets |= 1 << (drv[r] - 3);
    }
    int fail = 0;
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel num_threads(nthreads)
#endif
    {
        const size_t L = (size_t)slice_len;
        double *A = (double *)malloc(sizeof(double) * 7 * (size_t)nao_cart * L);
        double *Phi = (double *)malloc(sizeof(double) * 7 * (size_t)nmo * L);
        double *tmp = (double *)malloc(sizeof(double) * 6 * L);
        if (!A || !Phi || !tmp) {
#ifdef _OPENMP
#pragma omp atomic write
#endif
            fail = 1;
        } else {
#ifdef _OPENMP
#pragma omp for schedule(dynamic)
#endif
            for (int64_t s = 0; s < nslices; s++) {
                const int64_t p0 = s * slice_len, n = npts - p0 < slice_len ? npts - p0 : slice_len;
                slice_sets(A, sets, lxlylz, ncart_per_cont, prim_coeffs, nprim_per_cont, geo, cont_atom, ncont,
                           nao_cart, x + p0, y + p0, z + p0, n, is_normalized, tmp, tmp + L, tmp + 2 * L,
                           tmp + 3 * L, tmp + 4 * L, tmp + 5 * L);
                for (int c = 0; c <= 6; c++)
                    if (sets >> c & 1)
                        contract(Phi + (size_t)c * nmo * n, A + (size_t)c * nao_cart * n, TC, nao_cart, n, nmo);
                for (int64_t p = 0; p < n; p++) rho[p0 + p] = 0.0;
                for (int r = 0; r < ndrv; r++)
                    for (int64_t p = 0; p < n; p++) delta[(int64_t)r * npts + p0 + p] = 0.0;
                for (int i = 0; i < nmo; i++) {
                    const double w = occ[i];
                    const double *f0 = Phi + (int64_t)i * n;
                    for (int64_t p = 0; p < n; p++) rho[p0 + p] += w * (f0[p] * f0[p]);
                    for (int r = 0; r < ndrv; r++) {
                        const int c = drv[r];
                        const double *fc = Phi + ((int64_t)c * nmo + i) * n;
                        double *dr = delta + (int64_t)r * npts + p0;
                        if (c >= 4) {
                            const double *fa = Phi + ((int64_t)(c - 3) * nmo + i) * n;
                            for (int64_t p = 0; p < n; p++) dr[p] += 2.0 * w * (f0[p] * fc[p] + fa[p] * fa[p]);
                        } else {
                            for (int64_t p = 0; p < n; p++) dr[p] += 2.0 * w * (f0[p] * fc[p]);
                        }
                    }
                }
            }
        }
        free(A); free(Phi); free(tmp);
    }
    return fail;
}
