/*
 * CPU oracle for the grid-density path.  TEST INFRASTRUCTURE ONLY.
 *
 * This file is the checker, never the product: only tests/, the smoke check
 * and the benchmark's cpu_baseline leg load it.  It restates in plain C99 the
 * FP64 algorithm of orbkit's grid kernel so that the HIP path can be compared
 * with it on identical inputs:
 *
 *   orc_ipow / orc_xyz      orbkit/c_support.c:8-26
 *   orc_ao_xyz              orbkit/c_support.c:28-175 (drv 0-9, including the
 *                           reference's mixed second-derivative formula)
 *   orc_ao_norm, orc_dfact  orbkit/c_support.c:177-195
 *   orc_lcreator            orbkit/c_grid-based.c:9-79 (one contraction)
 *   orc_aocreator           orbkit/cy_core.pyx:49-78 (loop over contractions)
 *   orc_mocreator           orbkit/cy_core.pyx:80-101 (MO[i,j] = sum_k C[i,k] AO[k,j])
 *   orc_rho_slices          orbkit/core.py:179-308 + :504-561, density and
 *                           first/diagonal-second derivatives over slices
 *                           with OpenMP standing in for multiprocessing.Pool.
 *
 * Summation order follows the reference: per Cartesian component a sum over
 * primitives in file order, the monomial multiplied after the sum for drv 0
 * and inside it for drv != 0; the MO sum runs over AOs in index order.
 *
 * Pinning (what checks this file): bit-level comparison with the reference's
 * own c_lcreator compiled from its sources (oracle/_ref), the golden arrays
 * refdata_rho_compute.npz, Gaussian cubegen cubes, closed-form normalisation
 * and central finite differences (tests/test_oracle.py).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int orc_dfact(int n)
{
    int r = 1;
    while (n > 0) { r *= n; n -= 2; }
    return r;
}

double orc_ipow(double b, int e)
{
    double r = 1.0;
    for (; e > 0; e >>= 1) {
        if (e & 1) r *= b;
        b *= b;
    }
    return r;
}

double orc_xyz(double x, double y, double z, int lx, int ly, int lz)
{
    return orc_ipow(x, lx) * orc_ipow(y, ly) * orc_ipow(z, lz);
}

double orc_ao_norm(int lx, int ly, int lz, double a, int is_normalized)
{
    if (is_normalized > 0) return 1.0;
    const double pi = 3.14159265358979323846;
    int l = lx + ly + lz;
    double num = pow(2.0 / pi, 0.75) * (pow(2.0, l) * pow(a, (2.0 * l + 3.0) / 4.0));
    double den = pow((double)(orc_dfact(2 * lx - 1) * orc_dfact(2 * ly - 1) * orc_dfact(2 * lz - 1)), 0.5);
    return num / den;
}

/* first derivative along one axis: d/dq of q^lq exp(-a r^2) times the
 * other two monomials; l[] holds (lx,ly,lz), ax the axis. */
static double d1(double X, double Y, double Z, const int *l, int ax, double a)
{
    int lo[3] = {l[0], l[1], l[2]}, hi[3] = {l[0], l[1], l[2]};
    hi[ax] += 1;
    double up = -2 * a * orc_xyz(X, Y, Z, hi[0], hi[1], hi[2]);
    if (l[ax] == 0) return up;
    lo[ax] -= 1;
    return l[ax] * orc_xyz(X, Y, Z, lo[0], lo[1], lo[2]) + up;
}

/* diagonal second derivative along one axis */
static double d2(double X, double Y, double Z, const int *l, int ax, double a)
{
    double q = ax == 0 ? X : (ax == 1 ? Y : Z);
    double v = 2 * a * orc_xyz(X, Y, Z, l[0], l[1], l[2]) * (2 * a * pow(q, 2) - (2 * l[ax] + 1));
    if (l[ax] >= 2) {
        int lo[3] = {l[0], l[1], l[2]};
        lo[ax] -= 2;
        v += (pow(l[ax], 2) - l[ax]) * orc_xyz(X, Y, Z, lo[0], lo[1], lo[2]);
    }
    return v;
}

/* mixed second derivative with the reference's formula
 * (orbkit/c_support.c:121-168): 4a^2 q_i q_j m + a lowering term chosen by
 * which of l_i, l_j are nonzero.  It omits the -2a l terms, so it is not the
 * analytic mixed derivative for non-s functions; it is kept for parity. */
static double dmix(double X, double Y, double Z, const int *l, int i, int j, double a)
{
    int up[3] = {l[0], l[1], l[2]};
    up[i] += 1; up[j] += 1;
    double v = 4 * pow(a, 2) * orc_xyz(X, Y, Z, up[0], up[1], up[2]);
    int lo[3] = {l[0], l[1], l[2]};
    if (l[i] == 0 && l[j] > 0) {
        lo[j] -= 1;
        v += l[j] * orc_xyz(X, Y, Z, lo[0], lo[1], lo[2]);
    } else if (l[i] > 0 && l[j] == 0) {
        lo[i] -= 1;
        v += l[i] * orc_xyz(X, Y, Z, lo[0], lo[1], lo[2]);
    } else if (l[i] > 0 && l[j] > 0) {
        lo[i] -= 1; lo[j] -= 1;
        v += l[i] * l[j] * orc_xyz(X, Y, Z, lo[0], lo[1], lo[2]);
    }
    return v;
}

double orc_ao_xyz(double X, double Y, double Z, int lx, int ly, int lz, double a, int drv)
{
    const int l[3] = {lx, ly, lz};
    switch (drv) {
    case 0: return orc_xyz(X, Y, Z, lx, ly, lz);
    case 1: return d1(X, Y, Z, l, 0, a);
    case 2: return d1(X, Y, Z, l, 1, a);
    case 3: return d1(X, Y, Z, l, 2, a);
    case 4: return d2(X, Y, Z, l, 0, a);
    case 5: return d2(X, Y, Z, l, 1, a);
    case 6: return d2(X, Y, Z, l, 2, a);
    case 7: return dmix(X, Y, Z, l, 0, 1, a);
    case 8: return dmix(X, Y, Z, l, 0, 2, a);
    case 9: return dmix(X, Y, Z, l, 1, 2, a);
    default: return 0.0;   /* the reference prints a message and yields 0 */
    }
}

/* One contraction on one centre, all its Cartesian components, vector grid.
 * ao[il*ld + i] for il < ao_num, i < npts (ld = row stride). */
void orc_lcreator(double *ao, int64_t ld, const int *lxlylz, const double *coeffs,
                  const double *at, const double *x, const double *y, const double *z,
                  int64_t npts, int ao_num, int pnum, int drv, int is_normalized)
{
    double *nrm = (double *)malloc(sizeof(double) * (size_t)(ao_num * pnum + pnum));
    double *ce = nrm + ao_num * pnum;
    for (int il = 0; il < ao_num; il++)
        for (int p = 0; p < pnum; p++)
            nrm[il * pnum + p] = orc_ao_norm(lxlylz[3 * il], lxlylz[3 * il + 1], lxlylz[3 * il + 2],
                                             coeffs[2 * p], is_normalized);
    for (int64_t i = 0; i < npts; i++) {
        double X = x[i] - at[0], Y = y[i] - at[1], Z = z[i] - at[2];
        double rr = X * X + Y * Y + Z * Z;
        for (int p = 0; p < pnum; p++) ce[p] = coeffs[2 * p + 1] * exp(-coeffs[2 * p] * rr);
        for (int il = 0; il < ao_num; il++) {
            const int *l = lxlylz + 3 * il;
            double s = 0.0;
            if (drv == 0) {
                for (int p = 0; p < pnum; p++) s += nrm[il * pnum + p] * ce[p];
                s *= orc_xyz(X, Y, Z, l[0], l[1], l[2]);
            } else {
                for (int p = 0; p < pnum; p++)
                    s += nrm[il * pnum + p] * ce[p] * orc_ao_xyz(X, Y, Z, l[0], l[1], l[2], coeffs[2 * p], drv);
            }
            ao[il * ld + i] = s;
        }
    }
    free(nrm);
}

/* All contractions: ao[NAO_cart][npts] (row stride npts). */
void orc_aocreator(double *ao, const int *lxlylz, const int *ncart_per_cont,
                   const double *prim_coeffs, const int *nprim_per_cont,
                   const double *geo, const int *cont_atom, int ncont,
                   const double *x, const double *y, const double *z, int64_t npts,
                   int drv, int is_normalized)
{
    int row = 0, prim = 0;
    for (int c = 0; c < ncont; c++) {
        orc_lcreator(ao + (int64_t)row * npts, npts, lxlylz + 3 * row, prim_coeffs + 2 * prim,
                     geo + 3 * cont_atom[c], x, y, z, npts, ncart_per_cont[c], nprim_per_cont[c],
                     drv, is_normalized);
        row += ncart_per_cont[c];
        prim += nprim_per_cont[c];
    }
}

/* mo[i][j] = sum_k C[i][k] ao[k][j], k ascending. */
void orc_mocreator(double *mo, const double *ao, const double *C, int nao, int64_t npts, int nmo)
{
    for (int i = 0; i < nmo; i++)
        for (int64_t j = 0; j < npts; j++) {
            double v = 0.0;
            for (int k = 0; k < nao; k++) v += C[(int64_t)i * nao + k] * ao[(int64_t)k * npts + j];
            mo[(int64_t)i * npts + j] = v;
        }
}

/* Dense real transform out[r][j] = sum_c T[r][c] in[c][j] (used for the
 * Cartesian->spherical map, T from the cart2sph table). */
void orc_transform(double *out, const double *T, const double *in, int nout, int nin, int64_t npts)
{
    for (int r = 0; r < nout; r++) {
        double *o = out + (int64_t)r * npts;
        for (int64_t j = 0; j < npts; j++) o[j] = 0.0;
        for (int c = 0; c < nin; c++) {
            double t = T[(int64_t)r * nin + c];
            if (t == 0.0) continue;
            const double *v = in + (int64_t)c * npts;
            for (int64_t j = 0; j < npts; j++) o[j] += t * v[j];
        }
    }
}

/* One AO set of a slice as orbkit's ao_creator forms it: Cartesian rows,
 * optional per-row factor, optional transform T (spherical). */
static const double *slice_aos(double *ao, double *aos, const int *lxlylz, const int *ncart_per_cont,
                               const double *prim_coeffs, const int *nprim_per_cont, const double *geo,
                               const int *cont_atom, int ncont, int nao_cart, const double *row_scale,
                               const double *T, int nao_out, const double *x, const double *y,
                               const double *z, int64_t n, int drv, int is_normalized)
{
    orc_aocreator(ao, lxlylz, ncart_per_cont, prim_coeffs, nprim_per_cont, geo, cont_atom, ncont, x, y, z, n,
                  drv, is_normalized);
    if (row_scale)
        for (int r = 0; r < nao_cart; r++)
            for (int64_t j = 0; j < n; j++) ao[(int64_t)r * n + j] *= row_scale[r];
    if (!T) return ao;
    orc_transform(aos, T, ao, nao_out, nao_cart, n);
    return aos;
}

/*
 * rho_compute's density workload over slices of slice_len points, with
 * orbkit's loop structure (orbkit/core.py:244-304): AO(0) -> MOs -> rho and
 * mo_norm; then for each derivative code d (1-6) AO(d) -> MOs -> 2 occ phi_d phi,
 * and for d = 4..6 a further AO(d-3) -> MOs -> 2 occ phi_a^2.
 * delta[r][p] for r < ndrv.  Slices run in parallel (OpenMP dynamic schedule)
 * standing in for multiprocessing.Pool.  Returns 0 on success.
 */
int orc_rho_slices(double *rho, double *mo_norm, double *delta, const int *drv, int ndrv,
                   const int *lxlylz, const int *ncart_per_cont, const double *prim_coeffs,
                   const int *nprim_per_cont, const double *geo, const int *cont_atom, int ncont,
                   int nao_cart, const double *row_scale, const double *T, int nao_out,
                   const double *C, const double *occ, int nmo,
                   const double *x, const double *y, const double *z, int64_t npts,
                   int64_t slice_len, int is_normalized, int nthreads)
{
    if (slice_len <= 0) slice_len = npts;
    int64_t nslices = (npts + slice_len - 1) / slice_len;
    for (int i = 0; i < nmo; i++) mo_norm[i] = 0.0;
    for (int r = 0; r < ndrv; r++)
        if (drv[r] < 1 || drv[r] > 6) return 2;
    int fail = 0;
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel num_threads(nthreads)
#endif
    {
        double *part = (double *)calloc((size_t)nmo, sizeof(double));
        double *ao = (double *)malloc(sizeof(double) * (size_t)nao_cart * (size_t)slice_len);
        double *aos = T ? (double *)malloc(sizeof(double) * (size_t)nao_out * (size_t)slice_len) : NULL;
        double *mo = (double *)malloc(sizeof(double) * (size_t)nmo * (size_t)slice_len);
        double *md = (double *)malloc(sizeof(double) * (size_t)nmo * (size_t)slice_len);
        if (!part || !ao || !mo || !md || (T && !aos)) {
#ifdef _OPENMP
#pragma omp atomic write
#endif
            fail = 1;
        } else {
#ifdef _OPENMP
#pragma omp for schedule(dynamic)
#endif
            for (int64_t s = 0; s < nslices; s++) {
                int64_t p0 = s * slice_len, n = npts - p0 < slice_len ? npts - p0 : slice_len;
                const double *a = slice_aos(ao, aos, lxlylz, ncart_per_cont, prim_coeffs, nprim_per_cont, geo,
                                            cont_atom, ncont, nao_cart, row_scale, T, nao_out, x + p0, y + p0,
                                            z + p0, n, 0, is_normalized);
                orc_mocreator(mo, a, C, nao_out, n, nmo);
                for (int64_t j = 0; j < n; j++) rho[p0 + j] = 0.0;
                for (int i = 0; i < nmo; i++) {
                    const double *phi = mo + (int64_t)i * n;
                    double nn = 0.0;
                    for (int64_t j = 0; j < n; j++) {
                        nn += phi[j] * phi[j];
                        rho[p0 + j] += occ[i] * (phi[j] * phi[j]);
                    }
                    part[i] += nn;
                }
                for (int r = 0; r < ndrv; r++) {
                    double *dr = delta + (int64_t)r * npts + p0;
                    for (int64_t j = 0; j < n; j++) dr[j] = 0.0;
                    a = slice_aos(ao, aos, lxlylz, ncart_per_cont, prim_coeffs, nprim_per_cont, geo, cont_atom,
                                  ncont, nao_cart, row_scale, T, nao_out, x + p0, y + p0, z + p0, n, drv[r],
                                  is_normalized);
                    orc_mocreator(md, a, C, nao_out, n, nmo);
                    for (int i = 0; i < nmo; i++)
                        for (int64_t j = 0; j < n; j++)
                            dr[j] += occ[i] * 2 * md[(int64_t)i * n + j] * mo[(int64_t)i * n + j];
                    if (drv[r] >= 4) {
                        a = slice_aos(ao, aos, lxlylz, ncart_per_cont, prim_coeffs, nprim_per_cont, geo,
                                      cont_atom, ncont, nao_cart, row_scale, T, nao_out, x + p0, y + p0, z + p0,
                                      n, drv[r] - 3, is_normalized);
                        orc_mocreator(md, a, C, nao_out, n, nmo);
                        for (int i = 0; i < nmo; i++)
                            for (int64_t j = 0; j < n; j++) {
                                double v = md[(int64_t)i * n + j];
                                dr[j] += occ[i] * 2 * (v * v);
                            }
                    }
                }
            }
#ifdef _OPENMP
#pragma omp critical
#endif
            for (int i = 0; i < nmo; i++) mo_norm[i] += part[i];
        }
        free(part); free(ao); free(aos); free(mo); free(md);
    }
    return fail;
}

int orc_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
