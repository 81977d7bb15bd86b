// Grid analysis helpers that reuse the AO and MO kernels:
//   gross atomic densities   rho^a(r) = sum_i occ_i phi_i(r) phi_i^a(r),
//                            phi_i^a = sum_{mu on a} C_i,mu chi_mu
//                            (orbkit/extras.py:306-385).  Evaluated as
//                            rho^a = sum_{mu on a} chi_mu (D chi)_mu with the
//                            density matrix D = C^T diag(occ) C: one FP64
//                            MFMA GEMM per chunk (k_gemm_f64) and a per-atom
//                            segmented reduction (k_atom_reduce);
//   MO pair products         m[d][n][m][p] = bra[n][p] ket[d][m][p]
//                            (orbkit/core.py:841-941, k_mo_pairs);
//   transition flux          j[d][k][p] = -1/2 (phi_i d_d phi_j - phi_j d_d phi_i)
//                            for pairs k = (i, j) (orbkit/extras.py:441-492,
//                            k_jmo).
// All FP64; the pair kernels are HBM-write-bound (coalesced over points).
#include "okmi_internal.h"

namespace {

// D[mu][nu] = sum_i occ_i C[i][mu] C[i][nu]
__global__ void k_density_matrix(const double* C, const double* occ, int nmo, int nao, double* D)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)nao * nao) return;
    const int mu = (int)(e / nao), nu = (int)(e % nao);
    double s = 0.0;
    for (int i = 0; i < nmo; ++i) s = fma(occ[i] * C[(int64_t)i * nao + mu], C[(int64_t)i * nao + nu], s);
    D[e] = s;
}

// out[j][p] = sum_{rows r of selected atom j} A[r][p] Y[r][p]; rows of atom j
// are rows[start[j] .. start[j+1])
__global__ void k_atom_reduce(const double* A, const double* Y, int64_t ld, int64_t np, const int* start,
                              const int* rows, int nsel, double* out, int64_t ld_out)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (p >= np || j >= nsel) return;
    double s = 0.0;
    for (int k = start[j]; k < start[j + 1]; ++k) {
        const int64_t r = rows[k];
        s = fma(A[r * ld + p], Y[r * ld + p], s);
    }
    out[(int64_t)j * ld_out + p] = s;
}

// out[d][n][m][p] = bra[n][p] * ket[d][m][p]
__global__ void k_mo_pairs(const double* bra, int na, const double* ket, int nb, int nd, int64_t np, int64_t ld_in,
                           double* out, int64_t ld_out)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int dn = blockIdx.y;   // d * na + n
    if (p >= np) return;
    const int d = dn / na, n = dn % na;
    const double b = bra[(int64_t)n * ld_in + p];
    const double* k = ket + (int64_t)d * nb * ld_in + p;
    double* o = out + ((int64_t)dn * nb) * ld_out + p;
    for (int m = 0; m < nb; ++m) o[(int64_t)m * ld_out] = b * k[(int64_t)m * ld_in];
}

// out[d][k][p] = -1/2 (mo[0][i][p] mo[1+d][j][p] - mo[0][j][p] mo[1+d][i][p]),
// mo = [1 + nd][nmo][ld_in] (values, then the nd derivative sets)
__global__ void k_jmo(const double* mo, int nmo, int nd, int64_t np, int64_t ld_in, const int* pairs, int npair,
                      double* out, int64_t ld_out)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int dk = blockIdx.y;
    if (p >= np) return;
    const int d = dk / npair, k = dk % npair;
    const int i = pairs[2 * k], j = pairs[2 * k + 1];
    const double* v = mo;
    const double* g = mo + (int64_t)(1 + d) * nmo * ld_in;
    out[(int64_t)dk * ld_out + p] =
        -0.5 * (v[(int64_t)i * ld_in + p] * g[(int64_t)j * ld_in + p] - v[(int64_t)j * ld_in + p] * g[(int64_t)i * ld_in + p]);
}

}  // namespace

namespace {
// y[j][p] = sum_i occ_i M[i][p] sum_{r in rows} C[i][r] A[r][p]: the gross
// atomic density from caller-supplied AO and MO lists (orbkit's ao_list /
// mo_list arguments), which need not be consistent with C
__global__ void k_masked_mo_product(const double* C, const double* occ, int nmo, int nao, const double* A,
                                    const double* M, int64_t np, int64_t ld, const int* rows, int nrows, double* out)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    double s = 0.0;
    for (int i = 0; i < nmo; ++i) {
        double f = 0.0;
        for (int k = 0; k < nrows; ++k) {
            const int r = rows[k];
            f = fma(C[(int64_t)i * nao + r], A[(int64_t)r * ld + p], f);
        }
        s = fma(occ[i] * M[(int64_t)i * ld + p], f, s);
    }
    out[p] = s;
}
}  // namespace

size_t okmi_gross_ws_bytes(const OkmiBasis* b, int64_t chunk, int nrows_total, int nsel)
{
    return sizeof(double) * ((size_t)b->nao * b->nao + 2 * (size_t)b->nao * chunk) +
           sizeof(int) * ((size_t)nrows_total + nsel + 1) + 4 * 256;
}

int okmi_launch_gross_atomic(const OkmiBasis* b, const GridDesc& g, int64_t p0, int64_t np, int drv, const double* d_c,
                             const double* d_occ, int nmo, const int* sel_start, const int* sel_rows, int nsel,
                             double* d_out, int64_t ld_out, void* ws, size_t ws_bytes, hipStream_t st)
{
    if (np <= 0 || nsel <= 0) return 0;
    const int nrows = sel_start[nsel];
    for (int k = 0; k < nrows; ++k)
        if (sel_rows[k] < 0 || sel_rows[k] >= b->nao) {
            okmi_set_error("gross_atomic_density: AO row %d outside 0..%d", sel_rows[k], b->nao - 1);
            return 1;
        }
    const size_t fixed = okmi_gross_ws_bytes(b, 0, nrows, nsel);
    const size_t per = sizeof(double) * 2 * (size_t)b->nao;
    if (!ws || ws_bytes < fixed + per * 64) {
        okmi_set_error("gross_atomic_density: workspace of %zu bytes is too small (need at least %zu)", ws_bytes,
                       fixed + per * 64);
        return 1;
    }
    int64_t chunk = (int64_t)((ws_bytes - fixed) / per);
    if (chunk >= np) chunk = np;
    else chunk &= ~(int64_t)63;
    char* w = (char*)(((uintptr_t)ws + 255) & ~(uintptr_t)255);
    double* D = (double*)w;
    w += ((sizeof(double) * (size_t)b->nao * b->nao + 255) & ~(size_t)255);
    int* d_start = (int*)w;
    w += ((sizeof(int) * (size_t)(nsel + 1) + 255) & ~(size_t)255);
    int* d_rows = (int*)w;
    w += ((sizeof(int) * (size_t)std::max(1, nrows) + 255) & ~(size_t)255);
    double* A = (double*)w;
    double* Y = A + (size_t)b->nao * chunk;
    hipError_t e = hipMemcpyAsync(d_start, sel_start, sizeof(int) * (nsel + 1), hipMemcpyHostToDevice, st);
    if (e == hipSuccess && nrows > 0)
        e = hipMemcpyAsync(d_rows, sel_rows, sizeof(int) * nrows, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);   // host arrays may go away after return
    if (e != hipSuccess) return okmi_hip_fail(e, "gross_atomic_density: row lists");
    OkmiProfScope prof("gross_atomic", st);
    hipLaunchKernelGGL(k_density_matrix, dim3((unsigned)(((int64_t)b->nao * b->nao + 255) / 256)), dim3(256), 0, st,
                       d_c, d_occ, nmo, b->nao, D);
    for (int64_t c0 = 0; c0 < np; c0 += chunk) {
        const int64_t n = np - c0 < chunk ? np - c0 : chunk;
        int rc = okmi_launch_ao_eval(b, g, p0 + c0, n, &drv, 1, OKMI_F64, A, chunk, 0, st);
        if (!rc) rc = okmi_launch_gemm_f64(A, chunk, 0, 1, b->nao, D, b->nao, n, Y, chunk, st);
        if (rc) return rc;
        hipLaunchKernelGGL(k_atom_reduce, dim3((unsigned)((n + 255) / 256), (unsigned)nsel), dim3(256), 0, st, A, Y,
                           chunk, n, d_start, d_rows, nsel, d_out + c0, ld_out);
    }
    return okmi_check_launch("gross_atomic_density");
}

int okmi_launch_masked_mo_product(const double* d_c, const double* d_occ, int nmo, int nao, const double* d_ao,
                                  const double* d_mo, int64_t np, int64_t ld, const int* d_rows, int nrows,
                                  double* d_out, hipStream_t st)
{
    if (np <= 0) return 0;
    hipLaunchKernelGGL(k_masked_mo_product, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, st, d_c, d_occ, nmo, nao,
                       d_ao, d_mo, np, ld, d_rows, nrows, d_out);
    return okmi_check_launch("masked_mo_product");
}

int okmi_launch_mo_pairs(const double* d_bra, int na, const double* d_ket, int nb, int nd, int64_t np, int64_t ld_in,
                         double* d_out, int64_t ld_out, hipStream_t st)
{
    if (np <= 0 || na <= 0 || nb <= 0 || nd <= 0) return 0;
    OkmiProfScope prof("mo_pairs", st);
    hipLaunchKernelGGL(k_mo_pairs, dim3((unsigned)((np + 255) / 256), (unsigned)(nd * na)), dim3(256), 0, st, d_bra,
                       na, d_ket, nb, nd, np, ld_in, d_out, ld_out);
    return okmi_check_launch("mo_pair_products");
}

int okmi_launch_jmo(const double* d_mo, int nmo, int nd, int64_t np, int64_t ld_in, const int* d_pairs, int npair,
                    double* d_out, int64_t ld_out, hipStream_t st)
{
    if (np <= 0 || npair <= 0 || nd <= 0) return 0;
    OkmiProfScope prof("jmo", st);
    hipLaunchKernelGGL(k_jmo, dim3((unsigned)((np + 255) / 256), (unsigned)(nd * npair)), dim3(256), 0, st, d_mo, nmo,
                       nd, np, ld_in, d_pairs, npair, d_out, ld_out);
    return okmi_check_launch("transition_flux");
}
