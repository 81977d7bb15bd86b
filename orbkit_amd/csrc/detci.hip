// Grid reductions of orbkit's detCI module (orbkit/detci/cy_ci.pyx) over MO
// values on the device.  A CI density or flux is a sparse quadratic form in
// the MO values: terms (c_k, a_k, b_k) from identical determinants (a = b)
// and effective single excitations.  One thread per grid point sums its
// terms in the given order (orbkit's order: zero terms, then singles):
//   rho[p]           = sum_k (c_k phi_a(p)) phi_b(p)                       (get_rho)
//   jab[d][p]        = -sum_k 1/2 (c_k (phi_a d_d phi_b - phi_b d_d phi_a))  (get_jab)
//   a_nabla_b[d][p]  = sum_k c_k (phi_a d_d phi_b)                          (get_a_nabla_b)
// Reads are coalesced over points; every MO value of a point is reused from
// the L1/L2 caches across terms.  The time-dependent combinations
// (get_rho_full, get_j_full) are FP64 GEMMs of a coefficient matrix with
// these pair quantities and use okmi_mocreator.
#include "okmi_internal.h"

namespace {

__global__ void k_ci_rho(const double* __restrict__ mo, int64_t ld, int64_t np, const double* coef, const int* ia,
                         const int* ib, int nterm, double* __restrict__ out)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    double s = 0.0;
    for (int k = 0; k < nterm; ++k) {
        const double c = cptr(coef)[k];
        const int a = cptr(ia)[k], b = cptr(ib)[k];
        s += (c * mo[(int64_t)a * ld + p]) * mo[(int64_t)b * ld + p];
    }
    out[p] = s;
}

// mode 0: jab, mode 1: a_nabla_b; dmo is [nd][nmo][ld], out [nd][ld_out]
__global__ void k_ci_flux(const double* __restrict__ mo, const double* __restrict__ dmo, int nmo, int64_t ld,
                          int64_t np, const double* coef, const int* ia, const int* ib, int nterm, int mode,
                          double* __restrict__ out, int64_t ld_out)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int d = blockIdx.y;
    if (p >= np) return;
    const double* dm = dmo + (int64_t)d * nmo * ld;
    double s = 0.0;
    for (int k = 0; k < nterm; ++k) {
        const double c = cptr(coef)[k];
        const int a = cptr(ia)[k], b = cptr(ib)[k];
        const double fa = mo[(int64_t)a * ld + p], fb = mo[(int64_t)b * ld + p];
        if (mode == 0)
            s -= 0.5 * (c * (fa * dm[(int64_t)b * ld + p] - fb * dm[(int64_t)a * ld + p]));
        else
            s += c * (fa * dm[(int64_t)b * ld + p]);
    }
    out[(int64_t)d * ld_out + p] = s;
}

}  // namespace

size_t okmi_ci_ws_bytes(int nterm) { return (size_t)nterm * (sizeof(double) + 2 * sizeof(int)) + 3 * 256; }

int okmi_launch_ci(const double* d_mo, const double* d_dmo, int nmo, int nd, int64_t np, int64_t ld,
                   const double* coef, const int* ia, const int* ib, int nterm, int mode, double* d_out, int64_t ld_out,
                   void* ws, size_t ws_bytes, hipStream_t st)
{
    for (int k = 0; k < nterm; ++k)
        if (ia[k] < 0 || ia[k] >= nmo || ib[k] < 0 || ib[k] >= nmo) {
            okmi_set_error("ci: term %d refers to orbitals (%d, %d) of %d", k, ia[k], ib[k], nmo);
            return 1;
        }
    if (np <= 0) return 0;
    if (nterm > 0 && (!ws || ws_bytes < okmi_ci_ws_bytes(nterm))) {
        okmi_set_error("ci: workspace of %zu bytes is too small (need %zu)", ws_bytes, okmi_ci_ws_bytes(nterm));
        return 1;
    }
    char* w = (char*)(((uintptr_t)ws + 255) & ~(uintptr_t)255);
    double* dc = (double*)w;
    int* da = (int*)(w + (((size_t)nterm * sizeof(double) + 255) & ~(size_t)255));
    int* db = da + (((size_t)nterm * sizeof(int) + 255) & ~(size_t)255) / sizeof(int);
    if (nterm > 0) {
        hipError_t e = hipMemcpyAsync(dc, coef, sizeof(double) * nterm, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(da, ia, sizeof(int) * nterm, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(db, ib, sizeof(int) * nterm, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);   // host arrays may go away after return
        if (e != hipSuccess) return okmi_hip_fail(e, "ci: term lists");
    }
    OkmiProfScope prof("ci", st);
    if (mode < 0)
        hipLaunchKernelGGL(k_ci_rho, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, st, d_mo, ld, np, dc, da, db,
                           nterm, d_out);
    else
        hipLaunchKernelGGL(k_ci_flux, dim3((unsigned)((np + 255) / 256), (unsigned)nd), dim3(256), 0, st, d_mo, d_dmo,
                           nmo, ld, np, dc, da, db, nterm, mode, d_out, ld_out);
    return okmi_check_launch("ci");
}
