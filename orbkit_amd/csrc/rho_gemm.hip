// Host entries of the FP64 MO contraction kernels (rho_gemm_impl.h).
#include "rho_gemm_impl.h"

namespace {
// out[m][k] = C[m][colmap[k]] (0 where colmap[k] < 0): C in the fused row order
__global__ void k_permute_c(const double* C, int nmo, int nao, int64_t kpad, const int* colmap, double* out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)nmo * kpad) return;
    const int64_t m = i / kpad, k = i - m * kpad;
    const int src = colmap[k];
    out[i] = src >= 0 ? C[m * nao + src] : 0.0;
}
}  // namespace

int okmi_launch_permute_c(const double* C, int nmo, int nao, int64_t kpad, const int* colmap, double* out,
                          hipStream_t st)
{
    const int64_t n = (int64_t)nmo * kpad;
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_permute_c, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, C, nmo, nao, kpad, colmap,
                       out);
    return 0;
}

int okmi_launch_gemm_f64(const double* A, int64_t ld, int64_t set_stride, int nsets, int nao, const double* C,
                         int nmo, int64_t np, double* mo_out, int64_t ld_mo, hipStream_t st)
{
    if (np <= 0 || nmo <= 0) return 0;
    if (nsets < 1 || nsets > OKMI_MAX_SETS) {
        okmi_set_error("gemm_f64: %d AO sets (1..%d supported)", nsets, OKMI_MAX_SETS);
        return 1;
    }
    okmi_gemm::GemmArgs a{};
    a.A = A; a.ld = ld; a.set_stride = set_stride; a.nao = nao;
    a.C = C; a.nmo = nmo; a.np = np;
    a.mo_out = mo_out; a.ld_mo = ld_mo;
    OkmiProfScope prof("gemm_f64", st);
    int rc = okmi_gemm_f64_mode0(a, nsets, st);
    if (rc) return rc;
    return okmi_check_launch("gemm_f64");
}

int okmi_launch_rho_f64(const double* A, int64_t ld, const int64_t* set_off, int set_mask, int nao,
                        const double* C, int nmo, const double* occ, int64_t np, const int* out_code, int nout,
                        double* rho, double* delta, int64_t ld_delta, double* mo_norm, hipStream_t st,
                        const F64Screen* scr)
{
    if (np <= 0 || nmo <= 0) return 0;
    if (nout < 0 || nout > 16) {
        okmi_set_error("rho: %d derivative outputs (up to 16 supported)", nout);
        return 1;
    }
    okmi_gemm::GemmArgs a{};
    a.A = A; a.ld = ld; a.nao = nao;
    a.C = C; a.nmo = nmo; a.occ = occ; a.np = np;
    a.rho = rho; a.delta = delta; a.ld_delta = ld_delta; a.mo_norm = mo_norm;
    a.set_mask = set_mask;
    a.nout = nout;
    if (scr) {
        a.screened = 1;
        a.scr = *scr;
    }
    int maxc = 0;
    for (int c = 0; c < OKMI_MAX_SETS; ++c) {
        a.set_off[c] = (set_mask >> c & 1) ? set_off[c] : 0;
        if (set_mask >> c & 1) maxc = c;
    }
    for (int r = 0; r < nout; ++r) {
        a.out_code[r] = out_code[r];
        if (out_code[r] > 0) a.out_mask |= 1 << out_code[r];
    }
    const int nsc = maxc == 0 ? 1 : maxc <= 3 ? 4 : maxc <= 6 ? 7 : 10;
    OkmiProfScope prof("rho_f64", st);
    int rc = okmi_gemm_f64_mode1(a, nsc, st);
    if (rc) return rc;
    return okmi_check_launch("rho_f64");
}
