// FP64 MFMA contraction kernels, MO-output mode.
#include "rho_gemm_impl.h"

int okmi_gemm_f64_mode0(const okmi_gemm::GemmArgs& a, int nsets, hipStream_t st)
{
    return okmi_gemm::launch_mg(a, nsets, st);
}
