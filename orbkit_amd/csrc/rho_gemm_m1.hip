// FP64 MFMA contraction kernels, density mode.
#include "rho_gemm_impl.h"

int okmi_gemm_f64_mode1(const okmi_gemm::GemmArgs& a, int nsc, hipStream_t st)
{
    return okmi_gemm::launch_rho(a, nsc, st);
}
