// FP64 instances of the fused density kernel: AO values in FP64 (full-accuracy
// exponential) in the LDS tile, contracted with the FP64 MO coefficients on
// v_mfma_f64_16x16x4f64 with FP64 accumulation.  The same result as the staged
// FP64 path (AO sets written to HBM, then the FP64 GEMM) without the HBM round
// trip of the AO sets.  Compiled with the FP64 MFMAs in the VGPR form (no
// accumulator copies to AGPRs around each MFMA).
#include "fused_tile.h"

namespace okmi_fused {

int launch_pass_f64(const FusedArgs& a, int pass, int mt, int lm, hipStream_t st)
{
    return launch_pass<OKMI_PREC_F64>(a, pass, mt, lm, st);
}

}  // namespace okmi_fused
