// bf16 operand splitting for FP32-accurate products on the bf16 matrix
// cores, shared by the fused density kernel and the large-MO GEMM.
//   x = hi + mid (+ lo), each part rounded to nearest bf16 from the exact
//   residual of the previous ones: BF16X3 keeps (hi, mid) and 3 products,
//   BF16X6 keeps (hi, mid, lo) and 6 products (terms below 2^-24 dropped).
#pragma once
#include "okmi_common.h"

namespace okmi_split {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

template <int PREC>
struct Split {
    bf8 p[PREC == OKMI_PREC_BF16X6 ? 3 : 2];
};

typedef __bf16 bf2 __attribute__((ext_vector_type(2)));

// bf16 pair from its bit pattern, passed through an (empty, side-effect
// free) asm statement so that it reaches the dot product in a register.  As
// a compile-time constant the compiler would encode (-1, 0) as the 16-bit
// inline constant -1.0, which the dot product reads as the FP16 number -1.0
// (bit pattern 0xbc00, the bf16 number -2^-7): residuals formed that way were
// wrong on gfx950.  Not volatile, so it is hoisted and does not constrain
// instruction scheduling.
__device__ __forceinline__ bf2 pair_bits(unsigned bits)
{
    asm("" : "+s"(bits));
    return __builtin_bit_cast(bf2, bits);
}

// x - b.x and x - b.y for a bf16 pair b: one bf16 dot product with (-1, 0)
// or (0, -1) accumulated onto x.  When b holds x rounded to bf16 the
// difference needs at most 16 significant bits, so it is exact in FP32.
__device__ __forceinline__ float res_lo(bf2 b, float x)
{
    return __builtin_amdgcn_fdot2_f32_bf16(b, pair_bits(0x0000bf80u), x, false);
}
__device__ __forceinline__ float res_hi(bf2 b, float x)
{
    return __builtin_amdgcn_fdot2_f32_bf16(b, pair_bits(0xbf800000u), x, false);
}

// x = hi + mid (+ lo) exactly up to the last part: each part is the
// previous residual rounded to nearest bf16 (two per conversion), and every
// residual is exact in FP32 (res_lo / res_hi).  Two bf16 of consecutive k
// share a dword (low half: even k).
template <int PREC>
__device__ __forceinline__ void split8(const float (&x)[8], Split<PREC>& o)
{
    bf2 h[4], m[4], lo[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        h[i] = bf2{(__bf16)x[2 * i], (__bf16)x[2 * i + 1]};
        const float r0 = res_lo(h[i], x[2 * i]), r1 = res_hi(h[i], x[2 * i + 1]);
        m[i] = bf2{(__bf16)r0, (__bf16)r1};
        if constexpr (PREC == OKMI_PREC_BF16X6) {
            const float s0 = res_lo(m[i], r0), s1 = res_hi(m[i], r1);
            lo[i] = bf2{(__bf16)s0, (__bf16)s1};
        }
    }
    o.p[0] = bf8{h[0].x, h[0].y, h[1].x, h[1].y, h[2].x, h[2].y, h[3].x, h[3].y};
    o.p[1] = bf8{m[0].x, m[0].y, m[1].x, m[1].y, m[2].x, m[2].y, m[3].x, m[3].y};
    if constexpr (PREC == OKMI_PREC_BF16X6)
        o.p[2] = bf8{lo[0].x, lo[0].y, lo[1].x, lo[1].y, lo[2].x, lo[2].y, lo[3].x, lo[3].y};
}

template <int PREC>
__device__ __forceinline__ f4 mfma_split(const Split<PREC>& A, const Split<PREC>& B, f4 acc)
{
    // smallest terms first
    if constexpr (PREC == OKMI_PREC_BF16X6) {
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A.p[2], B.p[0], acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A.p[1], B.p[1], acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A.p[0], B.p[2], acc, 0, 0, 0);
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A.p[1], B.p[0], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A.p[0], B.p[1], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A.p[0], B.p[0], acc, 0, 0, 0);
    return acc;
}

namespace {

// one entry of the padded operand planes: [plane][mpad][kpad] bf16 (hi and
// mid rounded to nearest; nc = mpad * kpad entries per plane), FP32 or FP64
__device__ inline void store_c_split(double v, int prec, int64_t i, int64_t nc, void* out)
{
    if (prec == OKMI_PREC_F64) {
        ((double*)out)[i] = v;
    } else if (prec == OKMI_PREC_F32) {
        ((float*)out)[i] = (float)v;
    } else {
        __bf16* o = (__bf16*)out;
        const float x = (float)v;
        const __bf16 h = (__bf16)x;
        const float r1 = x - (float)h;
        const __bf16 md = (__bf16)r1;
        o[i] = h;
        o[nc + i] = md;
        if (prec == OKMI_PREC_BF16X6) o[2 * nc + i] = (__bf16)(r1 - (float)md);
    }
}

// C (FP64 [nmo][nao]) -> padded operand planes; with a column map, internal
// column k takes C's column colmap[k] (-1: zero)
__global__ void k_prep_c(const double* C, int nmo, int nao, int prec, int64_t mpad, int64_t kpad, const int* colmap,
                         void* out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= mpad * kpad) return;
    const int64_t m = i / kpad, k = i - m * kpad;
    const int64_t src = colmap ? colmap[k] : (k < nao ? k : -1);
    const double v = (m < nmo && src >= 0) ? C[m * nao + src] : 0.0;
    store_c_split(v, prec, i, mpad * kpad, out);
}

}  // namespace
}  // namespace okmi_split
