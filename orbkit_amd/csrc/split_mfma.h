// bf16 operand splitting for FP32-accurate products on the bf16 matrix
// cores, shared by the fused density kernel and the large-MO GEMM.
//   x = hi + mid (+ lo): BF16X3 keeps (hi, mid) and 3 products,
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

// x = hi + mid (+ lo) exactly up to the last part: hi and mid keep the top 16
// bits of x and of the residual x - hi (a truncating split, so every
// residual is exact in FP32); the last part is the residual rounded to
// nearest.  Two bf16 of consecutive k share a dword (low half: even k).
template <int PREC>
__device__ __forceinline__ void split8(const float (&x)[8], Split<PREC>& o)
{
    typedef unsigned u4 __attribute__((ext_vector_type(4)));
    u4 h, m, lo;
    float r1[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const unsigned xb = __float_as_uint(x[j]);
        r1[j] = x[j] - __uint_as_float(xb & 0xffff0000u);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
        h[i] = __builtin_amdgcn_perm(__float_as_uint(x[2 * i + 1]), __float_as_uint(x[2 * i]), 0x07060302u);
    if constexpr (PREC == OKMI_PREC_BF16X6) {
        float r2[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) r2[j] = r1[j] - __uint_as_float(__float_as_uint(r1[j]) & 0xffff0000u);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            m[i] = __builtin_amdgcn_perm(__float_as_uint(r1[2 * i + 1]), __float_as_uint(r1[2 * i]), 0x07060302u);
            const __bf16 a = (__bf16)r2[2 * i], b = (__bf16)r2[2 * i + 1];
            lo[i] = (unsigned)__builtin_bit_cast(unsigned short, a) | ((unsigned)__builtin_bit_cast(unsigned short, b) << 16);
        }
        o.p[2] = __builtin_bit_cast(bf8, lo);
    } else {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const __bf16 a = (__bf16)r1[2 * i], b = (__bf16)r1[2 * i + 1];
            m[i] = (unsigned)__builtin_bit_cast(unsigned short, a) | ((unsigned)__builtin_bit_cast(unsigned short, b) << 16);
        }
    }
    o.p[0] = __builtin_bit_cast(bf8, h);
    o.p[1] = __builtin_bit_cast(bf8, m);
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
// mid rounded to nearest; nc = mpad * kpad entries per plane) or FP32
__device__ inline void store_c_split(double v, int prec, int64_t i, int64_t nc, void* out)
{
    if (prec == OKMI_PREC_F32) {
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
