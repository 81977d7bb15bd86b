// FP64 fused density kernel: AO values in FP64 (full-accuracy exponential) in
// LDS, contracted with the FP64 MO coefficients on the FP64 matrix cores with
// FP64 accumulation.  The same result as the staged FP64 path (AO sets written
// to HBM, then the FP64 GEMM) without the HBM round trip of the AO sets.
// Compiled with the FP64 MFMAs in the VGPR form (no accumulator copies to
// AGPRs around each MFMA).
//
// On gfx950 the FP64 MFMAs and FP64 VALU instructions share one FP64
// throughput: a CU running MFMA waves and FP64-FMA waves side by side takes
// the sum of their times (3.56 + 1.10 -> 4.56 ms in a probe), so the AO
// evaluation (FP64 VALU) and the contraction (FP64 MFMA) cannot hide each
// other, and a warp-specialised version (4 evaluating + 4 contracting waves,
// double-buffered LDS) was slower (4.05 against 3.41 ms on the medium
// workload).  What pays is less FP64 work:
//   * MOs are padded to 4, not 16: full 16-MO tiles run on
//     v_mfma_f64_16x16x4f64, the 4-MO groups left over on
//     v_mfma_f64_4x4x4f64 (4 blocks), which has the same rate per flop (68
//     against 72 TFLOP/s measured) and the same B operand layout, so the two
//     share the AO operand (40 MOs: 2 tiles + 2 groups, 17% fewer MFMA cycles
//     than 3 padded tiles);
//   * AO rows are packed into 32-row chunks by first fit (fused_layout).
//
// A workgroup of four waves owns a tile of 64 points.  Per chunk: each thread
// loads its share of the chunk's C columns from global memory (held in
// registers through the evaluation), the waves evaluate the chunk's shells
// into the LDS tile (lane = point), the C columns go to LDS, barrier, each
// wave contracts its 16-point column with operands read from LDS one K step
// at a time, barrier.  Two workgroups per CU (64 KB AO tile + up to 15 KB of
// C at four AO sets).
//
// 4x4x4 lane layout (checked on the GPU with one-hot operands): block
// b = (lane % 16) / 4; A lane 16 k + 4 b + i holds A_b[i][k], B lane
// 16 k + 4 b + j holds B_b[k][j], D lane 16 i + 4 b + j holds D_b[i][j].
// With block b = point group (points 4 b .. 4 b + 3 of the column), i = MO
// and k = AO row, B lane l is AO row k0 + l / 16 at point l % 16 (the
// 16x16x4 B operand) and D lane l is MO m0 + l / 16 at point l % 16.
#include "fused_tile.h"

namespace okmi_fused {
namespace {

// MOs of a launch: MT tiles of 16 and MR groups of 4
template <int MT, int MR> struct MoSplit {
    static constexpr int MP = 16 * MT + 4 * MR;   // padded MO count
    static constexpr int T = MT > 0 ? MT : 1;     // array extents
    static constexpr int R = MR > 0 ? MR : 1;
    static constexpr int CR = (MP * KC + 255) / 256;   // C elements per thread per chunk
};

// acc[mt][s] (16x16x4 tiles) and accr[r][s] (4x4x4 groups) += C * AO_s over
// one chunk, for the 16-point column `col`; C from the chunk's LDS stage
// cs[k][m] (MP MOs per row), AO from the LDS tile
template <int PASS, int MT, int MR>
__device__ __forceinline__ void contract_f64(const double* L, const double* cs, int col, int g, int n16,
                                             d4 (&acc)[MoSplit<MT, MR>::T][PassCodes<PASS>::n],
                                             double (&accr)[MoSplit<MT, MR>::R][PassCodes<PASS>::n])
{
    constexpr int NS = PassCodes<PASS>::n;
    constexpr int LD = Tile<OKMI_PREC_F64>::LD;
    constexpr int MP = MoSplit<MT, MR>::MP;
#pragma unroll
    for (int ks = 0; ks < KC / 4; ++ks) {
        const int k = 4 * ks + g;
        double cv[MoSplit<MT, MR>::T], cr[MoSplit<MT, MR>::R], bv[NS];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) cv[mt] = cs[k * MP + 16 * mt + n16];
#pragma unroll
        for (int r = 0; r < MR; ++r) cr[r] = cs[k * MP + 16 * MT + 4 * r + (n16 & 3)];
#pragma unroll
        for (int q = 0; q < NS; ++q) bv[q] = L[(q * KC + k) * LD + col * 16 + n16];
#pragma unroll
        for (int q = 0; q < NS; ++q) {
#pragma unroll
            for (int mt = 0; mt < MT; ++mt)
                acc[mt][q] = __builtin_amdgcn_mfma_f64_16x16x4f64(cv[mt], bv[q], acc[mt][q], 0, 0, 0);
#pragma unroll
            for (int r = 0; r < MR; ++r) accr[r][q] = __builtin_amdgcn_mfma_f64_4x4x4f64(cr[r], bv[q], accr[r][q], 0, 0, 0);
        }
    }
}

// Per-tile MO norms sum_p Phi_m(p)^2 over the tile's points (padding points
// excluded): column sums over the 16 lanes of each MO, the four columns
// added through LDS (red: 4 x MP doubles, free after the chunk loop's last
// barrier), one row of norm_part per tile.  Called by every thread.
template <int PASS, int MT, int MR>
__device__ __forceinline__ void tile_norms(const FusedArgs& a, int64_t tile, int col, int lane, bool act, double* red,
                                           const d4 (&acc)[MoSplit<MT, MR>::T][PassCodes<PASS>::n],
                                           const double (&accr)[MoSplit<MT, MR>::R][PassCodes<PASS>::n])
{
    constexpr int MP = MoSplit<MT, MR>::MP;
    const int g = lane >> 4, n16 = lane & 15;
    auto put = [&](int m, double f) {
        double v = act ? f * f : 0.0;
        v += __shfl_xor(v, 1);
        v += __shfl_xor(v, 2);
        v += __shfl_xor(v, 4);
        v += __shfl_xor(v, 8);
        if (n16 == 0) red[col * MP + m] = v;
    };
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int i = 0; i < 4; ++i) put(16 * mt + 4 * i + g, acc[mt][0][i]);
#pragma unroll
    for (int r = 0; r < MR; ++r) put(16 * MT + 4 * r + g, accr[r][0]);
    __syncthreads();
    const int m = threadIdx.x;
    if (m < MP) a.norm_part[tile * MP + m] = (red[m] + red[MP + m]) + (red[2 * MP + m] + red[3 * MP + m]);
}

// Density terms of the column from its MO values: item i of acc[mt][s] is
// Phi_s[16 mt + 4 i + g][point n16], accr[r][s] is Phi_s[16 MT + 4 r + g][point n16]
template <int PASS, int MT, int MR>
__device__ __forceinline__ void epilogue_f64(const FusedArgs& a, int64_t tile, int col, int lane,
                                             const d4 (&acc)[MoSplit<MT, MR>::T][PassCodes<PASS>::n],
                                             const double (&accr)[MoSplit<MT, MR>::R][PassCodes<PASS>::n])
{
    constexpr int NS = PassCodes<PASS>::n;
    const int g = lane >> 4, n16 = lane & 15;
    double r0 = 0.0, t[NS > 1 ? NS - 1 : 1], u[6];
#pragma unroll
    for (int s = 0; s < (NS > 1 ? NS - 1 : 1); ++s) t[s] = 0.0;
#pragma unroll
    for (int q = 0; q < 6; ++q) u[q] = 0.0;
    auto add = [&](double w, const double* f) {   // f[s]: Phi_s of one MO at this point
        if constexpr (PASS <= 1) r0 += w * (f[0] * f[0]);
#pragma unroll
        for (int s = 1; s < NS; ++s) t[s - 1] += 2.0 * w * (f[0] * f[s]);
        if constexpr (PASS == 1) {
            u[0] += 2.0 * w * (f[1] * f[1]);
            u[1] += 2.0 * w * (f[2] * f[2]);
            u[2] += 2.0 * w * (f[3] * f[3]);
            u[3] += 2.0 * w * (f[1] * f[2]);
            u[4] += 2.0 * w * (f[1] * f[3]);
            u[5] += 2.0 * w * (f[2] * f[3]);
        }
    };
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            double f[NS];
#pragma unroll
            for (int s = 0; s < NS; ++s) f[s] = acc[mt][s][i];
            add(a.occ[16 * mt + 4 * i + g], f);
        }
#pragma unroll
    for (int r = 0; r < MR; ++r) {
        double f[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) f[s] = accr[r][s];
        add(a.occ[16 * MT + 4 * r + g], f);
    }
    r0 += __shfl_xor(r0, 16);
    r0 += __shfl_xor(r0, 32);
#pragma unroll
    for (int s = 0; s < NS - 1; ++s) {
        t[s] += __shfl_xor(t[s], 16);
        t[s] += __shfl_xor(t[s], 32);
    }
    if constexpr (PASS == 1) {
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            u[q] += __shfl_xor(u[q], 16);
            u[q] += __shfl_xor(u[q], 32);
        }
    }
    int64_t p;
    bool act;
    double qx, qy, qz;
    tile_point(a, tile, col * 16 + n16, p, act, qx, qy, qz);
    if (g == 0 && act) {
        if constexpr (PASS <= 1) a.rho[p] = r0;
#pragma unroll
        for (int s = 1; s < NS; ++s) {
            const int code = PassCodes<PASS>::c[s];
            if (a.need_mask >> code & 1) a.term[code * a.ld_term + p] = t[s - 1];
        }
        if constexpr (PASS == 1) {
#pragma unroll
            for (int q = 0; q < 6; ++q)
                if (a.need_mask >> (4 + q) & 1) a.term2[(4 + q) * a.ld_term + p] = u[q];
        }
    }
}

template <int PASS, int MT, int MR, int LM>
__global__ __launch_bounds__(256, 2) void k_rho_fused_f64(FusedArgs a)
{
    using M = MoSplit<MT, MR>;
    constexpr int NS = PassCodes<PASS>::n;
    constexpr int MP = M::MP;
    __shared__ double lds[NS * KC * Tile<OKMI_PREC_F64>::LD];
    __shared__ double cs[KC * MP];   // the chunk's C columns, [k][m]
    __shared__ int next[1];          // next work entry of the chunk being evaluated
    __shared__ int live[1];          // last chunk with a live shell
    __shared__ unsigned long long lmask[LIVE_WORDS];   // chunks with a shell in reach of the tile

    const int64_t tile = blockIdx.x;
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int g = lane >> 4, n16 = lane & 15;
    if (threadIdx.x == 0) {
        next[0] = 0;
        live[0] = -1;
    }
    __syncthreads();
    int64_t pidx;
    bool pact;
    double px, py, pz;
    tile_point(a, tile, lane, pidx, pact, px, py, pz);   // lane l of every wave: point l of the tile
    const Box box = tile_box(a, tile, px, py, pz);
    live_chunks(a, box, wave, lane, lmask);
    __syncthreads();
    // the coordinates are consumed here so that the evaluation does not count
    // their loads as outstanding (it would then also wait for the C loads)
    asm volatile("" ::"v"(px), "v"(py), "v"(pz));

    d4 acc[M::T][NS];
    double accr[M::R][NS];
#pragma unroll
    for (int mt = 0; mt < M::T; ++mt)
#pragma unroll
        for (int s = 0; s < NS; ++s) acc[mt][s] = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int r = 0; r < M::R; ++r)
#pragma unroll
        for (int s = 0; s < NS; ++s) accr[r][s] = 0.0;

    typedef const __attribute__((address_space(1))) double gdouble;
    gdouble* C = (gdouble*)a.C;
    for (int ch = next_live(a, lmask, 0); ch < a.nchunk; ch = next_live(a, lmask, ch + 1)) {
        const int k0 = ch * KC;
        // this thread's C elements e = threadIdx.x + 256 j of [m][k] (k
        // fastest: coalesced rows); clamped to the last element
        double creg[M::CR];
#pragma unroll
        for (int j = 0; j < M::CR; ++j) {
            const int e = min((int)threadIdx.x + 256 * j, MP * KC - 1);
            creg[j] = C[(int64_t)(e / KC) * a.kpad + k0 + e % KC];
        }
        eval_chunk<PASS, LM>(a, ch, box, px, py, pz, lds, &next[0], &live[0], ch, wave == 0, lane);
        // the stage is free: every wave read it before the previous barrier
#pragma unroll
        for (int j = 0; j < M::CR; ++j) {
            const int e = (int)threadIdx.x + 256 * j;
            if (e < MP * KC) cs[(e % KC) * MP + e / KC] = creg[j];
        }
        __syncthreads();
        // every wave finished taking from the counter before the barrier and
        // takes from it again after the next one
        if (threadIdx.x == 0) next[0] = 0;
        if (live[0] == ch) contract_f64<PASS, MT, MR>(lds, cs, wave, g, n16, acc, accr);
        __syncthreads();
    }
    epilogue_f64<PASS, MT, MR>(a, tile, wave, lane, acc, accr);
    if constexpr (PASS <= 1) {
        if (a.norm_part) {   // uniform: every thread takes the branch or none
            int64_t p;
            bool act;
            double qx, qy, qz;
            tile_point(a, tile, wave * 16 + n16, p, act, qx, qy, qz);
            tile_norms<PASS, MT, MR>(a, tile, wave, lane, act, lds, acc, accr);
        }
    }
}

// mo_norm[m] += sum over tiles of part[tile][m], m < nmo: one workgroup per
// MO, each thread a fixed stride of tiles, then a tree in LDS (the order of
// the additions does not depend on timing)
__global__ void k_norm_reduce(const double* part, int64_t ntile, int mpad, int nmo, double* mo_norm)
{
    __shared__ double red[256];
    const int m = blockIdx.x;
    double s = 0.0;
    for (int64_t t = threadIdx.x; t < ntile; t += 256) s += part[t * mpad + m];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0 && m < nmo) mo_norm[m] += red[0];
}

template <int PASS, int LM, int MT, int MR>
int launch_f64_one(const FusedArgs& a, hipStream_t st)
{
    const int64_t ntile = a.brick ? a.nbx * a.nby * a.nbz : (a.np + PTS - 1) / PTS;
    if (ntile >= ((int64_t)1 << 31)) {   // one workgroup per tile; brick indices are 32-bit
        okmi_set_error("rho_fused: %lld tiles of 64 points (at most 2^31 - 1 per call)", (long long)ntile);
        return 1;
    }
    if (ntile == 0) return 0;
    OkmiProfScope prof("rho_fused_f64", st);
    hipLaunchKernelGGL((k_rho_fused_f64<PASS, MT, MR, LM>), dim3((unsigned)ntile), dim3(256), 0, st, a);
    return 0;
}

// mpad: MOs padded to a multiple of 4 (4 .. 64)
template <int PASS, int LM>
int launch_f64_m(const FusedArgs& a, int mpad, hipStream_t st)
{
    switch (mpad) {
#define OKMI_F64_CASE(MT, MR) \
    case 16 * MT + 4 * MR: return launch_f64_one<PASS, LM, MT, MR>(a, st);
        OKMI_F64_CASE(0, 1) OKMI_F64_CASE(0, 2) OKMI_F64_CASE(0, 3)
        OKMI_F64_CASE(1, 0) OKMI_F64_CASE(1, 1) OKMI_F64_CASE(1, 2) OKMI_F64_CASE(1, 3)
        OKMI_F64_CASE(2, 0) OKMI_F64_CASE(2, 1) OKMI_F64_CASE(2, 2) OKMI_F64_CASE(2, 3)
        OKMI_F64_CASE(3, 0) OKMI_F64_CASE(3, 1) OKMI_F64_CASE(3, 2) OKMI_F64_CASE(3, 3)
        OKMI_F64_CASE(4, 0)
#undef OKMI_F64_CASE
    default: break;
    }
    okmi_set_error("rho_fused: %d padded MOs (FP64: a multiple of 4 up to 64)", mpad);
    return 1;
}

}  // namespace

int launch_norm_reduce(const double* part, int64_t ntile, int mpad, int nmo, double* mo_norm, hipStream_t st)
{
    if (ntile <= 0 || nmo <= 0) return 0;
    hipLaunchKernelGGL(k_norm_reduce, dim3((unsigned)nmo), dim3(256), 0, st, part, ntile, mpad, nmo, mo_norm);
    return 0;
}

int launch_pass_f64(const FusedArgs& a, int pass, int mpad, int lm, hipStream_t st)
{
    const bool l3 = lm <= 3;
    switch (pass) {
    case 0: return l3 ? launch_f64_m<0, 3>(a, mpad, st) : launch_f64_m<0, 4>(a, mpad, st);
    case 1: return l3 ? launch_f64_m<1, 3>(a, mpad, st) : launch_f64_m<1, 4>(a, mpad, st);
    case 2: return l3 ? launch_f64_m<2, 3>(a, mpad, st) : launch_f64_m<2, 4>(a, mpad, st);
    default: return l3 ? launch_f64_m<3, 3>(a, mpad, st) : launch_f64_m<3, 4>(a, mpad, st);
    }
}

}  // namespace okmi_fused
