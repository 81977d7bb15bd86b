// Density for many MOs (more than the fused kernel keeps in registers).
// Replaces slice_rho's density branch (orbkit/core.py:179-308) for large
// NMO: per chunk of points
//   1. k_ao_eval writes the AO sets of codes 0..S-1 in FP32 (FP64
//      evaluation rounded once), rows coalesced: ao[s][k][p];
//   2. k_rho_big contracts them with C on the matrix cores.  A workgroup owns
//      128 MOs x NP points x all S sets; the K loop (AO rows, 32 per step)
//      stages the C tile (pre-split bf16 planes or FP32) and the AO tile
//      (FP32) in LDS, double buffered with one barrier per step.  Each wave
//      holds 64 MOs x NPW 16-point columns x S sets of FP32 accumulators
//      (BigCfg), so the density epilogue (all sets of one MO and point in one
//      lane) runs in FP64 without leaving registers; the 128-MO partial sums
//      go to partial[mo tile][code][p];
//   3. k_reduce_big sums the MO tiles in a fixed order and writes rho and
//      orbkit's derivative rows.
// S is 1 (rho), 4 (+ first derivatives), 7 (+ diagonal second derivatives)
// or 10 (+ mixed second derivatives); a request is served by the smallest S
// holding its codes and their product terms.
#include "okmi_internal.h"
#include "split_mfma.h"

namespace {

using namespace okmi_split;

constexpr int MB = 128;   // MOs per workgroup
constexpr int KS = 32;    // AO rows per K step

// Workgroup shape per set count: NP points, NW waves in 2 (MO halves) x WN
// (point columns) with NPW 16-point columns each.  Up to 4 sets a wave holds
// 64 MOs x 16 points (S = 4) so that two waves fit per SIMD (8 waves share
// the LDS tile); with 7 or 10 sets four waves of 64 MOs x 16 points.
template <int S> struct BigCfg {
    static constexpr int NW = S <= 4 ? 8 : 4;
    static constexpr int WN = NW / 2;
    static constexpr int NP = S == 1 ? 128 : S == 4 ? 64 : 32;
    static constexpr int NPW = NP / 16 / WN;
    static constexpr int NT = 64 * NW;
};

struct BigArgs {
    const float* ao;       // [S][nao][ld]
    int64_t ld, set_stride;
    int nao;
    int64_t np;            // points in this chunk
    const void* C;         // [planes][mpad][kpad] bf16 or [mpad][kpad] FP32
    int64_t mpad, kpad;
    const double* occ;     // [mpad], zero padded
    double* partial;       // [nmt][S][ld]
    int nmt;
    int nmo;               // MOs (16-MO tiles at or past nmo are padding: their MFMAs are skipped)
    const int* live;       // [point tile][nchunk]: the chunks (32-row K steps) that reach the tile, in order
    const int* live_count; // [point tile]
    int nchunk;
    int64_t nblocks;
};

template <int PREC> struct Prec {
    static constexpr int planes = PREC == OKMI_PREC_F32 ? 1 : PREC == OKMI_PREC_BF16X3 ? 2 : 3;
};

// LDS row strides chosen for conflict-free operand reads:
//   bf16 B reads: rows 8g + j of lane group g -> stride = 2 (mod 8)
//   F32  B reads: rows 4ks + g              -> stride = 16 (mod 64)
template <int PREC, int NP> struct AoLd { static constexpr int v = PREC == OKMI_PREC_F32 ? NP + 16 : NP + 2; };
constexpr int CLD_BF = KS + 8;    // bf16 elements per C row in LDS (80 bytes)
constexpr int CLD_F = KS + 4;     // floats per C row in LDS (144 bytes)

template <int PREC, int S>
struct BigLds {
    static constexpr int NP = BigCfg<S>::NP;
    static constexpr int AOLD = AoLd<PREC, NP>::v;
    static constexpr int AO_FLOATS = S * KS * AOLD;
    static constexpr int C_BYTES = PREC == OKMI_PREC_F32 ? MB * CLD_F * 4 : Prec<PREC>::planes * MB * CLD_BF * 2;
    static constexpr int STAGE = AO_FLOATS * 4 + C_BYTES;          // bytes per buffer
    static constexpr int TOTAL = 2 * STAGE;
};

// XCD-aware block order: hardware places block b on XCD b % 8; blocks that
// share a point tile (its MO tiles) are given consecutive logical indices on
// one XCD so its L2 serves the AO tile to all of them.
__device__ __forceinline__ int64_t logical_block(int64_t b, int64_t nb)
{
    const int64_t xcd = b & 7, j = b >> 3, per = nb >> 3, rem = nb & 7;
    return (xcd < rem ? xcd * (per + 1) : rem * (per + 1) + (xcd - rem) * per) + j;
}

// LONG: live lists may exceed 128 entries (more than 128 chunks of 32 rows),
// and their tail is read from memory
template <int PREC, int S, bool LONG>
__global__ __launch_bounds__(64 * BigCfg<S>::NW, 1) void k_rho_big(BigArgs a)
{
    using LD = BigLds<PREC, S>;
    constexpr int NPW = BigCfg<S>::NPW, NP = LD::NP, AOLD = LD::AOLD, NT = BigCfg<S>::NT;
    constexpr int NPL = Prec<PREC>::planes;
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];

    const int64_t L = logical_block(blockIdx.x, a.nblocks);
    const int mt = (int)(L % a.nmt);
    const int64_t tn = L / a.nmt;
    const int64_t p0 = tn * NP;
    const int m0 = mt * MB;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int g = lane >> 4, n16 = lane & 15;
    // the K steps of this point tile: the 32-row chunks with a shell that
    // reaches one of its points (k_big_live), in row order
    const int* lst_mem = a.live + tn * a.nchunk;
    const int cnt = a.live_count[tn];
    // the list's first 128 entries held across the lanes (entry j in lane
    // j % 64 of lv[j / 64]) and read with readlane: a step's chunk index is
    // not a scalar load the step waits on
    const int lv0 = lane < cnt ? lst_mem[lane] : 0, lv1 = lane + 64 < cnt ? lst_mem[lane + 64] : 0;
    // 16-MO tiles of this wave that hold MOs (wave-uniform): the last MO
    // tile group of a padded count skips the MFMAs (and the AO split) of
    // its padding tiles, e.g. 400 MOs = 25 tiles of the 32 launched
    const int nv = min(4, max(0, (a.nmo - m0 - 64 * wm + 15) / 16));

    // ---- global -> register staging
    constexpr int AO_V4 = S * KS * NP / 4 / NT;           // float4 per thread per step
    static_assert(S * KS * NP / 4 % NT == 0, "AO tile must split evenly over the threads");
    constexpr int C_V4 = PREC == OKMI_PREC_F32 ? MB * KS / 4 / NT : NPL * MB * KS / 8 / NT;   // 16-byte pieces
    static_assert((PREC == OKMI_PREC_F32 ? MB * KS / 4 : NPL * MB * KS / 8) % NT == 0, "C tile split");
    typedef float fl4 __attribute__((ext_vector_type(4)));
    struct Stage {
        fl4 ao[AO_V4];
        fl4 c[C_V4];
    };
    auto gload = [&](int ks, Stage& R) {
        const int k0 = ks * KS;
#pragma unroll
        for (int v = 0; v < AO_V4; ++v) {
            const int e = v * NT + threadIdx.x;          // float4 index in [S][KS][NP/4]
            const int pc = e % (NP / 4), r = (e / (NP / 4)) % KS, s = e / (NP / 4 * KS);
            // a listed chunk's rows lie below the padded row count; the
            // clamp only keeps the load in bounds.  Unconditional loads: a
            // load behind a branch joins with a wait for every load in flight
            const int k = min(k0 + r, a.nao - 1);
            R.ao[v] = *(const fl4*)(a.ao + s * a.set_stride + (int64_t)k * a.ld + p0 + 4 * pc);
        }
#pragma unroll
        for (int v = 0; v < C_V4; ++v) {
            const int e = v * NT + threadIdx.x;
            if constexpr (PREC == OKMI_PREC_F32) {
                const int kc = e % (KS / 4), r = e / (KS / 4);
                R.c[v] = *(const fl4*)((const float*)a.C + (int64_t)(m0 + r) * a.kpad + k0 + 4 * kc);
            } else {
                const int kc = e % (KS / 8), r = (e / (KS / 8)) % MB, pl = e / (KS / 8 * MB);
                R.c[v] = *(const fl4*)((const __bf16*)a.C + (int64_t)pl * a.mpad * a.kpad + (int64_t)(m0 + r) * a.kpad +
                                       k0 + 8 * kc);
            }
        }
    };
    auto sstore = [&](int buf, const Stage& R) {
        float* ao = (float*)(smem + buf * LD::STAGE);
        unsigned char* cl = smem + buf * LD::STAGE + LD::AO_FLOATS * 4;
#pragma unroll
        for (int v = 0; v < AO_V4; ++v) {
            const int e = v * NT + threadIdx.x;
            const int pc = e % (NP / 4), r = (e / (NP / 4)) % KS, s = e / (NP / 4 * KS);
            float* d = ao + (s * KS + r) * AOLD + 4 * pc;   // 8-byte aligned (AOLD even)
            *(float2*)d = float2{R.ao[v][0], R.ao[v][1]};
            *(float2*)(d + 2) = float2{R.ao[v][2], R.ao[v][3]};
        }
#pragma unroll
        for (int v = 0; v < C_V4; ++v) {
            const int e = v * NT + threadIdx.x;
            if constexpr (PREC == OKMI_PREC_F32) {
                const int kc = e % (KS / 4), r = e / (KS / 4);
                *(fl4*)((float*)cl + r * CLD_F + 4 * kc) = R.c[v];
            } else {
                const int kc = e % (KS / 8), r = (e / (KS / 8)) % MB, pl = e / (KS / 8 * MB);
                *(fl4*)((__bf16*)cl + (pl * MB + r) * CLD_BF + 8 * kc) = R.c[v];
            }
        }
    };

    f4 acc[4][NPW][S];
#pragma unroll
    for (int rb = 0; rb < 4; ++rb)
#pragma unroll
        for (int c = 0; c < NPW; ++c)
#pragma unroll
            for (int s = 0; s < S; ++s) acc[rb][c][s] = f4{0.f, 0.f, 0.f, 0.f};

    auto compute = [&](int buf) {
        const float* ao = (const float*)(smem + buf * LD::STAGE);
        const unsigned char* cl = smem + buf * LD::STAGE + LD::AO_FLOATS * 4;
        if constexpr (PREC == OKMI_PREC_F32) {
#pragma unroll
            for (int kk = 0; kk < KS / 4; ++kk) {
                float av[4];
#pragma unroll
                for (int rb = 0; rb < 4; ++rb)
                    av[rb] = ((const float*)cl)[(64 * wm + 16 * rb + n16) * CLD_F + 4 * kk + g];
#pragma unroll
                for (int s = 0; s < S; ++s)
#pragma unroll
                    for (int c = 0; c < NPW; ++c) {
                        const float bv = ao[(s * KS + 4 * kk + g) * AOLD + (wn * NPW + c) * 16 + n16];
#pragma unroll
                        for (int rb = 0; rb < 4; ++rb)
                            if (rb < nv)
                                acc[rb][c][s] =
                                    __builtin_amdgcn_mfma_f32_16x16x4f32(av[rb], bv, acc[rb][c][s], 0, 0, 0);
                    }
            }
        } else {
            Split<PREC> A[4];
#pragma unroll
            for (int rb = 0; rb < 4; ++rb)
#pragma unroll
                for (int pl = 0; pl < NPL; ++pl)
                    A[rb].p[pl] = *(const bf8*)((const __bf16*)cl + (pl * MB + 64 * wm + 16 * rb + n16) * CLD_BF + 8 * g);
            if (nv > 0) {
#pragma unroll
                for (int s = 0; s < S; ++s)
#pragma unroll
                    for (int c = 0; c < NPW; ++c) {
                        float bx[8];
#pragma unroll
                        for (int j = 0; j < 8; ++j) bx[j] = ao[(s * KS + 8 * g + j) * AOLD + (wn * NPW + c) * 16 + n16];
                        Split<PREC> B;
                        split8<PREC>(bx, B);
#pragma unroll
                        for (int rb = 0; rb < 4; ++rb)
                            if (rb < nv) acc[rb][c][s] = mfma_split<PREC>(A[rb], B, acc[rb][c][s]);
                    }
            }
        }
    };

    // K loop over the live chunks: LDS double buffered, one barrier per
    // step; step i computes on buffer i & 1.  Up to 4 sets the next step's
    // global loads are issued before the step's MFMAs (one register stage).
    // With 7 or 10 sets (one wave per SIMD, nothing else to hide the
    // latency) they are issued two steps ahead in two register stages:
    // R_next holds step i + 1 (loading), R_free held step i (already in LDS)
    // and takes step i + 2 (measured at C5: S = 7 280 -> 266 ms; S = 4
    // 125 -> 137 ms, so S = 4 keeps one stage).
    // Without LONG the list is read by readlane alone (an entry that might
    // come from memory puts each step behind a branch whose join waits for
    // every load in flight).  The next steps' loads are unconditional: past
    // the last step they read the last chunk again, never stored to LDS.
    auto lst = [&](int j) {
        if constexpr (LONG)
            return j < 64 ? __builtin_amdgcn_readlane(lv0, j) : j < 128 ? __builtin_amdgcn_readlane(lv1, j - 64)
                                                                        : lst_mem[j];
        else
            return j < 64 ? __builtin_amdgcn_readlane(lv0, j) : __builtin_amdgcn_readlane(lv1, j - 64);
    };
    if (cnt > 0) {
        Stage R0, R1;
        gload(lst(0), R0);
        sstore(0, R0);
        if constexpr (S >= 7) {
            gload(lst(min(1, cnt - 1)), R1);
            __syncthreads();
            auto body = [&](int i, int buf, Stage& Rnext, Stage& Rfree) {
                gload(lst(min(i + 2, cnt - 1)), Rfree);
                compute(buf);
                if (i + 1 < cnt) sstore(buf ^ 1, Rnext);
                __syncthreads();
            };
            for (int i = 0; i < cnt; i += 2) {
                body(i, 0, R1, R0);
                if (i + 1 < cnt) body(i + 1, 1, R0, R1);
            }
        } else {
            __syncthreads();
            int buf = 0;
            for (int i = 0; i < cnt; ++i) {
                gload(lst(min(i + 1, cnt - 1)), R0);
                compute(buf);
                if (i + 1 < cnt) sstore(buf ^ 1, R0);
                __syncthreads();
                buf ^= 1;
            }
        }
    }

    // ---- epilogue: lane holds Phi_s[m = m0 + 64 wm + 16 rb + 4 g + i][p0 + (wn NPW + c) 16 + n16]
    double* red = (double*)smem;   // [2][S][NP], reused after the last barrier
#pragma unroll
    for (int c = 0; c < NPW; ++c) {
        double v[S];
#pragma unroll
        for (int s = 0; s < S; ++s) v[s] = 0.0;
#pragma unroll
        for (int rb = 0; rb < 4; ++rb)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const double w = a.occ[m0 + 64 * wm + 16 * rb + 4 * g + i];
                const double f0 = acc[rb][c][0][i];
                v[0] += w * (f0 * f0);
#pragma unroll
                for (int s = 1; s < S; ++s) {
                    double t = f0 * (double)acc[rb][c][s][i];
                    if (s >= 4 && s <= 6) {
                        const double fa = acc[rb][c][s - 3][i];
                        t += fa * fa;
                    } else if (s >= 7) {
                        const double fa = acc[rb][c][s == 9 ? 2 : 1][i], fb = acc[rb][c][s == 7 ? 2 : 3][i];
                        t += fa * fb;
                    }
                    v[s] += 2.0 * w * t;
                }
            }
#pragma unroll
        for (int s = 0; s < S; ++s) {
            v[s] += __shfl_xor(v[s], 16);
            v[s] += __shfl_xor(v[s], 32);
        }
        if (g == 0) {
#pragma unroll
            for (int s = 0; s < S; ++s) red[(wm * S + s) * NP + (wn * NPW + c) * 16 + n16] = v[s];
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < S * NP; e += NT) {
        const int s = e / NP, pt = e % NP;
        const int64_t p = p0 + pt;
        if (p < a.np)
            a.partial[((int64_t)mt * S + s) * a.ld + p] = red[s * NP + pt] + red[(S + s) * NP + pt];
    }
}

__global__ void k_pad_occ_big(const double* occ, int nmo, int mpad, double* out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < mpad) out[i] = i < nmo ? occ[i] : 0.0;
}

struct OutCodes {
    int n;
    int c[16];
};

struct BrickMap {
    int on;
    int64_t bk0, nbx, nby, nbz, nx, ny, nz;
};

// rho[p] = sum_mt partial[mt][0][col]; delta[r][p] = sum_mt partial[mt][c_r][col]
// (code 0: 2 rho, orbkit/core.py:299-300 with drv None).  Column col is point
// p = col (range order) or, in brick order, the grid point of lane col % 64
// of brick bk0 + col / 64 (columns outside the grid are dropped).
__global__ void k_reduce_big(const double* partial, int nmt, int S, int64_t ld, int64_t np, OutCodes oc,
                             double* rho, double* delta, int64_t ld_delta, BrickMap bm)
{
    const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= np) return;
    int64_t p = col;
    if (bm.on) {
        const int lane = (int)(col & 63);
        int64_t t = bm.bk0 + (col >> 6);
        const int64_t bz = t % bm.nbz;
        t /= bm.nbz;
        const int64_t by = t % bm.nby, bx = t / bm.nby;
        const int64_t ix = 4 * bx + (lane >> 4), iy = 4 * by + ((lane >> 2) & 3), iz = 4 * bz + (lane & 3);
        if (ix >= bm.nx || iy >= bm.ny || iz >= bm.nz) return;
        p = (ix * bm.ny + iy) * bm.nz + iz;
    }
    double r = 0.0;
    for (int m = 0; m < nmt; ++m) r += partial[(int64_t)m * S * ld + col];
    rho[p] = r;
    for (int j = 0; j < oc.n; ++j) {
        const int c = oc.c[j];
        double v = 0.0;
        if (c == 0) {
            v = 2.0 * r;
        } else {
            for (int m = 0; m < nmt; ++m) v += partial[((int64_t)m * S + c) * ld + col];
        }
        delta[j * ld_delta + p] = v;
    }
}

// Live chunks of each point tile of NP columns: one wave per tile forms the
// tile's bounding box from its points (brick order: the grid points of the
// bricks' lanes; range order: points p0 + column), then tests every 32-row
// chunk's shells against it (fused screening radii, the radii with which the
// AO kernel writes zeros) and writes the chunks with a shell in reach, in
// increasing order, to list[tile][..] and their number to count[tile].
struct LiveArgs {
    GridDesc grid;
    BrickMap bm;
    int64_t p0;          // range order: point index of column 0
    int64_t n;           // columns of this chunk of points
    int np;              // columns per tile
    const int2* chunks;  // per chunk: first work entry, shells | used rows << 16
    const ShellRec* shells;
    int nchunk;
    int* list;                 // null: no lists
    int* count;
    unsigned long long* mask;  // null: no masks; else [tile][words], bit c = chunk c live
    int words;
};

__device__ __forceinline__ double wave_reduce(double v, bool take_min)
{
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const double u = __shfl_xor(v, o);
        v = take_min ? fmin(v, u) : fmax(v, u);
    }
    return v;
}

__global__ __launch_bounds__(64) void k_big_live(LiveArgs a)
{
    const int64_t t = blockIdx.x;
    const int lane = threadIdx.x;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int j = lane; j < a.np; j += 64) {
        const int64_t col = t * a.np + j;
        if (col >= a.n) continue;
        double q[3];
        if (a.bm.on) {
            const int l = (int)(col & 63);
            int64_t b = a.bm.bk0 + (col >> 6);
            const int64_t bz = b % a.bm.nbz;
            b /= a.bm.nbz;
            const int64_t by = b % a.bm.nby, bx = b / a.bm.nby;
            const int64_t ix = 4 * bx + (l >> 4), iy = 4 * by + ((l >> 2) & 3), iz = 4 * bz + (l & 3);
            if (ix >= a.bm.nx || iy >= a.bm.ny || iz >= a.bm.nz) continue;
            q[0] = a.grid.x[ix]; q[1] = a.grid.y[iy]; q[2] = a.grid.z[iz];
        } else {
            grid_point(a.grid, a.p0 + col, q[0], q[1], q[2]);
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            lo[d] = fmin(lo[d], q[d]);
            hi[d] = fmax(hi[d], q[d]);
        }
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        lo[d] = wave_reduce(lo[d], true);
        hi[d] = wave_reduce(hi[d], false);
    }
    int base = 0;
    for (int c0 = 0; c0 < a.nchunk; c0 += 64) {
        const int c = c0 + lane;
        bool live = false;
        if (c < a.nchunk) {
            const int2 cw = a.chunks[c];
            const int nsh = cw.y & 0xffff;
            for (int e = 0; e < nsh && !live; ++e) {
                const ShellRec sh = a.shells[cw.x + e];
                const double cc[3] = {sh.cx, sh.cy, sh.cz};
                double d2 = 0.0;
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    const double dd = fmax(fmax(lo[d] - cc[d], cc[d] - hi[d]), 0.0);
                    d2 = fma(dd, dd, d2);
                }
                live = d2 < sh.rcut2;   // a tile with no point in the grid has lo > hi: d2 = inf
            }
        }
        const unsigned long long m = __builtin_amdgcn_ballot_w64(live);
        if (live && a.list) a.list[t * a.nchunk + base + __builtin_popcountll(m & ((1ull << lane) - 1))] = c;
        if (lane == 0 && a.mask) a.mask[t * a.words + c0 / 64] = m;
        base += __builtin_popcountll(m);
    }
    if (lane == 0 && a.count) a.count[t] = base;
}

// zero rows `rows[0 .. nrows)` of every AO set over columns [0, ncol)
template <typename T>
__global__ void k_zero_rows(T* ao, int S, int64_t set_stride, int64_t ld, const int* rows, int nrows, int64_t ncol)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t per = (int64_t)nrows * ncol;
    if (i >= per * S) return;
    const int64_t s = i / per, r = (i % per) / ncol, c = i % ncol;
    ao[s * set_stride + (int64_t)rows[r] * ld + c] = T(0);
}

template <int PREC, int S, bool LONG>
int launch_big_l(const BigArgs& a, hipStream_t st)
{
    using LD = BigLds<PREC, S>;
    static bool attr_set = false;
    if (!attr_set) {
        hipError_t e = hipFuncSetAttribute((const void*)k_rho_big<PREC, S, LONG>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, LD::TOTAL);
        if (e != hipSuccess) return okmi_hip_fail(e, "rho_big: LDS size");
        attr_set = true;
    }
    OkmiProfScope prof("rho_big", st);
    hipLaunchKernelGGL((k_rho_big<PREC, S, LONG>), dim3((unsigned)a.nblocks), dim3(BigCfg<S>::NT), LD::TOTAL, st, a);
    return 0;
}

template <int PREC, int S>
int launch_big_s(const BigArgs& a, hipStream_t st)
{
    return a.nchunk > 128 ? launch_big_l<PREC, S, true>(a, st) : launch_big_l<PREC, S, false>(a, st);
}

template <int PREC>
int launch_big(const BigArgs& a, int S, hipStream_t st)
{
    switch (S) {
    case 1: return launch_big_s<PREC, 1>(a, st);
    case 4: return launch_big_s<PREC, 4>(a, st);
    case 7: return launch_big_s<PREC, 7>(a, st);
    default: return launch_big_s<PREC, 10>(a, st);
    }
}

inline size_t a256(size_t x) { return (x + 255) & ~(size_t)255; }

int big_sets(const int* drv_out, int nout)
{
    int mx = 0;
    for (int r = 0; r < nout; ++r) mx = std::max(mx, drv_out[r]);
    return mx == 0 ? 1 : mx <= 3 ? 4 : mx <= 6 ? 7 : 10;
}

struct BigLayout {
    int64_t mpad, kpad, chunk;
    size_t c_bytes, occ_bytes, per_point, fixed;
};

BigLayout big_layout(const OkmiBasis* b, int nmo, int prec, int S)
{
    BigLayout L{};
    L.mpad = (int64_t)((nmo + MB - 1) / MB) * MB;
    L.kpad = std::max<int64_t>(b->nkpad, (int64_t)((b->nao + KS - 1) / KS) * KS);   // fused row order
    const size_t elem = prec == OKMI_PREC_F32 ? 4 : prec == OKMI_PREC_BF16X3 ? 4 : 6;
    L.c_bytes = a256(elem * L.mpad * L.kpad);
    L.occ_bytes = a256(8 * L.mpad);
    // AO sets (FP32), partials (FP64) and the live-chunk lists (at least 32
    // points per tile) per point
    L.per_point = (size_t)S * L.kpad * 4 + (size_t)(L.mpad / MB) * S * 8 + ((size_t)L.kpad / KS + 1) * 4 / 32 +
                  ((size_t)L.kpad / KS / 64 + 1) * 8 / 64 + 2;
    L.fixed = L.c_bytes + L.occ_bytes + 4 * 256 + 256;
    return L;
}

}  // namespace

int okmi_launch_live_chunks(const OkmiBasis* b, const GridDesc& g, int brick, int64_t bk0, int64_t p0, int64_t n,
                            int np, int* list, int* count, hipStream_t st, unsigned long long* mask, bool split)
{
    if (n <= 0) return 0;
    LiveArgs la{};
    la.grid = g;
    la.bm.on = brick;
    if (brick) {
        la.bm.nbx = (g.nx + 3) / 4; la.bm.nby = (g.ny + 3) / 4; la.bm.nbz = (g.nz + 3) / 4;
        la.bm.nx = g.nx; la.bm.ny = g.ny; la.bm.nz = g.nz;
        la.bm.bk0 = bk0;
    }
    la.p0 = p0; la.n = n; la.np = np;
    la.chunks = b->d_chunks; la.shells = split ? b->d_fshells_s : b->d_fshells; la.nchunk = b->nchunk;
    la.list = list; la.count = count;
    la.mask = mask; la.words = live_mask_words(b);
    hipLaunchKernelGGL(k_big_live, dim3((unsigned)((n + np - 1) / np)), dim3(64), 0, st, la);
    return 0;
}

int okmi_launch_zero_rows(double* a, int S, int64_t set_stride, int64_t ld, const int* rows, int nrows, int64_t ncol,
                          hipStream_t st)
{
    const int64_t nz = (int64_t)S * nrows * ncol;
    if (nz <= 0) return 0;
    hipLaunchKernelGGL(k_zero_rows<double>, dim3((unsigned)((nz + 255) / 256)), dim3(256), 0, st, a, S, set_stride,
                       ld, rows, nrows, ncol);
    return 0;
}

size_t okmi_big_ws_bytes(const OkmiBasis* b, int nmo, int prec, int S, int64_t np)
{
    const BigLayout L = big_layout(b, nmo, prec, S);
    const int64_t n = (np + 127) & ~(int64_t)127;
    return L.fixed + L.per_point * (size_t)n;
}

int okmi_launch_rho_big(OkmiBasis* b, const GridDesc& g, int64_t p0, int64_t np, const double* d_c,
                        const double* d_occ, int nmo, int prec, const int* drv_out, int nout, double* rho,
                        double* delta, int64_t ld_delta, void* ws, size_t ws_bytes, hipStream_t st)
{
    if (np <= 0) return 0;
    if (prec != OKMI_PREC_F32 && prec != OKMI_PREC_BF16X3 && prec != OKMI_PREC_BF16X6) {
        okmi_set_error("rho_big: unknown precision %d", prec);
        return 1;
    }
    if (nout > 16) {
        okmi_set_error("rho_big: %d derivative outputs (up to 16)", nout);
        return 1;
    }
    for (int r = 0; r < nout; ++r)
        if (drv_out[r] < 0 || drv_out[r] > 9) {
            okmi_set_error("The selection `drv=%d` is not valid!", drv_out[r]);
            return 1;
        }
    const int S = big_sets(drv_out, nout);
    const BigLayout L = big_layout(b, nmo, prec, S);
    int rc0 = okmi_prepare_fused(b);   // output-row shells and primitives with the fused radii
    if (rc0) return rc0;
    if (!ws || ws_bytes < L.fixed + L.per_point * 128) {
        okmi_set_error("rho_big: workspace of %zu bytes is too small (need at least %zu)", ws_bytes,
                       L.fixed + L.per_point * 128);
        return 1;
    }
    // a whole regular grid is processed in 4x4x4 bricks (compact point sets
    // for screening, 64-column runs per brick); otherwise in range order
    BrickMap bm{};
    int64_t total = np;
    if (g.regular && p0 == 0 && np == g.nx * g.ny * g.nz) {
        bm.on = 1;
        bm.nbx = (g.nx + 3) / 4; bm.nby = (g.ny + 3) / 4; bm.nbz = (g.nz + 3) / 4;
        bm.nx = g.nx; bm.ny = g.ny; bm.nz = g.nz;
        total = bm.nbx * bm.nby * bm.nbz * 64;
    }
    int64_t chunk = (int64_t)((ws_bytes - L.fixed) / L.per_point) & ~(int64_t)127;
    const int64_t npad = (total + 127) & ~(int64_t)127;
    if (chunk > npad) chunk = npad;
    char* w = (char*)(((uintptr_t)ws + 255) & ~(uintptr_t)255);
    void* cpl = w;
    double* occp = (double*)(w + L.c_bytes);
    float* ao = (float*)(w + L.c_bytes + L.occ_bytes);
    const int nmt = (int)(L.mpad / MB);
    double* part = (double*)((char*)ao + a256((size_t)S * L.kpad * 4 * chunk));
    const int NP = 32 * (S == 1 ? 4 : S == 4 ? 2 : 1);
    const int64_t ntile_max = (chunk + NP - 1) / NP;
    int* live = (int*)((char*)part + a256((size_t)nmt * S * 8 * chunk));
    int* live_count = live + ntile_max * b->nchunk;
    // per 64-point brick (128-point tile at S = 1) the live-chunk mask of the AO kernel
    const int nw = live_mask_words(b);
    const int mnp = NP == 128 ? 128 : 64;
    unsigned long long* lmask =
        (unsigned long long*)(((uintptr_t)(live_count + ntile_max) + 7) & ~(uintptr_t)7);
    const int64_t nmask = (chunk + mnp - 1) / mnp;
    if (L.kpad != b->nkpad || (size_t)((char*)(lmask + nmask * nw) - (char*)ws) > ws_bytes) {
        okmi_set_error("rho_big: inconsistent layout (%lld rows, %d fused rows)", (long long)L.kpad, b->nkpad);
        return 1;
    }

    // C in the fused row order (zero columns at the padding rows), the
    // padding rows of the AO buffer zeroed once (no shell writes them)
    hipLaunchKernelGGL(k_prep_c, dim3((unsigned)((L.mpad * L.kpad + 255) / 256)), dim3(256), 0, st, d_c, nmo, b->nao,
                       prec, L.mpad, L.kpad, (const int*)b->d_colmap, cpl);
    if (b->npadrows > 0) {
        const int64_t nz = (int64_t)S * b->npadrows * chunk;
        hipLaunchKernelGGL(k_zero_rows<float>, dim3((unsigned)((nz + 255) / 256)), dim3(256), 0, st, ao, S,
                           (int64_t)L.kpad * chunk, chunk, (const int*)b->d_padrows, b->npadrows, chunk);
    }
    hipLaunchKernelGGL(k_pad_occ_big, dim3((unsigned)((L.mpad + 255) / 256)), dim3(256), 0, st, d_occ, nmo,
                       (int)L.mpad, occp);
    int codes[10];
    for (int s = 0; s < S; ++s) codes[s] = s;
    OutCodes oc{};
    oc.n = nout;
    for (int r = 0; r < nout; ++r) oc.c[r] = drv_out[r];
    // the AO values end in FP32: the screening bound of that arithmetic
    const AoLiveMask am{lmask, nw, mnp == 128 ? 1 : 0, b->d_wshells_s};
    for (int64_t c0 = 0; c0 < total; c0 += chunk) {
        const int64_t n = total - c0 < chunk ? total - c0 : chunk;
        int rc;
        if (bm.on) bm.bk0 = c0 / 64;
        // live chunks first (the GEMM's lists per NP-point tile, the AO
        // kernel's masks per brick or 128-point tile): the AO kernel leaves
        // the rows of dead chunks unwritten
        const int64_t ntile = (n + NP - 1) / NP;
        LiveArgs la{};
        la.grid = g; la.bm = bm; la.p0 = p0 + c0; la.n = n; la.np = NP;
        la.chunks = b->d_chunks; la.shells = b->d_fshells_s; la.nchunk = b->nchunk;
        la.list = live; la.count = live_count;
        la.words = nw;
        if (NP == mnp) la.mask = lmask;
        hipLaunchKernelGGL(k_big_live, dim3((unsigned)ntile), dim3(64), 0, st, la);
        if (NP != mnp) {
            LiveArgs lm = la;
            lm.np = mnp; lm.list = nullptr; lm.count = nullptr; lm.mask = lmask;
            hipLaunchKernelGGL(k_big_live, dim3((unsigned)((n + mnp - 1) / mnp)), dim3(64), 0, st, lm);
        }
        if (bm.on) {
            AoBricks bk{c0 / 64, n / 64, bm.nbx, bm.nby, bm.nbz};
            rc = okmi_launch_ao_eval_ex(b, b->d_ishells_s, b->d_fprims_s, g, 0, n, &bk, codes, S, OKMI_F32, ao, chunk,
                                        (int64_t)L.kpad * chunk, st, &am);
        } else {
            rc = okmi_launch_ao_eval_ex(b, b->d_ishells_s, b->d_fprims_s, g, p0 + c0, n, nullptr, codes, S, OKMI_F32, ao,
                                        chunk, (int64_t)L.kpad * chunk, st, &am);
        }
        if (rc) return rc;
        BigArgs a{};
        a.ao = ao; a.ld = chunk; a.set_stride = (int64_t)L.kpad * chunk; a.nao = (int)L.kpad; a.np = n;
        a.C = cpl; a.mpad = L.mpad; a.kpad = L.kpad; a.occ = occp; a.partial = part; a.nmt = nmt; a.nmo = nmo;
        a.live = live; a.live_count = live_count; a.nchunk = b->nchunk;
        a.nblocks = (int64_t)nmt * ntile;
        switch (prec) {
        case OKMI_PREC_F32: rc = launch_big<OKMI_PREC_F32>(a, S, st); break;
        case OKMI_PREC_BF16X3: rc = launch_big<OKMI_PREC_BF16X3>(a, S, st); break;
        default: rc = launch_big<OKMI_PREC_BF16X6>(a, S, st); break;
        }
        if (rc) return rc;
        hipLaunchKernelGGL(k_reduce_big, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, part, nmt, S, chunk, n, oc,
                           bm.on ? rho : rho + c0, bm.on ? delta : (delta ? delta + c0 : nullptr), ld_delta, bm);
    }
    return okmi_check_launch("rho_big");
}
