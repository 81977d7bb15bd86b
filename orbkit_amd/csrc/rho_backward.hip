// Backward pass of rho(r) = sum_i occ_i Phi_i(r)^2, Phi = C A(r), for an
// upstream gradient g(r) = dL/drho(r).  orbkit has no derivative code; this
// differentiates the forward path of orbkit/core.py:rho_compute
// (orbkit/core.py:529-607: AO values, MO values C A, occ-weighted squares)
// with respect to the MO coefficients, the primitive exponents and the
// nuclear positions.  Per chunk of points (AO values are recomputed, never
// kept across chunks):
//
//   A        AO values (okmi_launch_ao_eval)
//   Phi      = C A                                  (k_gemm_nn, MFMA)
//   B_ir     = 2 occ_i g_r Phi_ir                   (k_bmat)
//   dL/dC   += B A^T                                (k_gemm_tn, MFMA, split over points,
//                                                    partials summed in a fixed order)
//   G        = C^T B   (= dL/dA)                    (k_gemm_nn with C^T)
//   per shell s (k_shell_grad, one block per shell and range of points):
//     w_c(r)  = sum_{rows mu of s} G_mu(r) T_s[mu][c]     (c: Cartesian component)
//     dL/dR_atom(s) -= sum_r sum_c w_c(r) grad_r chi_c(r)   (chi_c depends on r - R)
//     dL/dalpha_p   += sum_r q_s(r) d_p e_p(r) (nu_l / alpha_p - |r - R|^2),
//        q_s = sum_c w_c x^lx y^ly z^lz, e_p = exp(-alpha_p |r-R|^2),
//        nu_l = (2l+3)/4 for primitives normalised by orbkit's
//        N(alpha) ~ alpha^((2l+3)/4) (orbkit/c_support.c:177-188), 0 for
//        pre-normalised primitives.
//   Row factors of the basis (the 1/sqrt(df) and cart_scale folded into T)
//   are held fixed.
//
// Two arithmetics share these kernels (template argument R):
//   double  FP64 throughout (v_mfma_f64_16x16x4f64).
//   float   FP32 values and MFMA products (v_mfma_f32_16x16x4f32); kept in
//           FP64: the grid points, r - R and the exp argument alpha |r-R|^2
//           (tight exponents near their centre), the basis constants, and
//           every reduction across lanes, ranges and chunks (partial sums of
//           FP32 terms are added in FP64 in a fixed order).
#include "okmi_internal.h"
#include "ao_device.h"
#include <type_traits>

namespace {

constexpr int MAXP = 16;     // primitives per shell handled by one block (more: extra windows)
constexpr int RANGE = 4096;  // points per shell-gradient block (16 steps of 256: many blocks in flight)
constexpr int GEN_T_MAX = 15 * 15;   // elements of a general transform block the shell gradient holds in LDS

// 16x16x4 MFMA of each arithmetic.  Item i of lane l of the accumulator
// holds D[row(i, l/16)][l%16]: FP64 row 4i + l/16, FP32 row 4(l/16) + i.
template <typename R> struct Mfma;
template <> struct Mfma<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mac(double a, double b, v4 c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int i, int q) { return 4 * i + q; }
};
template <> struct Mfma<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mac(float a, float b, v4 c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int i, int q) { return 4 * q + i; }
};

template <typename R> __device__ __forceinline__ R gauss(double t);
template <> __device__ __forceinline__ double gauss<double>(double t) { return gauss_f64(t); }
template <> __device__ __forceinline__ float gauss<float>(double t) { return gauss_f32(t); }

template <typename R>
__global__ void k_transpose(const R* C, int nmo, int nao, R* CT)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)nmo * nao) return;
    const int64_t m = i / nao, k = i - m * nao;
    CT[k * nmo + m] = C[i];
}

// B[i][p] = 2 occ_i g(p) Phi[i][p] for p < np, 0 for np <= p < nfill
template <typename R>
__global__ void k_bmat(const R* Phi, const R* occ, const R* g, int nmo, int64_t np, int64_t nfill, int64_t ld, R* B)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nfill) return;
    const R gp = p < np ? R(2) * g[p] : R(0);
    for (int i = 0; i < nmo; ++i) B[i * ld + p] = p < np ? occ[i] * gp * Phi[i * ld + p] : R(0);
}

// W (M x K, row-major) packed as MFMA A operands: Wp[t][s][l] = W[16t +
// l%16][4s + l/16] for M tile t and K step s, zero outside W, so that a
// wave reads a tile's operand for one K step as one coalesced 64-lane load.
template <typename R>
__global__ void k_pack_w(const R* W, int M, int K, R* Wp)
{
    const int nmt = (M + 15) / 16, nks = (K + 3) / 4;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)nmt * nks * 64) return;
    const int l = (int)(i & 63);
    const int64_t ts = i >> 6;
    const int t = (int)(ts / nks), st = (int)(ts % nks);
    const int m = 16 * t + (l & 15), k = 4 * st + (l >> 4);
    Wp[i] = (m < M && k < K) ? W[(int64_t)m * K + k] : R(0);
}

// out[m][p] = sum_k W[m][k] X[k][p] for m < M and the points of [0, npad)
// (npad a multiple of 64, <= the row strides; X rows 16-byte aligned for
// FP32, 32-byte for FP64), W given packed by k_pack_w.  One wave per 64
// points computes every M tile, in groups of MG tiles: lane (r = l%16,
// q = l/16) reads X[k + q][p0 + 4r .. p0 + 4r + 3] (one vector load per row
// of a K step) and MFMA j of the step takes element j, so column r of point
// tile j is point p0 + 4r + j; the four point tiles of one output row then
// leave this lane as one vector store of four consecutive points.  The
// operands of K step s + D are loaded while step s is multiplied (D-deep
// register ring), which hides the HBM latency of X behind the MFMAs.  Rows
// k >= K of the last step read row K - 1 (finite), whose W entries are 0.
//
// Screened modes (rows in the fused order, 32-row chunks; the wave's 64
// points are one tile of the live-chunk lists live[tile][..], lcount[tile]):
//   MODE 1  K runs over the tile's live chunks only (Phi = C A: the rows of
//           the other chunks are not evaluated);
//   MODE 2  only the M tiles of the tile's live chunks are computed, two per
//           chunk (MG = 2; G = C^T B: the other rows are never read).
template <typename R, int MG, int MODE>
__global__ __launch_bounds__(256, 2) void k_gemm_nn(const R* __restrict__ Wp, int M, int K,
                                                    const R* __restrict__ X, int64_t ldx, int64_t npad,
                                                    R* __restrict__ out, int64_t ldo, const int* __restrict__ live,
                                                    const int* __restrict__ lcount, int nchunk)
{
    static_assert(MODE != 2 || MG == 2, "MODE 2 computes the two M tiles of one chunk per group");
    typedef typename Mfma<R>::v4 V;
    constexpr int D = 4;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = lane & 15, q = lane >> 4;
    const int64_t tile = (int64_t)blockIdx.x * 4 + wave;
    const int64_t p0 = tile * 64;
    if (p0 >= npad) return;
    const R* xp = X + p0 + 4 * r;
    const int nmt = (M + 15) / 16, nks = (K + 3) / 4;
    const int* lst = MODE ? live + tile * nchunk : nullptr;
    const int nlive = MODE ? __builtin_amdgcn_readfirstlane(lcount[tile]) : 0;
    // the list's first 128 entries held across the lanes, read with readlane
    // (no scalar load for a step to wait on)
    const int lv0 = MODE && lane < nlive ? lst[lane] : 0, lv1 = MODE && lane + 64 < nlive ? lst[lane + 64] : 0;
    // LONG: a list beyond 128 entries, whose tail is read from memory.  The
    // common short list is read by readlane alone: a possible memory read
    // would put each step's index behind a branch whose join waits for every
    // load in flight (the prefetched operands included).
    auto run = [&](auto longc) {
    constexpr bool LONG = decltype(longc)::value;
    auto entry = [&](int j) {
        if constexpr (LONG)
            return j < 64 ? __builtin_amdgcn_readlane(lv0, j) : j < 128 ? __builtin_amdgcn_readlane(lv1, j - 64) : lst[j];
        else
            return j < 64 ? __builtin_amdgcn_readlane(lv0, j) : __builtin_amdgcn_readlane(lv1, j - 64);
    };
    // K steps of this wave and the packed step index of step st
    const int nst = MODE == 1 ? 8 * nlive : nks;
    auto kstep = [&](int st) { return MODE == 1 ? entry(st >> 3) * 8 + (st & 7) : st; };
    const int ngroup = MODE == 2 ? nlive : (nmt + MG - 1) / MG;
    for (int gi = 0; gi < ngroup; ++gi) {
        const int g0 = MODE == 2 ? 2 * entry(gi) : gi * MG;   // first M tile of the group
        V xr[D];
        R wr[D][MG];
        // operands of step st (0 <= st < nst).  Every load is unconditional
        // and in bounds (M tiles past the last read the last tile: their
        // accumulators are never stored), so the loads of a step are not
        // split into branches of their own, each with its own wait for every
        // load in flight.
        auto load = [&](int st, V& xv, R (&wv)[MG]) {
            const int ks = kstep(st);
            xv = *(const V*)(xp + (int64_t)min(4 * ks + q, K - 1) * ldx);
#pragma unroll
            for (int mt = 0; mt < MG; ++mt) {
                const int t = min(g0 + mt, nmt - 1);
                wv[mt] = Wp[((int64_t)t * nks + ks) * 64 + lane];
            }
        };
        if (nst > 0) {
#pragma unroll
            for (int d = 0; d < D; ++d) load(min(d, nst - 1), xr[d], wr[d]);
        }
        V acc[MG][4];
#pragma unroll
        for (int mt = 0; mt < MG; ++mt)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[mt][j] = V{0, 0, 0, 0};
        for (int s0 = 0; s0 < nst; s0 += D) {
#pragma unroll
            for (int d = 0; d < D; ++d) {
                const int st = s0 + d;
                const R on = st < nst ? R(1) : R(0);   // steps past the end add nothing
                const V xv = xr[d];
                R wv[MG];
#pragma unroll
                for (int mt = 0; mt < MG; ++mt) wv[mt] = wr[d][mt] * on;
                load(min(st + D, nst - 1), xr[d], wr[d]);
#pragma unroll
                for (int j = 0; j < 4; ++j)
#pragma unroll
                    for (int mt = 0; mt < MG; ++mt) acc[mt][j] = Mfma<R>::mac(wv[mt], xv[j], acc[mt][j]);
            }
        }
#pragma unroll
        for (int mt = 0; mt < MG; ++mt)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int m = (g0 + mt) * 16 + Mfma<R>::row(i, q);
                if (m < M)
                    *(V*)(out + (int64_t)m * ldo + p0 + 4 * r) =
                        V{acc[mt][0][i], acc[mt][1][i], acc[mt][2][i], acc[mt][3][i]};
            }
    }
    };
    if (MODE && nlive > 128) run(std::true_type{});
    else run(std::false_type{});
}

// The same product for a short K (at most 4 KS rows, K <= 64: G = C^T B
// with the MOs as K): the wave loads its 64 points' K rows of X into
// registers once and then runs every M group from them (the streaming form
// above re-reads X once per group).  MODE 0: all M tiles in groups of MG;
// MODE 2: the M tiles of the tile's live chunks (MG = 2).
template <typename R, int MG, int MODE, int KS>
__global__ __launch_bounds__(256, 2) void k_gemm_nn_xr(const R* __restrict__ Wp, int M, int K,
                                                       const R* __restrict__ X, int64_t ldx, int64_t npad,
                                                       R* __restrict__ out, int64_t ldo, const int* __restrict__ live,
                                                       const int* __restrict__ lcount, int nchunk)
{
    static_assert(MODE != 2 || MG == 2, "MODE 2 computes the two M tiles of one chunk per group");
    typedef typename Mfma<R>::v4 V;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = lane & 15, q = lane >> 4;
    const int64_t tile = (int64_t)blockIdx.x * 4 + wave;
    const int64_t p0 = tile * 64;
    if (p0 >= npad) return;
    const R* xp = X + p0 + 4 * r;
    const int nmt = (M + 15) / 16, nks = (K + 3) / 4;
    V xr[KS];
#pragma unroll
    for (int st = 0; st < KS; ++st)
        xr[st] = st < nks ? *(const V*)(xp + (int64_t)min(4 * st + q, K - 1) * ldx) : V{0, 0, 0, 0};
    const int* lst = MODE ? live + tile * nchunk : nullptr;
    const int ngroup = MODE == 2 ? __builtin_amdgcn_readfirstlane(lcount[tile]) : (nmt + MG - 1) / MG;
    // the list's first 128 entries held across the lanes, read with readlane;
    // LONG: a longer list, whose tail is read from memory
    const int lv0 = MODE && lane < ngroup ? lst[lane] : 0, lv1 = MODE && lane + 64 < ngroup ? lst[lane + 64] : 0;
    auto run = [&](auto longc) {
    constexpr bool LONG = decltype(longc)::value;
    auto entry = [&](int j) {
        if constexpr (LONG)
            return j < 64 ? __builtin_amdgcn_readlane(lv0, j) : j < 128 ? __builtin_amdgcn_readlane(lv1, j - 64) : lst[j];
        else
            return j < 64 ? __builtin_amdgcn_readlane(lv0, j) : __builtin_amdgcn_readlane(lv1, j - 64);
    };
    for (int gi = 0; gi < ngroup; ++gi) {
        const int g0 = MODE == 2 ? 2 * entry(gi) : gi * MG;
        V acc[MG][4];
#pragma unroll
        for (int mt = 0; mt < MG; ++mt)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[mt][j] = V{0, 0, 0, 0};
        // the group's W operands, loaded in batches of WB K steps before
        // their MFMAs (one wait per batch, not one per K step; FP64 at 16
        // steps takes two batches, which keeps it within the registers); the
        // loads are unconditional and in bounds: M tiles past the last read
        // the last tile (never stored)
        constexpr int WB = sizeof(R) == 8 && KS > 12 ? KS / 4 : KS;
#pragma unroll
        for (int s0 = 0; s0 < KS; s0 += WB) {
            R wv[WB][MG];
#pragma unroll
            for (int u = 0; u < WB; ++u)
#pragma unroll
                for (int mt = 0; mt < MG; ++mt) {
                    const int t = min(g0 + mt, nmt - 1), ss = min(s0 + u, nks - 1);
                    wv[u][mt] = Wp[((int64_t)t * nks + ss) * 64 + lane];
                }
#pragma unroll
            for (int u = 0; u < WB; ++u) {
                if (s0 + u < nks) {
#pragma unroll
                    for (int j = 0; j < 4; ++j)
#pragma unroll
                        for (int mt = 0; mt < MG; ++mt)
                            acc[mt][j] = Mfma<R>::mac(wv[u][mt], xr[s0 + u][j], acc[mt][j]);
                }
            }
        }
#pragma unroll
        for (int mt = 0; mt < MG; ++mt)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int m = (g0 + mt) * 16 + Mfma<R>::row(i, q);
                if (m < M)
                    *(V*)(out + (int64_t)m * ldo + p0 + 4 * r) =
                        V{acc[mt][0][i], acc[mt][1][i], acc[mt][2][i], acc[mt][3][i]};
            }
    }
    };
    if (MODE && ngroup > 128) run(std::true_type{});
    else run(std::false_type{});
}

// dL/dC partials: part[s][m][n] = sum over the points k of range s of
// B[m][k] A[n][k].  One wave per (point range, tile of 16 NT AOs) computes
// every MO tile of that AO tile (groups of up to MG tiles), so each AO row is
// read from HBM once; the waves of one point range run on one XCD (blocks
// are dealt to the 8 XCDs round-robin: range s takes blocks with index
// % 8 == s % 8), where they share that range's B rows through the XCD's L2.
// The K index of lane (r = l%16, q = l/16) in MFMA j of a 16-point step is
// 4q + j, so that each lane reads 4 consecutive points (one vector load) per
// row.  Rows are padded with zeros up to a multiple of 16 points (krange a
// multiple of 16, rows aligned to 4 elements); the rows of the next step are
// loaded while a step is multiplied.  The partials are written (no atomics)
// and summed in FP64 in a fixed order by k_reduce_dc.  With `mask` (rows in
// the fused order, points in 64-point tiles with live-chunk masks
// mask[tile][nw]) the rows of a tile's dead chunks are not read: they count
// as zero (their values are below the screening bound and were not
// evaluated), and a step with both of the wave's chunks dead is skipped.
constexpr int NT = 4;
template <typename R, int MG>
__global__ __launch_bounds__(64, 2) void k_gemm_tn(const R* __restrict__ B, int nmo, const R* __restrict__ A, int nao,
                                                int64_t kend, int64_t ld, int64_t krange, int nsplit,
                                                R* __restrict__ part, const unsigned long long* __restrict__ mask,
                                                int nw, int nchunk)
{
    typedef typename Mfma<R>::v4 V;
    const int ntile = (nao + 16 * NT - 1) / (16 * NT);
    const int xcd = blockIdx.x & 7, local = blockIdx.x >> 3;
    const int t = local % ntile, split = (local / ntile) * 8 + xcd;
    if (split >= nsplit) return;
    const int lane = threadIdx.x;
    const int r = lane & 15, q = lane >> 4;
    const int n0 = t * (16 * NT);
    const int64_t k0 = (int64_t)split * krange;
    const int64_t k1 = min(kend, k0 + krange);
    const V* ap[NT];
    bool nrow[NT];
#pragma unroll
    for (int u = 0; u < NT; ++u) {
        nrow[u] = n0 + 16 * u + r < nao;
        ap[u] = (const V*)(A + (int64_t)(nrow[u] ? n0 + 16 * u + r : 0) * ld + 4 * q);
    }
    R* out = part + (int64_t)split * nmo * nao;
    const int nmt = (nmo + 15) / 16;
    for (int g0 = 0; g0 < nmt; g0 += MG) {
        const V* bp[MG];
        bool mrow[MG];
#pragma unroll
        for (int mt = 0; mt < MG; ++mt) {
            const int m = (g0 + mt) * 16 + r;
            mrow[mt] = m < nmo;
            bp[mt] = (const V*)(B + (int64_t)(mrow[mt] ? m : 0) * ld + 4 * q);
        }
        V acc[MG][NT];
#pragma unroll
        for (int mt = 0; mt < MG; ++mt)
#pragma unroll
            for (int u = 0; u < NT; ++u) acc[mt][u] = V{0, 0, 0, 0};
        // the next step's rows are loaded while this step is multiplied
        V nb[MG], na[NT];
        bool nany = true;
        auto load = [&](int64_t k) {
            bool lv[2] = {true, true};   // the wave's two 32-row chunks live in this point tile
            if (mask) {
                const unsigned long long* mrow = mask + (k >> 6) * nw;
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int c = (n0 >> 5) + h;
                    lv[h] = c < nchunk && ((mrow[c >> 6] >> (c & 63)) & 1);
                }
            }
            nany = lv[0] || lv[1];
            if (nany)
#pragma unroll
                for (int mt = 0; mt < MG; ++mt) nb[mt] = bp[mt][k >> 2];
#pragma unroll
            for (int u = 0; u < NT; ++u) na[u] = lv[u >> 1] ? ap[u][k >> 2] : V{0, 0, 0, 0};
        };
        if (k0 < k1) load(k0);
        for (int64_t k = k0; k < k1; k += 16) {
            // rows past nmo / nao read row 0: they only reach accumulator
            // rows / columns that are not stored
            V vb[MG], va[NT];
#pragma unroll
            for (int mt = 0; mt < MG; ++mt) vb[mt] = nb[mt];
#pragma unroll
            for (int u = 0; u < NT; ++u) va[u] = na[u];
            const bool any = nany;
            if (k + 16 < k1) load(k + 16);
            if (!any) continue;
#pragma unroll
            for (int j = 0; j < 4; ++j)
#pragma unroll
                for (int mt = 0; mt < MG; ++mt)
#pragma unroll
                    for (int u = 0; u < NT; ++u) acc[mt][u] = Mfma<R>::mac(vb[mt][j], va[u][j], acc[mt][u]);
        }
#pragma unroll
        for (int mt = 0; mt < MG; ++mt)
#pragma unroll
            for (int u = 0; u < NT; ++u)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int m = (g0 + mt) * 16 + Mfma<R>::row(i, q), n = n0 + 16 * u + r;
                    if (m < nmo && n < nao) out[(int64_t)m * nao + n] = acc[mt][u][i];
                }
    }
}

// dC[e] += sum_s part[s][e] for e < n over s = 0 .. nsplit - 1 in FP64, in
// a fixed order: eight interleaved partial sums (s mod 8, eight independent
// loads in flight per thread), then combined in index order
template <typename R>
__global__ void k_reduce_dc(const R* part, int nsplit, int64_t n, R* dC)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int s = 0;
    for (; s + 8 <= nsplit; s += 8)
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] += (double)part[(int64_t)(s + j) * n + e];
    for (int j = 0; s + j < nsplit; ++j) v[j] += (double)part[(int64_t)(s + j) * n + e];
    double t = 0.0;
#pragma unroll
    for (int j = 0; j < 8; ++j) t += v[j];
    dC[e] = (R)((double)dC[e] + t);
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) v += __shfl_xor(v, o);
    return v;
}

// Columns of a chunk laid out in 4x4x4 bricks of a whole regular grid, from
// brick bk0 (z fastest within a brick; bricks z fastest, then y, then x),
// lanes of a brick outside the grid repeating its nearest grid point.
struct BwdBricks {
    int on;
    int64_t bk0, nbx, nby, nbz, nx, ny, nz;
};

// grid indices of column `col`; returns whether the point lies in the grid
__device__ __forceinline__ bool brick_index(const BwdBricks& bm, int64_t col, int64_t& ix, int64_t& iy, int64_t& iz)
{
    const int l = (int)(col & 63);
    int64_t t = bm.bk0 + (col >> 6);
    const int64_t bz = t % bm.nbz;
    t /= bm.nbz;
    const int64_t by = t % bm.nby, bx = t / bm.nby;
    ix = 4 * bx + (l >> 4);
    iy = 4 * by + ((l >> 2) & 3);
    iz = 4 * bz + (l & 3);
    const bool in = ix < bm.nx && iy < bm.ny && iz < bm.nz;
    ix = ix < bm.nx ? ix : bm.nx - 1;
    iy = iy < bm.ny ? iy : bm.ny - 1;
    iz = iz < bm.nz ? iz : bm.nz - 1;
    return in;
}

template <typename R>
struct ShellGradArgs {
    const ShellRec* shells;
    const PrimRec* prims;
    const double* T;
    GridDesc grid;
    int64_t p0, np, ld;
    int nprim_total;
    int is_normalized;
    const R* G;          // [nao][ld]
    double* partial;     // [nrange][nprim_total + 3 nshell]
    int64_t ncol;
    // screened (whole regular grid): columns in bricks, G rows in the fused
    // order written only for the live chunks of each 64-point tile
    BwdBricks bm;
    const unsigned long long* mask;   // [tile][nw]
    int nw;
};

// Per-lane contributions of one point to the shell's position gradient and
// to the exponent gradients of primitives [pw0, pw0 + MP), accumulated into
// gr / ga when `near` (the lane's point is within the shell's radius).
// Every primitive's d e is computed once (kept for those of the window).
// w_c += sum over the rows r of a standard spherical shell of G_r T[r][c]
// (the transpose of sph_rows_to: at most three terms per row)
template <int L, int RW, typename R>
__device__ __forceinline__ void sph_rows_back(const R (&gv)[ncart(L)], R (&w)[ncart(L)])
{
    if constexpr (L <= 3 && RW < 2 * L + 1) {
        constexpr SphRow row = sph_row(L, RW);
        const R g = gv[RW];
        w[row.t[0].c] = fma(g, (R)row.t[0].v, w[row.t[0].c]);
        if constexpr (row.n > 1) w[row.t[1].c] = fma(g, (R)row.t[1].v, w[row.t[1].c]);
        if constexpr (row.n > 2) w[row.t[2].c] = fma(g, (R)row.t[2].v, w[row.t[2].c]);
        sph_rows_back<L, RW + 1, R>(gv, w);
    }
}

template <int L, int MP, typename R>
__device__ __forceinline__ void shell_point(const ShellGradArgs<R>& a, const ShellRec& sh, int pw0, int pend,
                                            double px, double py, double pz, int64_t i, double nfac, bool near,
                                            R (&gr)[3], R (&ga)[MP], const double* Ts)
{
    constexpr int NC = ncart(L);
    const double X = px - sh.cx, Y = py - sh.cy, Z = pz - sh.cz;
    const double rr = X * X + Y * Y + Z * Z;
    const R Xr = (R)X, Yr = (R)Y, Zr = (R)Z;
    // the G rows of a standard shell (2L + 1 spherical, NC Cartesian) are
    // loaded before the primitive loop, whose exponentials then cover their
    // latency; the loads are unconditional (rows past the shell's last read
    // its last row, unused)
    const R* Gi = a.G + (int64_t)sh.row0 * a.ld + i;
    R gv[NC];
    if (sh.kind != SHELL_T_GENERAL) {
#pragma unroll
        for (int k = 0; k < NC; ++k) gv[k] = Gi[(int64_t)min(k, sh.nrow - 1) * a.ld];
    }
    R S0 = 0, S1 = 0;
    R dw[MP];
#pragma unroll
    for (int j = 0; j < MP; ++j) dw[j] = 0;
    for (int p = 0; p < sh.nprim; ++p) {
        const PrimRec pr = load_prim(a.prims, sh.prim0 + p);
        // a primitive out of reach of every point of the wave adds nothing
        // above its screening bound
        if (!__builtin_amdgcn_ballot_w64(rr < pr.rcut2)) continue;
        const R de = (R)pr.d * gauss<R>(pr.alpha * rr);
        S0 += de;
        S1 = fma((R)(-2.0 * pr.alpha), de, S1);
#pragma unroll
        for (int j = 0; j < MP; ++j)
            if (p == pw0 + j) dw[j] = de;
    }
    // w_c = sum_rows G_row T[row][c]: the standard spherical and Cartesian
    // maps with compile-time coefficients, a general block from its LDS copy
    R w[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) w[k] = 0;
    if (sh.kind == SHELL_T_SPHERICAL) {
        sph_rows_back<L, 0, R>(gv, w);
    } else if (sh.kind == SHELL_T_CARTESIAN) {
#pragma unroll
        for (int k = 0; k < NC; ++k) w[k] = gv[k] * (R)cart_norm(L, k);
    } else {
        for (int rw = 0; rw < sh.nrow; ++rw) {
            const R g = Gi[(int64_t)rw * a.ld];
#pragma unroll
            for (int k = 0; k < NC; ++k) w[k] = fma(g, (R)Ts[rw * NC + k], w[k]);
        }
    }
    // q = sum_c w_c m_c and, with the lowered monomials u (times S0),
    // sum_c w_c d_a chi_c = R_a S1 q + sum_c w_c e_a(c) u_{c - 1_a}
    ShellPows<L, R> P;
    P.init(Xr, Yr, Zr);
    R m[NC];
    first_mono<L, 0, R>(P, m);
    R q = 0;
#pragma unroll
    for (int k = 0; k < NC; ++k) q = fma(w[k], m[k], q);
    if (pw0 == 0 && near) {
        R u[L > 0 ? ncart(L - 1) : 1];
        lower_mono<L, 0, R>(P, S0, u);
        gr[0] -= fma(Xr * S1, q, lowered_dot<L, 0, 0, R>(w, u, R(0)));
        gr[1] -= fma(Yr * S1, q, lowered_dot<L, 1, 0, R>(w, u, R(0)));
        gr[2] -= fma(Zr * S1, q, lowered_dot<L, 2, 0, R>(w, u, R(0)));
    }
    if (!near) q = 0;
#pragma unroll
    for (int j = 0; j < MP; ++j) {
        if (pw0 + j < pend) {
            const double alpha = load_prim(a.prims, sh.prim0 + pw0 + j).alpha;
            ga[j] = fma(q * dw[j], (R)(nfac / alpha - rr), ga[j]);
        }
    }
}

template <int L, int MP, typename R>
__device__ void shell_grad(const ShellGradArgs<R>& a, const ShellRec& sh, int pw0)
{
    __shared__ double red[4][3 + MP];
    __shared__ double Ts[GEN_T_MAX];   // a general shell's T block (nrow x ncart), read through LDS
    if (sh.kind == SHELL_T_GENERAL) {
        const OKMI_CONST double* T = cptr(a.T) + sh.t0;
        for (int t = threadIdx.x; t < sh.nrow * ncart(L); t += blockDim.x) Ts[t] = T[t];
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int pend = min(sh.nprim, pw0 + MP);
    const double nfac = a.is_normalized ? 0.0 : (2.0 * L + 3.0) / 4.0;
    R gr[3] = {0, 0, 0};
    R ga[MP];
#pragma unroll
    for (int j = 0; j < MP; ++j) ga[j] = 0;
    const int64_t r0 = (int64_t)blockIdx.x * RANGE;
    const int64_t r1 = min(a.np, r0 + RANGE);
    for (int64_t b = r0; b < r1; b += 256) {
        const int64_t i = b + threadIdx.x;
        bool act = i < r1;
        double px, py, pz;
        if (a.bm.on) {
            int64_t ix, iy, iz;
            act = brick_index(a.bm, act ? i : r0, ix, iy, iz) && act;
            px = a.grid.x[ix]; py = a.grid.y[iy]; pz = a.grid.z[iz];
        } else {
            grid_point(a.grid, a.p0 + (act ? i : r0), px, py, pz);
        }
        const double X = px - sh.cx, Y = py - sh.cy, Z = pz - sh.cz;
        bool near = act && X * X + Y * Y + Z * Z < sh.rcut2;
        if (a.mask && near) {   // the shell's rows of G exist where its chunk is live
            const int c = sh.row0 >> 5;
            near = (a.mask[(i >> 6) * a.nw + (c >> 6)] >> (c & 63)) & 1;
        }
        if (!__builtin_amdgcn_ballot_w64(near)) continue;   // no point of this wave reaches the shell
        shell_point<L, MP, R>(a, sh, pw0, pend, px, py, pz, act ? i : r0, nfac, near, gr, ga, Ts);
    }
    // block reduction in FP64, in a fixed order (deterministic)
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double v = wave_sum((double)gr[d]);
        if (lane == 0) red[wave][d] = v;
    }
#pragma unroll
    for (int j = 0; j < MP; ++j) {
        if (pw0 + j < pend) {
            const double v = wave_sum((double)ga[j]);
            if (lane == 0) red[wave][3 + j] = v;
        }
    }
    __syncthreads();
    const int t = threadIdx.x;
    double* out = a.partial + (int64_t)blockIdx.x * a.ncol;
    if (t < 3 && pw0 == 0) {
        out[a.nprim_total + 3 * blockIdx.y + t] = red[0][t] + red[1][t] + red[2][t] + red[3][t];
    } else if (t >= 3 && t < 3 + MP && pw0 + t - 3 < pend) {
        out[sh.prim0 + pw0 + t - 3] = red[0][t] + red[1][t] + red[2][t] + red[3][t];
    }
}

// MP primitives per block (4 when no shell has more, else MAXP; more
// primitives than MP: extra windows along grid z)
template <int MP, typename R>
__global__ __launch_bounds__(256) void k_shell_grad(ShellGradArgs<R> a)
{
    const ShellRec sh = load_shell(a.shells, blockIdx.y);
    const int pw0 = blockIdx.z * MP;
    if (pw0 >= sh.nprim) return;
    switch (sh.l) {
    case 0: shell_grad<0, MP, R>(a, sh, pw0); break;
    case 1: shell_grad<1, MP, R>(a, sh, pw0); break;
    case 2: shell_grad<2, MP, R>(a, sh, pw0); break;
    case 3: shell_grad<3, MP, R>(a, sh, pw0); break;
    default: shell_grad<4, MP, R>(a, sh, pw0); break;
    }
}

// column sums of partial[nrange][ncol] in range order, then added to the
// outputs: exponent columns to dalpha, shell position columns to the
// shell's atom (shells of one atom summed in shell order by one thread)
__global__ void k_reduce_cols(const double* partial, int nrange, int64_t ncol, double* colsum)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncol) return;
    double s = 0.0;
    for (int r = 0; r < nrange; ++r) s += partial[(int64_t)r * ncol + c];
    colsum[c] = s;
}

template <typename R>
__global__ void k_scatter_grad(const double* colsum, const ShellRec* shells, int nshell, int nprim_total,
                               int natoms, R* dalpha, R* dR)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nprim_total) {
        dalpha[t] = (R)((double)dalpha[t] + colsum[t]);
    } else if (t < nprim_total + 3 * natoms) {
        const int a = (t - nprim_total) / 3, d = (t - nprim_total) % 3;
        double s = 0.0;
        for (int sh = 0; sh < nshell; ++sh)
            if (load_shell(shells, sh).atom == a) s += colsum[nprim_total + 3 * sh + d];
        dR[3 * a + d] = (R)((double)dR[3 * a + d] + s);
    }
}

// zero-filled padding of a [rows][ld] buffer's columns [c0, c1) (fill kernel,
// no memset, so the sequence can sit in a captured graph)
template <typename R>
__global__ void k_zero_cols(R* X, int rows, int64_t ld, int64_t c0, int64_t c1)
{
    const int64_t n = c1 - c0;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * rows) return;
    X[(i / n) * ld + c0 + i % n] = R(0);
}

// screened backward helpers (rows in the fused order of the basis, colmap:
// the basis row of each fused row, -1 for padding rows)
template <typename R>
__global__ void k_permute_cols(const R* C, int nmo, int nao, int64_t kpad, const int* colmap, R* out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)nmo * kpad) return;
    const int64_t m = i / kpad, k = i - m * kpad;
    const int c = colmap[k];
    out[i] = c >= 0 ? C[m * nao + c] : R(0);
}

// dC[m][colmap[k]] += dCp[m][k] (each basis column has one fused row: no races)
template <typename R>
__global__ void k_unpermute_add(const R* dCp, int nmo, int64_t kpad, const int* colmap, int nao, R* dC)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)nmo * kpad) return;
    const int64_t m = i / kpad, k = i - m * kpad;
    const int c = colmap[k];
    if (c >= 0) dC[m * nao + c] = (R)((double)dC[m * nao + c] + (double)dCp[i]);
}

template <typename R>
__global__ void k_fill_zero(R* x, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = R(0);
}

// rows[0 .. nrows) of X [.][ld] zeroed over columns [0, ncol)
template <typename R>
__global__ void k_zero_rows_r(R* X, int64_t ld, const int* rows, int nrows, int64_t ncol)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)nrows * ncol) return;
    X[(int64_t)rows[i / ncol] * ld + i % ncol] = R(0);
}

// dL/drho of each brick column (0 for lanes outside the grid)
template <typename R>
__global__ void k_gather_bricks(const R* g, BwdBricks bm, int64_t n, R* out)
{
    const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= n) return;
    int64_t ix, iy, iz;
    const bool in = brick_index(bm, col, ix, iy, iz);
    out[col] = in ? g[(ix * bm.ny + iy) * bm.nz + iz] : R(0);
}

inline size_t pad256(size_t n) { return (n + 255) & ~(size_t)255; }

struct BwdLayout {
    int64_t chunk = 0;
    int nrange = 0;
    int64_t ncol = 0;
    int nsplit_max = 0;   // point ranges of the dL/dC contraction
    int krows = 0;        // rows of A and G: the AO count, or the fused rows of the screened path if more
    size_t ct = 0, wc = 0, wct = 0, a = 0, phi = 0, bm = 0, g = 0, part = 0, cols = 0, dcp = 0, total = 0;   // bytes
    size_t cp = 0, dca = 0, gb = 0, live = 0, lcount = 0, mask = 0;   // screened path
};

// rows of A and G (the screened path stages them in the fused row order)
inline int bwd_krows(const OkmiBasis* b) { return std::max(b->nao, b->nkpad); }

// dL/dC partials: up to 512 point ranges, at most 64 MB of them
constexpr size_t DC_PART_BYTES = (size_t)64 << 20;
int dc_split_max(const OkmiBasis* b, int nmo, size_t es)
{
    const size_t one = (size_t)nmo * bwd_krows(b) * es;
    return (int)std::max<size_t>(1, std::min<size_t>(512, DC_PART_BYTES / std::max<size_t>(one, 1)));
}

// elements of W (M x K) packed by k_pack_w
inline size_t packed_w_elems(int M, int K) { return (size_t)((M + 15) / 16) * ((K + 3) / 4) * 64; }

// es: bytes per value of the arithmetic (8 FP64, 4 FP32)
BwdLayout bwd_layout(const OkmiBasis* b, int nmo, int64_t chunk, size_t es)
{
    BwdLayout L;
    L.chunk = (chunk + 63) & ~(int64_t)63;
    L.nrange = (int)((L.chunk + RANGE - 1) / RANGE);
    L.ncol = b->nprim + 3 * (int64_t)b->nshell;
    L.krows = bwd_krows(b);
    const size_t K = (size_t)L.krows;
    const int64_t ntile = L.chunk / 64;
    const int nw = b->nchunk > 0 ? live_mask_words(b) : 0;
    size_t o = 0;
    L.ct = o; o += pad256(K * nmo * es);
    L.wc = o; o += pad256(packed_w_elems(nmo, L.krows) * es);     // C packed (Phi = C A)
    L.wct = o; o += pad256(packed_w_elems(L.krows, nmo) * es);    // C^T packed (G = C^T B)
    L.a = o; o += pad256(K * L.chunk * es);
    L.phi = o; o += pad256((size_t)nmo * L.chunk * es);
    L.bm = o; o += pad256((size_t)nmo * L.chunk * es);
    L.g = o; o += pad256(K * L.chunk * es);
    L.part = o; o += pad256((size_t)L.nrange * L.ncol * sizeof(double));
    L.cols = o; o += pad256((size_t)L.ncol * sizeof(double));
    // ranges of at least 256 points: a short chunk needs few partials
    L.nsplit_max = (int)std::max<int64_t>(1, std::min<int64_t>(dc_split_max(b, nmo, es), L.chunk / 256));
    L.dcp = o; o += pad256((size_t)L.nsplit_max * nmo * K * es);
    // screened path: C in the fused row order, dL/dC in that order, dL/drho
    // in brick order, live-chunk lists, counts and masks per 64-point tile
    L.cp = o; o += pad256((size_t)nmo * K * es);
    L.dca = o; o += pad256((size_t)nmo * K * es);
    L.gb = o; o += pad256((size_t)L.chunk * es);
    L.live = o; o += pad256((size_t)ntile * std::max(b->nchunk, 1) * sizeof(int));
    L.lcount = o; o += pad256((size_t)ntile * sizeof(int));
    L.mask = o; o += pad256((size_t)ntile * std::max(nw, 1) * sizeof(unsigned long long));
    L.total = o;
    return L;
}

size_t ws_bytes_of(const OkmiBasis* b, int nmo, int64_t chunk, size_t es)
{
    return bwd_layout(b, nmo, chunk, es).total + 256;
}

// out[M][p] = W[M][K] X[K][p] for the points of [0, npad), W packed by
// k_pack_w: M tiles in groups of at most 3 (FP64) or 5 (FP32) tiles, so that
// the accumulators leave room for two waves per SIMD, the group size chosen
// so that the groups divide the tiles as evenly as possible.  mode 1 / 2:
// the screened forms of k_gemm_nn (live-chunk lists per 64-point tile).
template <typename R>
void launch_gemm_nn(const R* Wp, int M, int K, const R* X, int64_t ldx, int64_t npad, R* out, int64_t ldo,
                    hipStream_t st, int mode = 0, const int* live = nullptr, const int* lcount = nullptr,
                    int nchunk = 0)
{
    const dim3 grd((unsigned)((npad + 255) / 256)), blk(256);
    const int nks = (K + 3) / 4;
    if (nks <= 16) {   // short K (the MO count): X rows held in registers
#define OKMI_GXR(MG, MODE, KS) \
    hipLaunchKernelGGL((k_gemm_nn_xr<R, MG, MODE, KS>), grd, blk, 0, st, Wp, M, K, X, ldx, npad, out, ldo, live, lcount, \
                       nchunk)
#define OKMI_GXR_KS(MG, MODE)                      \
    do {                                           \
        if (nks <= 4) OKMI_GXR(MG, MODE, 4);       \
        else if (nks <= 8) OKMI_GXR(MG, MODE, 8);  \
        else if (nks <= 12) OKMI_GXR(MG, MODE, 12); \
        else OKMI_GXR(MG, MODE, 16);               \
    } while (0)
        if (mode == 2) OKMI_GXR_KS(2, 2);
        else if (mode == 0) OKMI_GXR_KS(2, 0);
#undef OKMI_GXR_KS
#undef OKMI_GXR
        if (mode != 1) return;
    }
#define OKMI_GNN(MG, MODE) \
    hipLaunchKernelGGL((k_gemm_nn<R, MG, MODE>), grd, blk, 0, st, Wp, M, K, X, ldx, npad, out, ldo, live, lcount, nchunk)
    if (mode == 2) {
        OKMI_GNN(2, 2);
        return;
    }
    const int nmt = (M + 15) / 16;
    const int gmax = sizeof(R) == 8 ? 3 : 5;
    const int ngroup = (nmt + gmax - 1) / gmax;
    const int mg = (nmt + ngroup - 1) / ngroup;
    switch (mg) {
    case 1: if (mode) OKMI_GNN(1, 1); else OKMI_GNN(1, 0); break;
    case 2: if (mode) OKMI_GNN(2, 1); else OKMI_GNN(2, 0); break;
    case 3: if (mode) OKMI_GNN(3, 1); else OKMI_GNN(3, 0); break;
    case 4:
        if constexpr (sizeof(R) == 4) {
            if (mode) OKMI_GNN(4, 1); else OKMI_GNN(4, 0);
        }
        break;
    default:
        if constexpr (sizeof(R) == 4) {
            if (mode) OKMI_GNN(5, 1); else OKMI_GNN(5, 0);
        }
        break;
    }
#undef OKMI_GNN
}

template <typename R>
int launch_backward(const OkmiBasis* b, const GridDesc& g, int64_t p0, int64_t np, const R* d_c, const R* d_occ,
                    int nmo, const R* d_grad, R* d_dc, R* d_dalpha, R* d_dr, void* ws, size_t ws_bytes,
                    hipStream_t st)
{
    constexpr size_t es = sizeof(R);
    const char* name = es == 8 ? "rho_backward" : "rho_backward_f32";
    if (np <= 0 || nmo <= 0) return 0;
    // largest chunk (a multiple of 64 points, at most np rounded up) whose layout fits
    int64_t lo = 64, hi = (np + 63) & ~(int64_t)63;
    if (!ws || ws_bytes_of(b, nmo, lo, es) > ws_bytes) {
        okmi_set_error("%s: workspace of %zu bytes is too small (need at least %zu)", name, ws_bytes,
                       ws_bytes_of(b, nmo, lo, es));
        return 1;
    }
    while (lo < hi) {
        const int64_t mid = ((lo + hi) / 2 + 64) & ~(int64_t)63;
        if (mid > hi) break;
        if (ws_bytes_of(b, nmo, mid, es) <= ws_bytes) lo = mid;
        else hi = mid - 64;
    }
    const BwdLayout Lo = bwd_layout(b, nmo, lo, es);
    const int64_t chunk = Lo.chunk;
    for (int s = 0; s < b->nshell; ++s)
        if (b->h_shells[s].nprim > 256 * MAXP) {
            okmi_set_error("%s: shell %d has %d primitives (up to %d)", name, s, b->h_shells[s].nprim, 256 * MAXP);
            return 1;
        }
    int maxp = 1;
    for (int s = 0; s < b->nshell; ++s) {
        const ShellRec& sh = b->h_shells[s];
        maxp = std::max(maxp, sh.nprim);
        if (sh.kind == SHELL_T_GENERAL && sh.nrow * ncart(sh.l) > GEN_T_MAX) {
            okmi_set_error("%s: shell %d has a %d x %d transform block (up to %d elements)", name, s, sh.nrow,
                           ncart(sh.l), GEN_T_MAX);
            return 1;
        }
    }
    char* w = (char*)(((uintptr_t)ws + 255) & ~(uintptr_t)255);
    R* CT = (R*)(w + Lo.ct);
    R* WC = (R*)(w + Lo.wc);
    R* WCT = (R*)(w + Lo.wct);
    R* A = (R*)(w + Lo.a);
    R* Phi = (R*)(w + Lo.phi);
    R* B = (R*)(w + Lo.bm);
    R* G = (R*)(w + Lo.g);
    double* part = (double*)(w + Lo.part);
    double* cols = (double*)(w + Lo.cols);
    R* dcp = (R*)(w + Lo.dcp);
    // Screened form for a whole regular grid: points in 4x4x4 bricks, AO rows
    // in the fused order of the basis (32-row chunks), and per 64-point tile
    // only the chunks with a shell in reach (the bound of the fused kernels)
    // are evaluated (A), contracted (Phi = C A, dL/dC = B A^T) and produced
    // (G = C^T B); the shell gradients read G where the shell's chunk is live.
    const bool scr = g.regular && p0 == 0 && np == g.nx * g.ny * g.nz && b->nchunk > 0 && b->d_ishells &&
                     b->nkpad > 0 && b->nkpad <= Lo.krows;
    const int K = scr ? b->nkpad : b->nao;   // rows of A and G
    // the FP32 backward screens with the bound of the paths ending in FP32
    const bool split = es == 4;
    const R* dg = d_grad;
    R* Cp = (R*)(w + Lo.cp);
    R* dCa = (R*)(w + Lo.dca);
    R* gb = (R*)(w + Lo.gb);
    int* live = (int*)(w + Lo.live);
    int* lcount = (int*)(w + Lo.lcount);
    unsigned long long* lmask = (unsigned long long*)(w + Lo.mask);
    const int nw = scr ? live_mask_words(b) : 0;
    BwdBricks bm{};
    int64_t total = np;
    if (scr) {
        bm.on = 1;
        bm.nbx = (g.nx + 3) / 4; bm.nby = (g.ny + 3) / 4; bm.nbz = (g.nz + 3) / 4;
        bm.nx = g.nx; bm.ny = g.ny; bm.nz = g.nz;
        total = bm.nbx * bm.nby * bm.nbz * 64;
    }
    OkmiProfScope prof(name, st);
    const R* Cw = d_c;   // C in the row order of A
    if (scr) {
        const int64_t ne = (int64_t)nmo * K;
        hipLaunchKernelGGL(k_permute_cols<R>, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, st, d_c, nmo, b->nao,
                           (int64_t)K, (const int*)b->d_colmap, Cp);
        hipLaunchKernelGGL(k_fill_zero<R>, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, st, dCa, ne);
        // rows no shell writes (padding of the chunks) are zero in A for good
        if (b->npadrows > 0)
            hipLaunchKernelGGL(k_zero_rows_r<R>, dim3((unsigned)(((int64_t)b->npadrows * chunk + 255) / 256)),
                               dim3(256), 0, st, A, chunk, (const int*)b->d_padrows, b->npadrows, chunk);
        Cw = Cp;
    }
    hipLaunchKernelGGL(k_transpose<R>, dim3((unsigned)(((int64_t)nmo * K + 255) / 256)), dim3(256), 0, st, Cw, nmo,
                       K, CT);
    hipLaunchKernelGGL(k_pack_w<R>, dim3((unsigned)((packed_w_elems(nmo, K) + 255) / 256)), dim3(256), 0, st, Cw, nmo,
                       K, WC);
    hipLaunchKernelGGL(k_pack_w<R>, dim3((unsigned)((packed_w_elems(K, nmo) + 255) / 256)), dim3(256), 0, st, CT, K,
                       nmo, WCT);
    const int set0 = 0;
    for (int64_t c0 = 0; c0 < total; c0 += chunk) {
        const int64_t n = total - c0 < chunk ? total - c0 : chunk;
        const int64_t n16 = (n + 15) & ~(int64_t)15;
        const int64_t n64 = (n + 63) & ~(int64_t)63;
        int rc;
        if (scr) {
            bm.bk0 = c0 / 64;
            rc = okmi_launch_live_chunks(b, g, 1, bm.bk0, 0, n, 64, live, lcount, st, lmask, split);
            if (rc) return rc;
            const AoBricks bk{bm.bk0, n / 64, bm.nbx, bm.nby, bm.nbz};
            const AoLiveMask am{lmask, nw, 0, split ? b->d_wshells_s : nullptr};
            rc = okmi_launch_ao_eval_ex(b, split ? b->d_ishells_s : b->d_ishells, split ? b->d_fprims_s : b->d_fprims,
                                        g, 0, n, &bk, &set0, 1, es == 8 ? OKMI_F64 : OKMI_F32, A, chunk, 0, st, &am);
            if (rc) return rc;
            hipLaunchKernelGGL(k_gather_bricks<R>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, d_grad, bm, n,
                               gb);
            dg = gb;
        } else {
            rc = okmi_launch_ao_eval(b, g, p0 + c0, n, &set0, 1, es == 8 ? OKMI_F64 : OKMI_F32, A, chunk, 0, st);
            if (rc) return rc;
            if (n16 > n)
                hipLaunchKernelGGL(k_zero_cols<R>, dim3((unsigned)((b->nao * (n16 - n) + 255) / 256)), dim3(256), 0,
                                   st, A, b->nao, chunk, n, n16);
            dg = d_grad + c0;
        }
        {
            OkmiProfScope p1(es == 8 ? "gemm_nn" : "gemm_nn_f32", st);
            launch_gemm_nn<R>(WC, nmo, K, A, chunk, n64, Phi, chunk, st, scr ? 1 : 0, live, lcount, b->nchunk);
        }
        hipLaunchKernelGGL(k_bmat<R>, dim3((unsigned)((n16 + 255) / 256)), dim3(256), 0, st, Phi, d_occ, dg, nmo, n,
                           n16, chunk, B);
        // point ranges: about 2048 waves over the chip, at least 256 points
        // per range, at most the partial buffer's ranges
        const int64_t tiles = (K + 16 * NT - 1) / (16 * NT);
        int64_t nsplit = std::max<int64_t>(1, std::min<int64_t>({(2048 + tiles - 1) / tiles, n16 / 256,
                                                                 (int64_t)Lo.nsplit_max}));
        int64_t krange = ((n16 + nsplit - 1) / nsplit + 15) & ~(int64_t)15;
        if (scr) krange = (krange + 63) & ~(int64_t)63;   // whole 64-point tiles per range
        nsplit = (n16 + krange - 1) / krange;
        // blocks: (range group of 8, AO tile, XCD); range = group * 8 + XCD
        const unsigned nblk = (unsigned)(((nsplit + 7) / 8) * tiles * 8);
        const int nmt = (nmo + 15) / 16;
        const unsigned long long* tn_mask = scr ? lmask : nullptr;
        const int nchunk_tn = b->nchunk;
        {
            OkmiProfScope p2(es == 8 ? "gemm_tn" : "gemm_tn_f32", st);
            if (nmt == 1)
                hipLaunchKernelGGL((k_gemm_tn<R, 1>), dim3(nblk), dim3(64), 0, st, B, nmo, A, K, n16, chunk,
                                   krange, (int)nsplit, dcp, tn_mask, nw, nchunk_tn);
            else if (nmt == 2)
                hipLaunchKernelGGL((k_gemm_tn<R, 2>), dim3(nblk), dim3(64), 0, st, B, nmo, A, K, n16, chunk,
                                   krange, (int)nsplit, dcp, tn_mask, nw, nchunk_tn);
            else
                hipLaunchKernelGGL((k_gemm_tn<R, 3>), dim3(nblk), dim3(64), 0, st, B, nmo, A, K, n16, chunk,
                                   krange, (int)nsplit, dcp, tn_mask, nw, nchunk_tn);
        }
        const int64_t ndc = (int64_t)nmo * K;
        hipLaunchKernelGGL(k_reduce_dc<R>, dim3((unsigned)((ndc + 255) / 256)), dim3(256), 0, st, dcp, (int)nsplit,
                           ndc, scr ? dCa : d_dc);
        {
            OkmiProfScope p1(es == 8 ? "gemm_nn" : "gemm_nn_f32", st);
            launch_gemm_nn<R>(WCT, K, nmo, B, chunk, n64, G, chunk, st, scr ? 2 : 0, live, lcount, b->nchunk);
        }
        ShellGradArgs<R> sa{};
        sa.shells = scr ? (split ? b->d_ishells_s : b->d_ishells) : b->d_shells;
        sa.prims = scr ? (split ? b->d_fprims_s : b->d_fprims) : b->d_prims;
        sa.T = b->d_T; sa.grid = g;
        sa.p0 = p0 + c0; sa.np = n; sa.ld = chunk; sa.nprim_total = b->nprim; sa.is_normalized = b->is_normalized;
        sa.G = G; sa.partial = part; sa.ncol = Lo.ncol;
        sa.bm = bm;
        sa.mask = scr ? lmask : nullptr;
        sa.nw = nw;
        const int nrange = (int)((n + RANGE - 1) / RANGE);
        {
            OkmiProfScope p3(es == 8 ? "shell_grad" : "shell_grad_f32", st);
            if (maxp <= 4)
                hipLaunchKernelGGL((k_shell_grad<4, R>), dim3((unsigned)nrange, (unsigned)b->nshell, 1u), dim3(256),
                                   0, st, sa);
            else
                hipLaunchKernelGGL((k_shell_grad<MAXP, R>),
                                   dim3((unsigned)nrange, (unsigned)b->nshell, (unsigned)((maxp + MAXP - 1) / MAXP)),
                                   dim3(256), 0, st, sa);
        }
        hipLaunchKernelGGL(k_reduce_cols, dim3((unsigned)((Lo.ncol + 255) / 256)), dim3(256), 0, st, part, nrange,
                           Lo.ncol, cols);
        hipLaunchKernelGGL(k_scatter_grad<R>, dim3((unsigned)((b->nprim + 3 * b->natoms + 255) / 256)), dim3(256), 0,
                           st, cols, b->d_shells, b->nshell, b->nprim, b->natoms, d_dalpha, d_dr);
    }
    if (scr) {
        const int64_t ne = (int64_t)nmo * K;
        hipLaunchKernelGGL(k_unpermute_add<R>, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, st, dCa, nmo,
                           (int64_t)K, (const int*)b->d_colmap, b->nao, d_dc);
    }
    return okmi_check_launch(name);
}

}  // namespace

size_t okmi_backward_ws_bytes(const OkmiBasis* b, int nmo, int64_t chunk)
{
    return ws_bytes_of(b, nmo, chunk, sizeof(double));
}

size_t okmi_backward_f32_ws_bytes(const OkmiBasis* b, int nmo, int64_t chunk)
{
    return ws_bytes_of(b, nmo, chunk, sizeof(float));
}

int okmi_launch_rho_backward(const OkmiBasis* b, const GridDesc& g, int64_t p0, int64_t np, const double* d_c,
                             const double* d_occ, int nmo, const double* d_grad, double* d_dc, double* d_dalpha,
                             double* d_dr, void* ws, size_t ws_bytes, hipStream_t st)
{
    return launch_backward<double>(b, g, p0, np, d_c, d_occ, nmo, d_grad, d_dc, d_dalpha, d_dr, ws, ws_bytes, st);
}

int okmi_launch_rho_backward_f32(const OkmiBasis* b, const GridDesc& g, int64_t p0, int64_t np, const float* d_c,
                                 const float* d_occ, int nmo, const float* d_grad, float* d_dc, float* d_dalpha,
                                 float* d_dr, void* ws, size_t ws_bytes, hipStream_t st)
{
    return launch_backward<float>(b, g, p0, np, d_c, d_occ, nmo, d_grad, d_dc, d_dalpha, d_dr, ws, ws_bytes, st);
}
