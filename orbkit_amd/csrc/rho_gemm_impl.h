// FP64 MO contraction on MFMA (v_mfma_f64_16x16x4f64) with the density
// assembly in the epilogue.
//
//   Phi_s[m][p] = sum_k C[m][k] A_s[k][p]         (orbkit/cy_core.pyx:80-101)
//   rho[p]      = sum_m occ_m Phi_0[m][p]^2       (orbkit/core.py:266-274)
//   delta_r[p]  = 2 sum_m occ_m (Phi_0 Phi_d + Phi_a Phi_b)   (orbkit/core.py:280-304)
//
// A workgroup is 4 waves; each wave owns 16 points and loops over MO tiles
// of 16 in groups of MG, accumulating every AO set (NS) for those tiles in
// registers.  In density mode the MOs are reduced in registers and never
// written; in MO mode Phi is stored [set][mo][point].  With more than one MO
// group (density mode) the waves of a workgroup instead share 16 points and
// split the MO tiles between them in even contiguous shares (one wave per
// group up to 4 groups, 8 waves for 5 to 8, else 4 waves), so that the AO
// rows of a K step are fetched once for several groups.
#pragma once
#include "okmi_internal.h"
#include <type_traits>

namespace okmi_gemm {

typedef double d4 __attribute__((ext_vector_type(4)));

struct GemmArgs {
    const double* A;
    int64_t ld, set_stride;
    int nao;
    const double* C;
    int nmo;
    const double* occ;
    int64_t np;
    double* mo_out;
    int64_t ld_mo;
    double* rho;
    double* delta;
    int64_t ld_delta;
    double* mo_norm;
    // density mode: accumulators are indexed by derivative code
    int set_mask;            // codes whose AO set is present
    int out_mask;            // codes with a density output
    int64_t set_off[OKMI_MAX_SETS];   // offset of code's AO set in A
    int nout;
    int out_code[16];        // derivative code of each output row
    // density mode, screened (okmi_launch_rho_f64 with F64Screen): A rows and
    // C columns in the fused row order, each 64-column block contracts its
    // live chunks only, brick-ordered columns mapped to grid points for the
    // outputs
    int screened;
    F64Screen scr;
    // density mode: the W waves of a workgroup (W = blockDim / 64) share 16
    // points and split the MO tiles in contiguous shares of nmt / W (A rows
    // read by W waves together, not once per group), their sums added in
    // wave order at the end; else 4 waves each own 16 points of the
    // workgroup's 64 and run every group
    int wave_groups;
};

template <int MG, int NS>
__global__ __launch_bounds__(256) void k_gemm_f64(GemmArgs a)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int n = lane & 15, kq = lane >> 4;
    const int64_t p = (int64_t)blockIdx.x * 64 + wave * 16 + n;
    const bool pin = p < a.np;
    const int nmt = (a.nmo + 15) >> 4;
    const int ngroups = (nmt + MG - 1) / MG;

    for (int g = 0; g < ngroups; ++g) {
        d4 acc[MG][NS];
#pragma unroll
        for (int mt = 0; mt < MG; ++mt)
#pragma unroll
            for (int s = 0; s < NS; ++s) acc[mt][s] = d4{0.0, 0.0, 0.0, 0.0};

        const double* cbase[MG];
        bool min_[MG];
#pragma unroll
        for (int mt = 0; mt < MG; ++mt) {
            const int m = (g * MG + mt) * 16 + n;
            min_[mt] = m < a.nmo;
            cbase[mt] = a.C + (int64_t)(min_[mt] ? m : 0) * a.nao;
        }
        const double* abase = a.A + (pin ? p : 0);

        for (int k0 = 0; k0 < a.nao; k0 += 4) {
            const int k = k0 + kq;
            const bool kin = k < a.nao;
            double av[MG], bv[NS];
#pragma unroll
            for (int mt = 0; mt < MG; ++mt) av[mt] = (kin && min_[mt]) ? cbase[mt][k] : 0.0;
#pragma unroll
            for (int s = 0; s < NS; ++s)
                bv[s] = (kin && pin) ? abase[s * a.set_stride + (int64_t)k * a.ld] : 0.0;
#pragma unroll
            for (int mt = 0; mt < MG; ++mt)
#pragma unroll
                for (int s = 0; s < NS; ++s)
                    acc[mt][s] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[mt], bv[s], acc[mt][s], 0, 0, 0);
        }

        // v_mfma_f64_16x16x4f64 leaves D[4*i + lane/16][lane%16] in item i,
        // so this lane holds Phi_s[m = tile*16 + 4*i + kq][p]
#pragma unroll
        for (int mt = 0; mt < MG; ++mt) {
            const int mb = (g * MG + mt) * 16 + kq;
#pragma unroll
            for (int s = 0; s < NS; ++s)
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    if (pin && mb + 4 * i < a.nmo)
                        a.mo_out[(int64_t)s * a.nmo * a.ld_mo + (int64_t)(mb + 4 * i) * a.ld_mo + p] = acc[mt][s][i];
        }
    }
}

template <int MG>
int launch_ns(const GemmArgs& a, int nsets, hipStream_t st)
{
    dim3 blk(256), grd((unsigned)((a.np + 63) / 64));
    switch (nsets) {
#define OKMI_NS_CASE(NS)                                                          \
    case NS:                                                                      \
        if constexpr (MG * NS <= 24) {                                            \
            hipLaunchKernelGGL((k_gemm_f64<MG, NS>), grd, blk, 0, st, a);   \
            return 0;                                                             \
        }                                                                         \
        break;
        OKMI_NS_CASE(1) OKMI_NS_CASE(2) OKMI_NS_CASE(3) OKMI_NS_CASE(4) OKMI_NS_CASE(5)
        OKMI_NS_CASE(6) OKMI_NS_CASE(7) OKMI_NS_CASE(8) OKMI_NS_CASE(9) OKMI_NS_CASE(10)
#undef OKMI_NS_CASE
    default: break;
    }
    okmi_set_error("gemm_f64: unsupported set count %d for MO group %d", nsets, MG);
    return 1;
}

// MO tiles per group: at most 24 accumulator tiles (MG x sets, 4 doubles
// each), which keeps every accumulator in VGPRs with 2 waves per SIMD; more
// (4 x 7 and up) overflowed into AGPR copies at 1 wave per SIMD
inline int mo_group(int nsets, int nmt)
{
    int mg = 24 / (nsets > 0 ? nsets : 1);
    if (mg > 4) mg = 4;
    if (mg < 1) mg = 1;
    return nmt < mg ? nmt : mg;
}

inline int launch_mg(const GemmArgs& a, int nsets, hipStream_t st)
{
    const int mg = mo_group(nsets, (a.nmo + 15) / 16);
    switch (mg) {
    case 1: return launch_ns<1>(a, nsets, st);
    case 2: return launch_ns<2>(a, nsets, st);
    case 3: return launch_ns<3>(a, nsets, st);
    default: return launch_ns<4>(a, nsets, st);
    }
}


// dsum[D] += 2 w (f0 f_D + e_D) for every output code D >= 1, unrolled at
// compile time (e_D: f_a^2 for D = aa, f_a f_b for D = ab)
template <int D, int NSC>
__device__ __forceinline__ void rho_terms(const d4 (&acc)[NSC], int i, double f0, double w, int out_mask,
                                          double (&dsum)[NSC])
{
    if constexpr (D < NSC) {
        if (out_mask & (1 << D)) {
            double t = f0 * acc[D][i];
            if constexpr (D >= 4 && D <= 6) t += acc[D - 3][i] * acc[D - 3][i];
            if constexpr (D == 7) t += acc[1][i] * acc[2][i];
            if constexpr (D == 8) t += acc[1][i] * acc[3][i];
            if constexpr (D == 9) t += acc[2][i] * acc[3][i];
            dsum[D] += 2.0 * w * t;
        }
        rho_terms<D + 1, NSC>(acc, i, f0, w, out_mask, dsum);
    }
}

// Density mode.  Accumulators are indexed by derivative code (NSC = 1, 4, 7
// or 10 codes 0..NSC-1); codes absent from set_mask issue no loads and no
// MFMAs (uniform branches).  Per point, with w = occ:
//   rho   += w f0^2
//   d_c   += 2 w (f0 f_c + e_c),  e_c = f_a^2 (c = aa), f_a f_b (c = ab), 0 (c = a)
// FULL: every code 0..NSC-1 present (no per-set branches) and the loads
// unconditional (rows clamped, the zeroing done by selects), a straight
// sequence of NT x NSC MFMAs per K step (NT <= MG tiles in the pass)
// LONG: live lists may exceed 128 entries (more than 128 chunks of 32 rows),
// and their tail is read from memory
template <int MG, int NSC, bool FULL, bool LONG>
__global__ __launch_bounds__(512) void k_rho_f64(GemmArgs a)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
    const int n = lane & 15, kq = lane >> 4;
    const bool wg = a.wave_groups;
    const int64_t p = wg ? (int64_t)blockIdx.x * 16 + n : (int64_t)blockIdx.x * 64 + wave * 16 + n;
    bool pin = p < a.np;
    int64_t q = p;   // output index of this lane's point
    if (a.screened && a.scr.brick) {
        const int l = (int)(p & 63);
        int64_t bk = a.scr.bk0 + (p >> 6);
        const int64_t bz = bk % a.scr.nbz;
        bk /= a.scr.nbz;
        const int64_t by = bk % a.scr.nby, bx = bk / a.scr.nby;
        const int64_t ix = 4 * bx + (l >> 4), iy = 4 * by + ((l >> 2) & 3), iz = 4 * bz + (l & 3);
        pin = pin && ix < a.scr.nx && iy < a.scr.ny && iz < a.scr.nz;   // brick lanes outside the grid add nothing
        q = (ix * a.scr.ny + iy) * a.scr.nz + iz;
    }
    // K steps of 4 rows: every row, or the 8 steps of each live chunk of
    // the 64-column tile holding this block's columns
    const int* lst = nullptr;
    int nsteps = (a.nao + 3) / 4;
    // the list's first 128 entries are held across the wave's lanes (entry j
    // in lane j % 64 of lv[j / 64]) and read with readlane, so that a K step
    // does not wait on a scalar load of its chunk index
    int lv0 = 0, lv1 = 0;
    if (a.screened) {
        const int64_t tile = wg ? blockIdx.x >> 2 : blockIdx.x;
        lst = a.scr.live + tile * a.scr.nchunk;
        const int cnt = a.scr.live_count[tile];
        nsteps = 8 * cnt;
        lv0 = lane < cnt ? lst[lane] : 0;
        lv1 = lane + 64 < cnt ? lst[lane + 64] : 0;
    }
    // entry j of the list (wave-uniform j).  Without LONG by readlane alone:
    // an entry that might come from memory puts each K step behind a branch
    // whose join waits for every load in flight (the prefetched operands
    // included)
    auto chunk_of = [&](int j) {
        if constexpr (LONG)
            return j < 64 ? __builtin_amdgcn_readlane(lv0, j) : j < 128 ? __builtin_amdgcn_readlane(lv1, j - 64) : lst[j];
        else
            return j < 64 ? __builtin_amdgcn_readlane(lv0, j) : __builtin_amdgcn_readlane(lv1, j - 64);
    };
    const int ldc = a.screened ? a.scr.ldc : a.nao;
    const int nmt = (a.nmo + 15) >> 4;
    // this wave's MO tiles.  With four-tile groups where a pass short of MG
    // tiles can run the instance of its exact count (the accumulators leave
    // room: MG x NSC <= 16): all nmt tiles, or (wave groups) a contiguous
    // share as even as the wave count allows, so that no wave runs MFMAs on
    // tiles past the last MO and the waves sharing a SIMD carry about the
    // same work.  Otherwise groups of MG tiles, round-robin over the waves,
    // the last group masked past the last MO: the larger instances sit at
    // the register limit, where the extra code spilled, and with fewer than
    // four tiles per group a single group always holds exactly the nmt tiles
    // (MG = nmt) while several groups only occur beyond 16 products per K
    // step (no exact-count passes there).
    constexpr bool EXACT = MG == 4 && MG * NSC <= 16;
    const int t_beg = wg ? (int)((int64_t)nmt * wave / nwave) : 0;
    const int t_end = wg ? (int)((int64_t)nmt * (wave + 1) / nwave) : nmt;

    double rho_acc = 0.0;
    double dsum[NSC];
#pragma unroll
    for (int d = 0; d < NSC; ++d) dsum[d] = 0.0;

    // one pass over the K steps for the NT (<= MG) tiles from t0
    auto pass = [&](const int t0, auto ntc) {
        constexpr int NT = decltype(ntc)::value;
        d4 acc[NT][NSC];
#pragma unroll
        for (int mt = 0; mt < NT; ++mt)
#pragma unroll
            for (int d = 0; d < NSC; ++d) acc[mt][d] = d4{0.0, 0.0, 0.0, 0.0};
        const double* cbase[NT];
        bool min_[NT];
#pragma unroll
        for (int mt = 0; mt < NT; ++mt) {
            const int m = (t0 + mt) * 16 + n;
            min_[mt] = m < a.nmo;
            cbase[mt] = a.C + (int64_t)(min_[mt] ? m : 0) * ldc;
        }
        const double* abase = a.A + (p < a.np ? p : 0);
        // A row of K step ks for this lane, and the C column it pairs with
        // (the same index: C is given in the A row order)
        auto row_of = [&](int ks, int& k, bool& kin, int& kcol, bool& cin) {
            k = (a.screened ? chunk_of(ks >> 3) * 32 + 4 * (ks & 7) : 4 * ks) + kq;
            kin = k < a.nao;
            kcol = k;
            cin = kin;
        };
        if constexpr (FULL) {
            // Rows past the end read row nao - 1 and are zeroed through av;
            // points past the end read point 0 and are zeroed through bv
            // (they must add nothing to mo_norm).  The zeroing multiplies by
            // 0 or 1 (the operands are finite) instead of selecting: the
            // compiler turns a select of a load into a masked load in its own
            // branch with its own wait, which serialised the K step's loads.
            // Up to 16 accumulator tiles the operands of step ks + 1 are
            // loaded before the MFMAs of step ks (raw loads; the masks
            // applied when they are used); with more, the registers of a
            // second operand set would spill.
            // K steps whose operands are in flight: two with groups of up to
            // three MO tiles, one with four-tile groups (whose waves share a
            // workgroup's AO rows: waves with one and with two steps in
            // flight side by side ran the large workload 9% slower, two in
            // every wave 1% slower), none beyond 16 accumulator tiles (their
            // registers would spill)
            constexpr int DEPTH = NT * NSC > 16 ? 0 : MG == 4 ? 1 : 2;
            constexpr int NSET = DEPTH == 2 ? 3 : 1;
            const double pm = pin ? 1.0 : 0.0;
            // operands of K step ks into set st (a step past the end reads
            // step nsteps - 1 again: in bounds, never used)
            double ra[NSET][NT], rb[NSET][NSC], rm[NSET];
            auto fetch = [&](int ks, int st) {
                int k, kcol;
                bool kin, cin;
                row_of(ks < nsteps ? ks : nsteps - 1, k, kin, kcol, cin);
                const int kc = kin ? k : a.nao - 1;
                const int cc = cin ? kcol : 0;
#pragma unroll
                for (int mt = 0; mt < NT; ++mt) ra[st][mt] = cbase[mt][cc];
#pragma unroll
                for (int d = 0; d < NSC; ++d) rb[st][d] = abase[a.set_off[d] + (int64_t)kc * a.ld];
                rm[st] = cin ? 1.0 : 0.0;
            };
            auto prep = [&](int st, double (&av)[NT], double (&bv)[NSC]) {
#pragma unroll
                for (int mt = 0; mt < NT; ++mt) av[mt] = ra[st][mt] * (min_[mt] ? rm[st] : 0.0);
#pragma unroll
                for (int d = 0; d < NSC; ++d) bv[d] = rb[st][d] * pm;
            };
            auto mma = [&](const double (&av)[NT], const double (&bv)[NSC]) {
#pragma unroll
                for (int d = 0; d < NSC; ++d)
#pragma unroll
                    for (int mt = 0; mt < NT; ++mt)
                        acc[mt][d] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[mt], bv[d], acc[mt][d], 0, 0, 0);
                // nothing of the next step is scheduled among the MFMAs: a
                // copy of a prefetched operand there would wait for its load
                __builtin_amdgcn_sched_barrier(0);
            };
            auto step = [&](int st) {
                double av[NT], bv[NSC];
                prep(st, av, bv);
                mma(av, bv);
            };
            if constexpr (DEPTH == 2) {
                // three operand sets in turn: the loads of step ks + 2 are
                // issued before the MFMAs of step ks, so two steps of MFMAs
                // (and the other waves') cover their latency; no operand is
                // copied between sets (a copy waits for its load)
                if (nsteps > 0) {
                    fetch(0, 0);
                    fetch(1, 1);
                }
                int ks = 0;
                for (; ks + 3 <= nsteps; ks += 3) {
                    fetch(ks + 2, 2);
                    step(0);
                    fetch(ks + 3, 0);
                    step(1);
                    fetch(ks + 4, 1);
                    step(2);
                }
                if (ks < nsteps) step(0);       // the one or two steps left are in sets 0 and 1
                if (ks + 1 < nsteps) step(1);
            } else if constexpr (DEPTH == 1) {
                // one set: the next step's loads are issued once the step's
                // operands are taken, before its MFMAs
                double ra1[NT], rb1[NSC], rm1 = 0.0;
                auto fetch1 = [&](int ks) {
                    int k, kcol;
                    bool kin, cin;
                    row_of(ks, k, kin, kcol, cin);
                    const int kc = kin ? k : a.nao - 1;
                    const int cc = cin ? kcol : 0;
#pragma unroll
                    for (int mt = 0; mt < NT; ++mt) ra1[mt] = cbase[mt][cc];
#pragma unroll
                    for (int d = 0; d < NSC; ++d) rb1[d] = abase[a.set_off[d] + (int64_t)kc * a.ld];
                    rm1 = cin ? 1.0 : 0.0;
                };
                if (nsteps > 0) fetch1(0);
                for (int ks = 0; ks < nsteps; ++ks) {
                    double av[NT], bv[NSC];
#pragma unroll
                    for (int mt = 0; mt < NT; ++mt) av[mt] = ra1[mt] * (min_[mt] ? rm1 : 0.0);
#pragma unroll
                    for (int d = 0; d < NSC; ++d) bv[d] = rb1[d] * pm;
                    if (ks + 1 < nsteps) fetch1(ks + 1);
                    mma(av, bv);
                }
            } else {
                for (int ks = 0; ks < nsteps; ++ks) {
                    fetch(ks, 0);
                    step(0);
                }
            }
        } else {
            for (int ks = 0; ks < nsteps; ++ks) {
                int k, kcol;
                bool kin, cin;
                row_of(ks, k, kin, kcol, cin);
                double av[NT];
#pragma unroll
                for (int mt = 0; mt < NT; ++mt) av[mt] = (cin && min_[mt]) ? cbase[mt][kcol] : 0.0;
#pragma unroll
                for (int d = 0; d < NSC; ++d) {
                    if (a.set_mask & (1 << d)) {
                        const double bv = (kin && pin) ? abase[a.set_off[d] + (int64_t)k * a.ld] : 0.0;
#pragma unroll
                        for (int mt = 0; mt < NT; ++mt)
                            acc[mt][d] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[mt], bv, acc[mt][d], 0, 0, 0);
                    }
                }
            }
        }
        // item i of this lane is Phi[m = tile*16 + 4*i + kq][p]
#pragma unroll
        for (int mt = 0; mt < NT; ++mt) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int m = (t0 + mt) * 16 + kq + 4 * i;
                const double w = m < a.nmo ? a.occ[m] : 0.0;
                const double f0 = acc[mt][0][i];
                rho_acc += w * (f0 * f0);
                rho_terms<1, NSC>(acc[mt], i, f0, w, a.out_mask, dsum);
                if (a.mo_norm) {
                    double q = f0 * f0;
                    q += __shfl_xor(q, 1);
                    q += __shfl_xor(q, 2);
                    q += __shfl_xor(q, 4);
                    q += __shfl_xor(q, 8);
                    if (n == 0 && m < a.nmo) atomicAdd(a.mo_norm + m, q);
                }
            }
        }
    };
    if constexpr (EXACT) {
        for (int t0 = t_beg; t0 < t_end; t0 += MG) {
            const int nt = t_end - t0 < MG ? t_end - t0 : MG;   // wave-uniform
            if (nt == MG) pass(t0, std::integral_constant<int, MG>{});
            else if (nt == 1) pass(t0, std::integral_constant<int, 1>{});
            else if constexpr (MG > 2) {
                if (nt == 2) pass(t0, std::integral_constant<int, 2>{});
                else if constexpr (MG > 3) pass(t0, std::integral_constant<int, 3>{});
            }
        }
    } else {
        const int ngroups = (nmt + MG - 1) / MG;
        for (int g = wg ? wave : 0; g < ngroups; g += wg ? nwave : 1) pass(g * MG, std::integral_constant<int, MG>{});
    }
    rho_acc += __shfl_xor(rho_acc, 16);
    rho_acc += __shfl_xor(rho_acc, 32);
#pragma unroll
    for (int d = 1; d < NSC; ++d) {
        if (a.out_mask & (1 << d)) {
            dsum[d] += __shfl_xor(dsum[d], 16);
            dsum[d] += __shfl_xor(dsum[d], 32);
        }
    }
    if (wg) {
        // the waves' sums over their MO groups, added in wave order
        __shared__ double red[8][16][NSC];
        if (kq == 0) {
            red[wave][n][0] = rho_acc;
#pragma unroll
            for (int d = 1; d < NSC; ++d) red[wave][n][d] = dsum[d];
        }
        __syncthreads();
        if (wave != 0) return;
        rho_acc = red[0][n][0];
#pragma unroll
        for (int d = 1; d < NSC; ++d) dsum[d] = red[0][n][d];
        for (int w = 1; w < nwave; ++w) {
            rho_acc += red[w][n][0];
#pragma unroll
            for (int d = 1; d < NSC; ++d) dsum[d] += red[w][n][d];
        }
    }
    if (kq == 0 && pin) {
        a.rho[q] = rho_acc;
        for (int r = 0; r < a.nout; ++r) {
            const int c = a.out_code[r];
            double v = 0.0;
#pragma unroll
            for (int d = 0; d < NSC; ++d)
                if (d == c) v = d == 0 ? 2.0 * rho_acc : dsum[d];
            a.delta[(int64_t)r * a.ld_delta + q] = v;
        }
    }
}

template <int MG, bool FULL>
int launch_rho_full(const GemmArgs& a, int nsc, hipStream_t st)
{
    // wave_groups: one wave per MO group up to 4 groups; for 5 to 8 groups
    // 8 waves where passes run their exact tile count (four-tile groups, at
    // most 16 products per K step: the tiles dealt evenly, so each SIMD
    // holds two waves of about equal work), else one wave per group; beyond
    // 8 groups 4 waves
    const int nmt = (a.nmo + 15) / 16, ng = (nmt + MG - 1) / MG;
    const int w8 = MG == 4 && nsc * MG <= 16 ? 8 : ng;
    dim3 blk(a.wave_groups ? 64 * (ng <= 4 ? ng : ng <= 8 ? w8 : 4) : 256);
    dim3 grd((unsigned)(a.wave_groups ? (a.np + 15) / 16 : (a.np + 63) / 64));
    const bool lng = a.screened && a.scr.nchunk > 128;
#define OKMI_RHO_LAUNCH(NS)                                                                   \
    do {                                                                                      \
        if (lng) hipLaunchKernelGGL((k_rho_f64<MG, NS, FULL, true>), grd, blk, 0, st, a);     \
        else hipLaunchKernelGGL((k_rho_f64<MG, NS, FULL, false>), grd, blk, 0, st, a);        \
        return 0;                                                                             \
    } while (0)
    switch (nsc) {
    case 1: OKMI_RHO_LAUNCH(1);
    case 4: OKMI_RHO_LAUNCH(4);
    case 7:
        if constexpr (MG <= 3) OKMI_RHO_LAUNCH(7);
        break;
    default:
        if constexpr (MG <= 2) OKMI_RHO_LAUNCH(10);
        break;
    }
#undef OKMI_RHO_LAUNCH
    okmi_set_error("rho_f64: unsupported code range %d for MO group %d", nsc, MG);
    return 1;
}

// All codes 0..nsc-1 present and at least 4 of them: the branch-free
// instance (measured faster there; with one set the masked loop is faster).
template <int MG>
int launch_rho_nsc(const GemmArgs& a, int nsc, hipStream_t st)
{
    if (nsc >= 4 && a.set_mask == (1 << nsc) - 1) return launch_rho_full<MG, true>(a, nsc, st);
    return launch_rho_full<MG, false>(a, nsc, st);
}

inline int launch_rho(const GemmArgs& args, int nsc, hipStream_t st)
{
    const int nmt = (args.nmo + 15) / 16, mg = mo_group(nsc, nmt);
    GemmArgs a = args;
    a.wave_groups = (nmt + mg - 1) / mg >= 2;
    switch (mg) {
    case 1: return launch_rho_nsc<1>(a, nsc, st);
    case 2: return launch_rho_nsc<2>(a, nsc, st);
    case 3: return launch_rho_nsc<3>(a, nsc, st);
    default: return launch_rho_nsc<4>(a, nsc, st);
    }
}
}  // namespace okmi_gemm

// one translation unit per kernel family (compiled in parallel)
int okmi_gemm_f64_mode0(const okmi_gemm::GemmArgs& a, int nsets, hipStream_t st);
int okmi_gemm_f64_mode1(const okmi_gemm::GemmArgs& a, int nsc, hipStream_t st);
