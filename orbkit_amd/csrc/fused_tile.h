// Tile body of the fused density kernels: AO evaluation -> LDS -> MFMA
// contraction with the MO coefficients -> density terms, in one launch.  MOs
// never leave registers.  Shared by the single-molecule kernels (split and
// FP32 precisions in rho_fused.hip, FP64 in rho_fused_f64.hip) and the batch
// kernel.
//
// A workgroup of four waves owns a tile of 64 grid points: a 4x4x4 brick of a
// regular grid (compact in space) or 64 consecutive points of a vector grid.
// The AO rows are processed in chunks of 32 (one K step of the bf16 MFMA).
// For each chunk the four waves take the chunk's shells one at a time from a
// shared counter, costliest first (lane = point), and evaluate them in FP64
// (coordinates, exponentials, polynomials, the Cartesian-to-spherical map),
// storing each value in LDS [row][point]: rounded once to FP32 for the
// reduced precisions, in FP64 (with the full-accuracy exponential) for the
// FP64 contraction.  Reduced precisions use one buffer and three workgroups
// per CU for rho / rho + grad rho with l <= 3, with the chunk's C columns
// shared through LDS, else double buffering with one barrier per chunk; FP64
// uses one 64 KB buffer and two workgroups per CU (see FusedPlan).  Each wave
// then multiplies the rows of its 16-point column with the matching 32
// columns of C:
//   F64     v_mfma_f64_16x16x4_f64, C and AO in FP64, FP64 accumulation
//   F32     v_mfma_f32_16x16x4_f32, C and AO in FP32
//   BF16X3  v_mfma_f32_16x16x32_bf16 on (hi, lo) splits, 3 products
//   BF16X6  (hi, mid, lo) splits, 6 products (terms below 2^-24 dropped)
// accumulating in FP32 except F64.  The epilogue forms the density terms in
// FP64.
//
// Screening: a shell whose screening radius (every value and derivative
// below 1e-40 beyond it) does not reach the tile's bounding box is not
// evaluated (its rows are zero); a chunk with no such shell is skipped,
// contraction included.
//
// Derivative codes are handled in passes of at most 4 AO sets each:
//   PASS 0  {0}          rho
//   PASS 1  {0,x,y,z}    rho, first derivatives, and the f_a f_b halves of
//                        every second derivative (orbkit/core.py:287-302)
//   PASS 2  {0,xx,yy,zz} f0 f_aa halves of the diagonal second derivatives
//   PASS 3  {0,xy,xz,yz} f0 f_ab halves of the mixed ones
// Each term is written to its own FP64 buffer (no read-modify-write); a
// gather kernel then forms orbkit's output rows.
#pragma once
#include "okmi_internal.h"
#include "ao_device.h"
#include "split_mfma.h"

#include <type_traits>

namespace okmi_fused {

using namespace okmi_split;

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int KC = 32;        // AO rows per chunk
constexpr int PTS = 64;       // points per wave

// LDS AO tile of a precision: element type and row stride.  FP32 rows are
// padded to 65 floats (conflict-free 16-lane column reads); FP64 rows are
// 64 doubles, whose reads by the FP64 MFMA operand layout (4 rows x 16
// points) are not the limit next to the 64-cycle FP64 MFMAs they feed.
template <int PREC> struct Tile {
    using T = float;
    using Acc = f4;
    static constexpr int LD = 65;
};
template <> struct Tile<OKMI_PREC_F64> {
    using T = double;
    using Acc = d4;
    static constexpr int LD = 64;
};
template <typename T> constexpr int tile_ld() { return std::is_same<T, double>::value ? 64 : 65; }

struct FusedArgs {
    const ShellRec* shells;   // in work order: per chunk its shells, costliest first
                              // (prim0: first primitive group)
    const PrimGroup* groups;  // primitives, groups of 4 per shell
    const double* T;          // general transform blocks (kind 0 shells)
    const int2* chunks;       // per chunk: first shell, shell count | (used rows << 16)
    GridDesc grid;
    int64_t p0, np;
    int nao, nchunk;
    const void* C;            // F32: [mpad][kpad] float; BF16: [plane][mpad][kpad] bf16
    int64_t kpad, mpad;
    const double* occ;        // [mpad], zero padded
    double* rho;              // PASS 0/1
    double* term;             // [10][ld_term]: code c term (first derivative or f0 f_c half)
    double* term2;            // [10][ld_term]: f_a f_b halves (codes 4..9)
    int64_t ld_term;
    int need_mask;            // codes whose terms are wanted
    int brick;                // regular grid tiled in bricks of 64 points (p0 = 0, all points)
    int tlx, tly, tlz;        // brick extents 2^tlx x 2^tly x 2^tlz (tlx + tly + tlz = 6)
    int mt_live;              // MO tiles of 16 holding a real MO (0: all MT of the instance)
    int64_t nbx, nby, nbz;    // bricks per axis
    double* norm_part;        // FP64 kernel: per tile sum_p Phi_m(p)^2, [tile][mpad] (null: not wanted)
};

template <int PASS> struct PassCodes;
template <> struct PassCodes<0> { static constexpr int n = 1; static constexpr int c[4] = {0, 0, 0, 0}; };
template <> struct PassCodes<1> { static constexpr int n = 4; static constexpr int c[4] = {0, 1, 2, 3}; };
template <> struct PassCodes<2> { static constexpr int n = 4; static constexpr int c[4] = {0, 4, 5, 6}; };
template <> struct PassCodes<3> { static constexpr int n = 4; static constexpr int c[4] = {0, 7, 8, 9}; };

// brick coordinates of a tile (32-bit arithmetic: the host keeps the brick
// count below 2^32)
__device__ __forceinline__ void brick_of(const FusedArgs& a, int64_t tile, int64_t& bx, int64_t& by, int64_t& bz)
{
    unsigned t = (unsigned)tile;
    const unsigned nbz = (unsigned)a.nbz, nby = (unsigned)a.nby;
    bz = t % nbz;
    t /= nbz;
    by = t % nby;
    bx = t / nby;
}

// point of lane `ln` of this block's tile: index relative to p0, validity,
// and the (clamped) coordinates
__device__ __forceinline__ void tile_point(const FusedArgs& a, int64_t tile, int ln, int64_t& p, bool& act, double& px,
                                           double& py, double& pz)
{
    if (a.brick) {
        int64_t bx, by, bz;
        brick_of(a, tile, bx, by, bz);
        int64_t ix = (bx << a.tlx) + (ln >> (a.tly + a.tlz));
        int64_t iy = (by << a.tly) + ((ln >> a.tlz) & ((1 << a.tly) - 1));
        int64_t iz = (bz << a.tlz) + (ln & ((1 << a.tlz) - 1));
        act = ix < a.grid.nx && iy < a.grid.ny && iz < a.grid.nz;
        ix = ix < a.grid.nx ? ix : a.grid.nx - 1;
        iy = iy < a.grid.ny ? iy : a.grid.ny - 1;
        iz = iz < a.grid.nz ? iz : a.grid.nz - 1;
        p = (ix * a.grid.ny + iy) * a.grid.nz + iz;
        px = a.grid.x[ix]; py = a.grid.y[iy]; pz = a.grid.z[iz];
    } else {
        const int64_t pl = tile * PTS + ln;
        act = pl < a.np;
        p = act ? pl : a.np - 1;
        grid_point(a.grid, a.p0 + p, px, py, pz);
    }
}

// value known to be equal in all lanes -> scalar register, so that the
// compiler treats everything derived from it as wave-uniform (scalar loads
// of the shell tables, scalar branches)
__device__ __forceinline__ double uniform(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readfirstlane((int)b);
    const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ double wave_min(double v)
{
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) v = fmin(v, __shfl_xor(v, o));
    return uniform(v);
}

__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) v = fmax(v, __shfl_xor(v, o));
    return uniform(v);
}

struct Box {
    double lo[3], hi[3];
    // squared distance from a point to the box (wave-uniform)
    __device__ __forceinline__ double dist2(const ShellRec& sh) const
    {
        const double c[3] = {sh.cx, sh.cy, sh.cz};
        double d2 = 0.0;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double d = fmax(fmax(lo[k] - c[k], c[k] - hi[k]), 0.0);
            d2 = fma(d, d, d2);
        }
        return d2;
    }
};

// output rows of one shell and AO set from its canonical Cartesian values
// into the LDS tile (element type T: rounded once to FP32, or FP64)
template <int L, typename E>
__device__ __forceinline__ void store_rows(const FusedArgs& a, const ShellRec& sh, const double* c, E* dst)
{
    constexpr int NC = ncart(L);
    constexpr int LD = tile_ld<E>();
    auto store = [&](int r, double v) { dst[r * LD] = (E)v; };
    if (sh.kind == SHELL_T_SPHERICAL) {
        sph_rows_to<L, 0>(c, store);
    } else if (sh.kind == SHELL_T_CARTESIAN) {
        cart_rows_to<L, 0>(c, store);
    } else {
        const OKMI_CONST double* T = cptr(a.T) + sh.t0;
        for (int r = 0; r < sh.nrow; ++r) {
            double v = 0.0;
#pragma unroll
            for (int k = 0; k < NC; ++k) v = fma(T[r * NC + k], c[k], v);
            store(r, v);
        }
    }
}

// radial sums of the first N primitives of a group:
// S0 = sum d e, S1 = sum -2 alpha d e, S2 = sum 4 alpha^2 d e, e = exp(-alpha r^2)
// (EXACT: the full-accuracy FP64 exponential, else the one accurate to
// 2.6e-9 for values that end in FP32)
template <int N, bool need1, bool need2, bool EXACT>
__device__ __forceinline__ void prim_sums(const PrimGroup& pg, double rr, double& S0, double& S1, double& S2)
{
    double de[N];
#pragma unroll
    for (int j = 0; j < N; ++j)
        de[j] = pg.d[j] * (EXACT ? gauss_f64(pg.alpha[j] * rr) : gauss_f64_fast(pg.alpha[j] * rr));
#pragma unroll
    for (int j = 0; j < N; ++j) {
        S0 += de[j];
        if constexpr (need1) S1 = fma(-2.0 * pg.alpha[j], de[j], S1);
        if constexpr (need2) S2 = fma(4.0 * pg.alpha[j] * pg.alpha[j], de[j], S2);
    }
}

template <int L, int PASS, typename T>
__device__ __forceinline__ void eval_shell_chunk(const FusedArgs& a, const ShellRec& sh, double d2, double px,
                                                 double py, double pz, int k0, T* lds, int lane)
{
    constexpr bool EXACT = std::is_same<T, double>::value;
    constexpr int LD = tile_ld<T>();
    constexpr int NS = PassCodes<PASS>::n;
    constexpr bool need1 = PASS == 1 || PASS == 2;
    constexpr bool need2 = PASS >= 2;
    const double X = px - sh.cx, Y = py - sh.cy, Z = pz - sh.cz;
    const double rr = X * X + Y * Y + Z * Z;
    double S0 = 0.0, S1 = 0.0, S2 = 0.0;
    // Primitives four at a time (one group: independent exponential chains,
    // one wait for its loads).  A shell's primitives are sorted by screening
    // radius, largest first, so those that reach the tile lead: only they
    // are evaluated, and a group that is not fully live ends the shell.
    for (int q = 0; q < sh.nprim; q += 4) {
        const PrimGroup pg = load_group(a.groups, sh.prim0 + (q >> 2));
        int nl = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) nl += d2 < pg.rcut2[j];   // wave-uniform
        switch (nl) {
        case 4: prim_sums<4, need1, need2, EXACT>(pg, rr, S0, S1, S2); break;
        case 3: prim_sums<3, need1, need2, EXACT>(pg, rr, S0, S1, S2); break;
        case 2: prim_sums<2, need1, need2, EXACT>(pg, rr, S0, S1, S2); break;
        case 1: prim_sums<1, need1, need2, EXACT>(pg, rr, S0, S1, S2); break;
        default: break;
        }
        if (nl < 4) break;
    }
    ShellPows<L, double> P;
    P.init(X, Y, Z);
    constexpr int NC = ncart(L);
    if constexpr (PASS == 1) {
        // value and x, y, z derivatives from shared monomials
        double m[NC], u[L > 0 ? ncart(L - 1) : 1], c[NC];
        first_mono<L, 0, double>(P, m);
        lower_mono<L, 0, double>(P, S0, u);
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            if (s == 0) {
#pragma unroll
                for (int k = 0; k < NC; ++k) c[k] = m[k] * S0;
            } else if (s == 1) {
                first_deriv<L, 0, 0, double>(m, u, X * S1, c);
            } else if (s == 2) {
                first_deriv<L, 1, 0, double>(m, u, Y * S1, c);
            } else {
                first_deriv<L, 2, 0, double>(m, u, Z * S1, c);
            }
            store_rows<L>(a, sh, c, lds + s * KC * LD + (sh.row0 - k0) * LD + lane);
        }
        return;
    }
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        double c[NC];
        switch (PassCodes<PASS>::c[s]) {
        case 0: cart_all<L, 0, double>(P, S0, S1, S2, c); break;
        case 1: cart_all<L, 1, double>(P, S0, S1, S2, c); break;
        case 2: cart_all<L, 2, double>(P, S0, S1, S2, c); break;
        case 3: cart_all<L, 3, double>(P, S0, S1, S2, c); break;
        case 4: cart_all<L, 4, double>(P, S0, S1, S2, c); break;
        case 5: cart_all<L, 5, double>(P, S0, S1, S2, c); break;
        case 6: cart_all<L, 6, double>(P, S0, S1, S2, c); break;
        case 7: cart_all<L, 7, double>(P, S0, S1, S2, c); break;
        case 8: cart_all<L, 8, double>(P, S0, S1, S2, c); break;
        default: cart_all<L, 9, double>(P, S0, S1, S2, c); break;
        }
        store_rows<L>(a, sh, c, lds + s * KC * LD + (sh.row0 - k0) * LD + lane);
    }
}

template <int PREC, int PASS>
struct FusedLds {
    static constexpr int BUF = PassCodes<PASS>::n * KC * Tile<PREC>::LD;   // elements per buffer
};

// Operand A (MO coefficients) of one chunk, in the lane layout of one wave.
template <int PREC, int MT>
struct COperand {
    Split<PREC> ca[MT];
    float cf[MT][KC / 4];
    double cd[PREC == OKMI_PREC_F64 ? MT : 1][KC / 4];
};

template <int PREC, int MT>
__device__ __forceinline__ void load_c(const FusedArgs& a, int k0, int g, int n16, COperand<PREC, MT>& c)
{
    if constexpr (PREC == OKMI_PREC_F64) {
        const double* C = (const double*)a.C;
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
            for (int ks = 0; ks < KC / 4; ++ks)
                c.cd[mt][ks] = C[(int64_t)(mt * 16 + n16) * a.kpad + k0 + 4 * ks + g];
    } else if constexpr (PREC == OKMI_PREC_F32) {
        // K step ks of lane group g takes chunk row 8 g + ks (the same in
        // contract), so a lane's 8 values are contiguous: two 16-byte loads
        const float* C = (const float*)a.C;
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
            const f4* src = (const f4*)(C + (int64_t)(mt * 16 + n16) * a.kpad + k0 + 8 * g);
            const f4 lo = src[0], hi = src[1];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                c.cf[mt][j] = lo[j];
                c.cf[mt][4 + j] = hi[j];
            }
        }
    } else {
        const __bf16* C = (const __bf16*)a.C;
        constexpr int NP = PREC == OKMI_PREC_BF16X6 ? 3 : 2;
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
            for (int q = 0; q < NP; ++q)
                c.ca[mt].p[q] = *(const bf8*)(C + ((int64_t)q * a.mpad + mt * 16 + n16) * a.kpad + k0 + 8 * g);
    }
}

// The split C operand of one chunk, shared by the four waves through LDS in
// the one-buffer schedule: [plane][MT * 16 rows] of 80-byte rows, each the
// chunk's 32 bf16 (four 16-byte slots) and a 16-byte pad slot, so that the
// contraction's 16-byte reads (lanes n16 = 0..15 on rows 20 dwords apart)
// fall in distinct banks.  The workgroup loads it cooperatively when the
// chunk's evaluation starts (one copy instead of one per wave, its latency
// hidden behind the evaluation) and writes it to LDS before the barrier.
template <int PREC, int MT> struct CStage {
    static constexpr bool on = PREC == OKMI_PREC_BF16X3 || PREC == OKMI_PREC_BF16X6;
    static constexpr int NP = PREC == OKMI_PREC_BF16X6 ? 3 : 2;
    static constexpr int data = on ? NP * MT * 16 * 4 : 0;    // 16-byte data slots
    static constexpr int slots = on ? NP * MT * 16 * 5 : 0;   // with the pad slots
    static constexpr int rounds = (data + 255) / 256;          // per thread (256 threads)
};

template <int PREC, int MT>
struct CRegs {
    bf8 v[CStage<PREC, MT>::rounds > 0 ? CStage<PREC, MT>::rounds : 1];
};

// Data slot d = 4 r + c (row r = plane q, MO m; k0 + 8c .. k0 + 8c + 7) for
// d = 256 t + thread.  The loads are unconditional (threads past the end
// reload the last slot) and use the global address space, so that they are
// counted by vmcnt alone and the evaluation never waits for them.
template <int PREC, int MT>
__device__ __forceinline__ void fetch_c(const FusedArgs& a, int k0, CRegs<PREC, MT>& cr)
{
    using S = CStage<PREC, MT>;
    typedef const __attribute__((address_space(1))) __bf16 gbf16;
    typedef const __attribute__((address_space(1))) bf8 gbf8;
    gbf16* C = (gbf16*)a.C;
#pragma unroll
    for (int t = 0; t < S::rounds; ++t) {
        const int d = min(t * 256 + (int)threadIdx.x, S::data - 1);
        const int r = d >> 2, c = d & 3;
        const int q = r / (MT * 16), m = r - q * (MT * 16);
        cr.v[t] = *(gbf8*)(C + ((int64_t)q * a.mpad + m) * a.kpad + k0 + 8 * c);
    }
}

template <int PREC, int MT>
__device__ __forceinline__ void put_c(const CRegs<PREC, MT>& cr, bf8* cs)
{
    using S = CStage<PREC, MT>;
#pragma unroll
    for (int t = 0; t < S::rounds; ++t) {
        const int d = t * 256 + (int)threadIdx.x;
        if (d < S::data) cs[(d >> 2) * 5 + (d & 3)] = cr.v[t];
    }
}

// the lane's operand A fragments (MO row 16 mt + n16, k = 8g .. 8g + 7)
template <int PREC, int MT>
__device__ __forceinline__ void read_c(const bf8* cs, int g, int n16, COperand<PREC, MT>& c)
{
    constexpr int NP = CStage<PREC, MT>::NP;
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int q = 0; q < NP; ++q) c.ca[mt].p[q] = cs[((q * MT + mt) * 16 + n16) * 5 + g];
}

// AO rows of chunk `ch` for the tile's 64 points (lane = point) into the LDS
// tile L.  The calling waves take the chunk's shells one at a time from the
// counter *next (costliest first), so that screening cannot leave one wave
// with most of the work; a wave that evaluates a shell sets *live = tag.
// Waves with `pad` set also clear the rows after the chunk's last shell.
template <int PASS, int LM, typename T>
__device__ __forceinline__ void eval_chunk(const FusedArgs& a, int ch, const Box& box, double px, double py, double pz,
                                           T* L, int* next, int* live, int tag, bool pad, int lane)
{
    constexpr int NS = PassCodes<PASS>::n;
    constexpr int LD = tile_ld<T>();
    const int k0 = ch * KC;
    const int2 cw = load_int2(a.chunks, ch);
    const int nsh = cw.y & 0xffff, used = cw.y >> 16;
    if (pad && used < KC) {
        for (int s = 0; s < NS; ++s)
            for (int r = used; r < KC; ++r) L[s * KC * LD + r * LD + lane] = T(0);
    }
    for (;;) {
        int e = 0;
        if (lane == 0) e = atomicAdd(next, 1);
        e = __builtin_amdgcn_readfirstlane(e);   // all lanes active again: lane 0's value
        if (e >= nsh) break;
        const ShellRec sh = load_shell(a.shells, cw.x + e);
        const double d2 = box.dist2(sh);
        if (d2 < sh.rcut2) {
            if (lane == 0) *live = tag;
            switch (sh.l) {
            case 0: eval_shell_chunk<0, PASS>(a, sh, d2, px, py, pz, k0, L, lane); break;
            case 1: eval_shell_chunk<1, PASS>(a, sh, d2, px, py, pz, k0, L, lane); break;
            case 2: eval_shell_chunk<2, PASS>(a, sh, d2, px, py, pz, k0, L, lane); break;
            case 3: eval_shell_chunk<3, PASS>(a, sh, d2, px, py, pz, k0, L, lane); break;
            default:
                if constexpr (LM >= 4) eval_shell_chunk<4, PASS>(a, sh, d2, px, py, pz, k0, L, lane);
                break;
            }
        } else {
            for (int s = 0; s < NS; ++s)
                for (int r = 0; r < sh.nrow; ++r) L[s * KC * LD + (sh.row0 - k0 + r) * LD + lane] = T(0);
        }
    }
}

// Contraction of one chunk (LDS tile L) for the 16-point column `col` of the
// tile: acc[mt][s] += C[16 mt .. 16 mt + 15][chunk] * AO_s[chunk][column].
// SKIP (the batch kernel): MO tiles from mtl on (mtl > 0) hold only
// padding, their MFMAs are skipped (their accumulators stay zero, and their
// occupations are zero).
template <int PREC, int PASS, int MT, bool SKIP = false>
__device__ __forceinline__ void contract(const typename Tile<PREC>::T* L, int col, int g, int n16,
                                         const COperand<PREC, MT>& c,
                                         typename Tile<PREC>::Acc (&acc)[MT][PassCodes<PASS>::n], int mtl = 0)
{
    const int nmt = SKIP && mtl > 0 && mtl < MT ? mtl : MT;   // wave-uniform
    constexpr int NS = PassCodes<PASS>::n;
    constexpr int LD = Tile<PREC>::LD;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const typename Tile<PREC>::T* src = L + s * KC * LD + col * 16 + n16;
        if constexpr (PREC == OKMI_PREC_F64) {
#pragma unroll
            for (int ks = 0; ks < KC / 4; ++ks) {
                const double b = src[(4 * ks + g) * LD];
#pragma unroll
                for (int mt = 0; mt < MT; ++mt)
                    if (mt < nmt) acc[mt][s] = __builtin_amdgcn_mfma_f64_16x16x4f64(c.cd[mt][ks], b, acc[mt][s], 0, 0, 0);
            }
        } else if constexpr (PREC == OKMI_PREC_F32) {
#pragma unroll
            for (int ks = 0; ks < KC / 4; ++ks) {
                const float b = src[(8 * g + ks) * LD];   // chunk row 8 g + ks, as load_c
#pragma unroll
                for (int mt = 0; mt < MT; ++mt)
                    if (mt < nmt) acc[mt][s] = __builtin_amdgcn_mfma_f32_16x16x4f32(c.cf[mt][ks], b, acc[mt][s], 0, 0, 0);
            }
        } else {
            float bx[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) bx[j] = src[(8 * g + j) * LD];
            Split<PREC> bs;
            split8<PREC>(bx, bs);
#pragma unroll
            for (int mt = 0; mt < MT; ++mt)
                if (mt < nmt) acc[mt][s] = mfma_split<PREC>(c.ca[mt], bs, acc[mt][s]);
        }
    }
}

// Density terms of the 16-point column `col` of the tile from its MO values:
// item i of the lane holds Phi[m][tile point 16 col + n16] with
// m = 16 mt + 4 g + i (FP32 accumulator layout) or m = 16 mt + 4 i + g
// (v_mfma_f64_16x16x4f64 leaves D[4 i + lane / 16][lane % 16] in item i).
template <int PREC, int PASS, int MT>
__device__ __forceinline__ void epilogue(const FusedArgs& a, int64_t tile, int col, int lane,
                                         const typename Tile<PREC>::Acc (&acc)[MT][PassCodes<PASS>::n])
{
    constexpr int NS = PassCodes<PASS>::n;
    const int g = lane >> 4, n16 = lane & 15;
    double r0 = 0.0, t[NS > 1 ? NS - 1 : 1], u[6];
#pragma unroll
    for (int s = 0; s < (NS > 1 ? NS - 1 : 1); ++s) t[s] = 0.0;
#pragma unroll
    for (int q = 0; q < 6; ++q) u[q] = 0.0;
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const double w = a.occ[mt * 16 + (PREC == OKMI_PREC_F64 ? 4 * i + g : 4 * g + i)];
            const double f0 = acc[mt][0][i];
            if constexpr (PASS <= 1) r0 += w * (f0 * f0);
            if constexpr (NS > 1) {
#pragma unroll
                for (int s = 1; s < NS; ++s) t[s - 1] += 2.0 * w * (f0 * (double)acc[mt][s][i]);
            }
            if constexpr (PASS == 1) {
                const double fx = acc[mt][1][i], fy = acc[mt][2][i], fz = acc[mt][3][i];
                u[0] += 2.0 * w * (fx * fx);
                u[1] += 2.0 * w * (fy * fy);
                u[2] += 2.0 * w * (fz * fz);
                u[3] += 2.0 * w * (fx * fy);
                u[4] += 2.0 * w * (fx * fz);
                u[5] += 2.0 * w * (fy * fz);
            }
        }
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

// (min, max) of the axis values of brick b along one axis (2^tl values,
// fewer at the grid's end), wave-uniform
__device__ __forceinline__ double2 axis_span(const double* ax, int64_t b, int tl, int64_t n)
{
    const int64_t i0 = b << tl, i1 = min(i0 + (int64_t(1) << tl), n);
    double lo = ax[i0], hi = lo;
    for (int64_t i = i0 + 1; i < i1; ++i) {
        lo = fmin(lo, ax[i]);
        hi = fmax(hi, ax[i]);
    }
    return double2{uniform(lo), uniform(hi)};
}

// Bounding box of the tile's points: for a brick of a regular grid from the
// (at most 2^tl) axis values it spans, otherwise by wave reductions over the
// lanes' points (px, py, pz: lane l holds point l of the tile).
__device__ __forceinline__ Box tile_box(const FusedArgs& a, int64_t tile, double px, double py, double pz)
{
    Box box;
    if (a.brick) {
        int64_t bx, by, bz;
        brick_of(a, tile, bx, by, bz);
        const double2 sx = axis_span(a.grid.x, bx, a.tlx, a.grid.nx);
        const double2 sy = axis_span(a.grid.y, by, a.tly, a.grid.ny);
        const double2 sz = axis_span(a.grid.z, bz, a.tlz, a.grid.nz);
        box.lo[0] = sx.x; box.hi[0] = sx.y;
        box.lo[1] = sy.x; box.hi[1] = sy.y;
        box.lo[2] = sz.x; box.hi[2] = sz.y;
        return box;
    }
    box.lo[0] = wave_min(px); box.hi[0] = wave_max(px);
    box.lo[1] = wave_min(py); box.hi[1] = wave_max(py);
    box.lo[2] = wave_min(pz); box.hi[2] = wave_max(pz);
    return box;
}

// Chunk liveness of a tile, computed once per tile before the chunk loop:
// bit c of mask[c / 64] is set when a shell of chunk c reaches the tile's
// box (the test eval_chunk applies shell by shell, so a chunk without the
// bit would evaluate nothing and skip its contraction).  The chunk loop
// visits only these chunks, which saves a dead chunk's shell claims, row
// clearing and barriers.  Word j is formed by wave j % 4 (lane l tests chunk
// 64 j + l) and written whole, so no clearing is needed; callers run this
// after a barrier that follows every read of the previous tile's mask and
// put a barrier before the first read.  Bases with more chunks than the
// mask holds have every chunk visited.
constexpr int LIVE_WORDS = 32;

__device__ __forceinline__ void live_chunks(const FusedArgs& a, const Box& box, int wave, int lane,
                                            unsigned long long* mask)
{
    const int nw = (a.nchunk + 63) / 64;
    if (nw > LIVE_WORDS) return;
    for (int j = wave; j < nw; j += 4) {
        const int c = 64 * j + lane;
        bool hit = false;
        if (c < a.nchunk) {
            const int2 cw = load_int2(a.chunks, c);
            const int nsh = cw.y & 0xffff;
            for (int e = 0; e < nsh && !hit; ++e) {
                const ShellRec sh = load_shell(a.shells, cw.x + e);
                hit = box.dist2(sh) < sh.rcut2;
            }
        }
        const unsigned long long m = __builtin_amdgcn_ballot_w64(hit);
        if (lane == 0) mask[j] = m;
    }
}

// first chunk c' >= c with its bit set (a.nchunk if none), wave-uniform
__device__ __forceinline__ int next_live(const FusedArgs& a, const unsigned long long* mask, int c)
{
    if (a.nchunk > 64 * LIVE_WORDS) return c;
    while (c < a.nchunk) {
        const unsigned long long w0 = mask[c >> 6];
        const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)w0);
        const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(w0 >> 32));
        const unsigned long long w = (((unsigned long long)hi << 32) | lo) >> (c & 63);
        if (w) return c + __builtin_ctzll(w);
        c = (c | 63) + 1;
    }
    return a.nchunk;
}

// Whether an instance uses one LDS buffer and three workgroups per CU: rho
// and rho + grad rho with l <= 3 and up to 48 MOs (32 in the batch kernel,
// which needs a few more registers) fit in 168 VGPRs without spilling (33 KB
// AO tile plus up to 11.5 KB of C stage) and run at 3 waves/SIMD; the others
// keep two buffers (66.5 KB, 2 waves/SIMD), where holding C through the
// evaluation hides its latency.  FP64 always has one buffer (64 KB for four
// AO sets) and two workgroups per CU.  FP64 and FP32 load their C operand
// into registers before the evaluation and hold it through it (hold_c).
template <int PREC, int PASS, int LM, int MT, bool BATCH> struct FusedPlan {
    static constexpr bool f64 = PREC == OKMI_PREC_F64;
    static constexpr bool single = f64 || (PASS <= 1 && LM <= 3 && MT <= (BATCH ? 2 : 3));
    static constexpr int buffers = single ? 1 : 2;
    static constexpr int occupancy = f64 ? 2 : single ? 3 : 2;
    static constexpr bool c_early = f64 ? MT <= 3 : PASS <= 1 ? MT <= 3 : MT <= 2;
    // one-buffer schedule, C in registers (FP64, FP32): loaded before the
    // evaluation and held through it, its latency hidden there (FP32 with 3
    // MO tiles and four AO sets would exceed the 168-VGPR budget)
    static constexpr bool hold_c = c_early && (f64 || (PREC == OKMI_PREC_F32 && (PASS == 0 || MT <= 2)));
};

// One tile of 64 points by four waves that all evaluate and all contract:
// every thread of the workgroup calls this with the same tile.  `lds` holds
// FusedPlan::buffers AO buffers, `cst` the shared C operand (one-buffer
// schedule, split precisions), `next` two work counters, `live` two chunk
// tags.  The first barrier orders it after any previous tile's LDS reads.
// LM: highest angular momentum of the basis (3 or 4).
template <int PREC, int PASS, int MT, int LM, bool BATCH>
__device__ __forceinline__ void rho_tile(const FusedArgs& a, int64_t tile, typename Tile<PREC>::T* lds, bf8* cst,
                                         int* next, int* live, unsigned long long* lmask)
{
    using P = FusedPlan<PREC, PASS, LM, MT, BATCH>;
    using Acc = typename Tile<PREC>::Acc;
    constexpr int NS = PassCodes<PASS>::n;
    constexpr int BUF = FusedLds<PREC, PASS>::BUF;
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int g = lane >> 4, n16 = lane & 15;
    if (threadIdx.x < 2) next[threadIdx.x] = 0;
    __syncthreads();
    int64_t pidx;
    bool pact;
    double px, py, pz;
    tile_point(a, tile, lane, pidx, pact, px, py, pz);   // lane l of every wave: point l of the tile
    const Box box = tile_box(a, tile, px, py, pz);
    live_chunks(a, box, wave, lane, lmask);
    __syncthreads();

    Acc acc[MT][NS];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int s = 0; s < NS; ++s) acc[mt][s] = Acc{0, 0, 0, 0};

    if constexpr (P::single) {
        // One buffer: evaluate (C fetched meanwhile), C to LDS, barrier,
        // contract, barrier.  The point coordinates are consumed here so
        // that the evaluation does not count their loads as outstanding
        // (which would make it wait for the C loads issued after them).
        asm volatile("" ::"v"(px), "v"(py), "v"(pz));
        for (int ch = next_live(a, lmask, 0); ch < a.nchunk; ch = next_live(a, lmask, ch + 1)) {
            CRegs<PREC, MT> cr;
            COperand<PREC, MT> c;
            if constexpr (CStage<PREC, MT>::on) fetch_c<PREC, MT>(a, ch * KC, cr);
            else if constexpr (P::hold_c) load_c<PREC, MT>(a, ch * KC, g, n16, c);
            eval_chunk<PASS, LM>(a, ch, box, px, py, pz, lds, &next[0], &live[0], ch, wave == 0, lane);
            // the staging area is free: every wave read it before the
            // previous barrier
            if constexpr (CStage<PREC, MT>::on) put_c<PREC, MT>(cr, cst);
            __syncthreads();
            // every wave finished taking from the counter before the barrier
            // and takes from it again after the next one
            if (threadIdx.x == 0) next[0] = 0;
            if (live[0] == ch) {
                if constexpr (CStage<PREC, MT>::on) read_c<PREC, MT>(cst, g, n16, c);
                else if constexpr (!P::hold_c) load_c<PREC, MT>(a, ch * KC, g, n16, c);
                contract<PREC, PASS, MT, BATCH>(lds, wave, g, n16, c, acc, a.mt_live);
            }
            __syncthreads();
        }
        epilogue<PREC, PASS, MT>(a, tile, wave, lane, acc);
        return;
    }
    // Two buffers.  The C operand is loaded before the AO evaluation (its
    // latency hidden there) unless holding it through the evaluation would
    // exceed the register budget: second-derivative passes and 4 MO tiles
    // load it after the barrier.
    constexpr bool c_early = P::c_early;
    int buf = 0;
    for (int ch = next_live(a, lmask, 0); ch < a.nchunk; ch = next_live(a, lmask, ch + 1)) {
        typename Tile<PREC>::T* L = lds + buf * BUF;
        COperand<PREC, MT> c;
        if constexpr (c_early) load_c<PREC, MT>(a, ch * KC, g, n16, c);
        // the other buffer's counter: every wave finished taking from it
        // before the previous barrier, the next chunk reads it after the next
        if (threadIdx.x == 0) next[buf ^ 1] = 0;
        eval_chunk<PASS, LM>(a, ch, box, px, py, pz, L, &next[buf], &live[buf], ch, wave == 0, lane);
        // All rows of buffer `buf` written; the other buffer is free because
        // every wave has passed the previous barrier after reading it.
        __syncthreads();
        // No shell of the chunk reaches the tile: every row is zero, skip
        // the contraction (same decision in all waves).  live[buf] is tagged
        // with the chunk index instead of being reset: it is written before
        // this barrier and next written (chunk ch + 2) after the following
        // one, and a stale value can only let a zero chunk through.
        if (live[buf] == ch) {
            if constexpr (!c_early) load_c<PREC, MT>(a, ch * KC, g, n16, c);
            contract<PREC, PASS, MT, BATCH>(L, wave, g, n16, c, acc, a.mt_live);
        }
        buf ^= 1;
    }
    epilogue<PREC, PASS, MT>(a, tile, wave, lane, acc);
}

// Single-molecule fused kernel: one workgroup per tile of 64 points.
template <int PREC, int PASS, int MT, int LM>
__global__ __launch_bounds__(256, (FusedPlan<PREC, PASS, LM, MT, false>::occupancy)) void k_rho_fused(FusedArgs a)
{
    using P = FusedPlan<PREC, PASS, LM, MT, false>;
    __shared__ typename Tile<PREC>::T lds[P::buffers * FusedLds<PREC, PASS>::BUF];
    __shared__ bf8 cst[P::single && CStage<PREC, MT>::on ? CStage<PREC, MT>::slots : 1];
    __shared__ int next[2];   // per buffer: next work entry to evaluate
    __shared__ int live[2];   // per buffer: index of the last chunk with a live shell
    __shared__ unsigned long long lmask[LIVE_WORDS];   // chunks with a shell in reach of the tile
    rho_tile<PREC, PASS, MT, LM, false>(a, blockIdx.x, lds, cst, next, live, lmask);
}

template <int PREC, int PASS, int LM>
int launch_mt(const FusedArgs& a, int mt, hipStream_t st)
{
    const int64_t ntile = a.brick ? a.nbx * a.nby * a.nbz : (a.np + PTS - 1) / PTS;
    if (ntile >= ((int64_t)1 << 31)) {   // one workgroup per tile; brick indices are 32-bit
        okmi_set_error("rho_fused: %lld tiles of 64 points (at most 2^31 - 1 per call)", (long long)ntile);
        return 1;
    }
    dim3 blk(256), grd((unsigned)ntile);
    OkmiProfScope prof(PREC == OKMI_PREC_F64 ? "rho_fused_f64" : "rho_fused", st);
    switch (mt) {
    case 1: hipLaunchKernelGGL((k_rho_fused<PREC, PASS, 1, LM>), grd, blk, 0, st, a); return 0;
    case 2: hipLaunchKernelGGL((k_rho_fused<PREC, PASS, 2, LM>), grd, blk, 0, st, a); return 0;
    case 3: hipLaunchKernelGGL((k_rho_fused<PREC, PASS, 3, LM>), grd, blk, 0, st, a); return 0;
    case 4: hipLaunchKernelGGL((k_rho_fused<PREC, PASS, 4, LM>), grd, blk, 0, st, a); return 0;
    default: break;
    }
    okmi_set_error("rho_fused: %d MO tiles (up to 4 supported)", mt);
    return 1;
}

// lm: highest angular momentum of the basis (l <= 3 selects the l <= 3
// instances; only rho and rho + grad rho have them, the second-derivative
// passes use the general ones)
template <int PREC>
int launch_pass(const FusedArgs& a, int pass, int mt, int lm, hipStream_t st)
{
    const bool l3 = lm <= 3;
    switch (pass) {
    case 0: return l3 ? launch_mt<PREC, 0, 3>(a, mt, st) : launch_mt<PREC, 0, 4>(a, mt, st);
    case 1: return l3 ? launch_mt<PREC, 1, 3>(a, mt, st) : launch_mt<PREC, 1, 4>(a, mt, st);
    case 2: return launch_mt<PREC, 2, 4>(a, mt, st);
    default: return launch_mt<PREC, 3, 4>(a, mt, st);
    }
}

// FP64 kernel (rho_fused_f64.hip, compiled with FP64 MFMAs in the VGPR form);
// mpad: MOs padded to a multiple of 4, a.mpad the same
int launch_pass_f64(const FusedArgs& a, int pass, int mpad, int lm, hipStream_t st);
// mo_norm[m] += sum_t part[t][m] for m < nmo (the FP64 kernel's per-tile norms)
int launch_norm_reduce(const double* part, int64_t ntile, int mpad, int nmo, double* mo_norm, hipStream_t st);

}  // namespace okmi_fused
