// Fused density kernel: AO evaluation -> LDS -> MFMA contraction with the MO
// coefficients -> density terms, in one launch.  MOs never leave registers.
//
// A workgroup of four waves owns a tile of 64 grid points: a 4x4x4 brick of a
// regular grid (compact in space) or 64 consecutive points of a vector grid.
// The AO rows are processed in chunks of 32 (one K step of the bf16 MFMA).
// For each chunk the four waves take the chunk's shells one at a time from a
// shared counter, costliest first (lane = point), and evaluate them in FP64
// (coordinates, exponentials, polynomials, the Cartesian-to-spherical map),
// storing each value rounded once to FP32 in LDS [row][point] (one buffer
// and three workgroups per CU for rho / rho + grad rho with l <= 3, with the
// chunk's C columns shared through LDS; else double buffered with one
// barrier per chunk; see FusedPlan).  Each wave then multiplies the rows of
// its 16-point column with the matching 32 columns of C:
//   F32     v_mfma_f32_16x16x4_f32, C and AO in FP32
//   BF16X3  v_mfma_f32_16x16x32_bf16 on (hi, lo) splits, 3 products
//   BF16X6  (hi, mid, lo) splits, 6 products (terms below 2^-24 dropped)
// accumulating in FP32.  The epilogue forms the density terms in FP64.
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
#include "okmi_internal.h"
#include "ao_device.h"
#include "split_mfma.h"

namespace {

using namespace okmi_split;

constexpr int KC = 32;        // AO rows per chunk
constexpr int LDS_LD = 65;    // padded row stride (floats) of the LDS AO tile
constexpr int PTS = 64;       // points per wave

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
    int tl_pad;
    int64_t nbx, nby, nbz;    // bricks per axis
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

// output rows of one shell and AO set from its canonical Cartesian values,
// rounded once to FP32 into the LDS tile (row stride LDS_LD)
template <int L>
__device__ __forceinline__ void store_rows(const FusedArgs& a, const ShellRec& sh, const double* c, float* dst)
{
    constexpr int NC = ncart(L);
    auto store = [&](int r, double v) { dst[r * LDS_LD] = (float)v; };
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
template <int N, bool need1, bool need2>
__device__ __forceinline__ void prim_sums(const PrimGroup& pg, double rr, double& S0, double& S1, double& S2)
{
    double de[N];
#pragma unroll
    for (int j = 0; j < N; ++j) de[j] = pg.d[j] * gauss_f64_fast(pg.alpha[j] * rr);
#pragma unroll
    for (int j = 0; j < N; ++j) {
        S0 += de[j];
        if constexpr (need1) S1 = fma(-2.0 * pg.alpha[j], de[j], S1);
        if constexpr (need2) S2 = fma(4.0 * pg.alpha[j] * pg.alpha[j], de[j], S2);
    }
}

template <int L, int PASS>
__device__ __forceinline__ void eval_shell_chunk(const FusedArgs& a, const ShellRec& sh, double d2, double px,
                                                 double py, double pz, int k0, float* lds, int lane)
{
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
        case 4: prim_sums<4, need1, need2>(pg, rr, S0, S1, S2); break;
        case 3: prim_sums<3, need1, need2>(pg, rr, S0, S1, S2); break;
        case 2: prim_sums<2, need1, need2>(pg, rr, S0, S1, S2); break;
        case 1: prim_sums<1, need1, need2>(pg, rr, S0, S1, S2); break;
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
            store_rows<L>(a, sh, c, lds + s * KC * LDS_LD + (sh.row0 - k0) * LDS_LD + lane);
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
        store_rows<L>(a, sh, c, lds + s * KC * LDS_LD + (sh.row0 - k0) * LDS_LD + lane);
    }
}

template <int PASS>
struct FusedLds {
    static constexpr int BUF = PassCodes<PASS>::n * KC * LDS_LD;   // floats per buffer
};

// Operand A (MO coefficients) of one chunk, in the lane layout of one wave.
template <int PREC, int MT>
struct COperand {
    Split<PREC> ca[MT];
    float cf[MT][KC / 4];
};

template <int PREC, int MT>
__device__ __forceinline__ void load_c(const FusedArgs& a, int k0, int g, int n16, COperand<PREC, MT>& c)
{
    if constexpr (PREC == OKMI_PREC_F32) {
        const float* C = (const float*)a.C;
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
            for (int ks = 0; ks < KC / 4; ++ks)
                c.cf[mt][ks] = C[(int64_t)(mt * 16 + n16) * a.kpad + k0 + 4 * ks + g];
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
    static constexpr bool on = PREC != OKMI_PREC_F32;
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
template <int PASS, int LM>
__device__ __forceinline__ void eval_chunk(const FusedArgs& a, int ch, const Box& box, double px, double py, double pz,
                                           float* L, int* next, int* live, int tag, bool pad, int lane)
{
    constexpr int NS = PassCodes<PASS>::n;
    const int k0 = ch * KC;
    const int2 cw = load_int2(a.chunks, ch);
    const int nsh = cw.y & 0xffff, used = cw.y >> 16;
    if (pad && used < KC) {
        for (int s = 0; s < NS; ++s)
            for (int r = used; r < KC; ++r) L[s * KC * LDS_LD + r * LDS_LD + lane] = 0.f;
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
                for (int r = 0; r < sh.nrow; ++r) L[s * KC * LDS_LD + (sh.row0 - k0 + r) * LDS_LD + lane] = 0.f;
        }
    }
}

// Contraction of one chunk (LDS tile L) for the 16-point column `col` of the
// tile: acc[mt][s] += C[16 mt .. 16 mt + 15][chunk] * AO_s[chunk][column].
template <int PREC, int PASS, int MT>
__device__ __forceinline__ void contract(const float* L, int col, int g, int n16, const COperand<PREC, MT>& c,
                                         f4 (&acc)[MT][PassCodes<PASS>::n])
{
    constexpr int NS = PassCodes<PASS>::n;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const float* src = L + s * KC * LDS_LD + col * 16 + n16;
        if constexpr (PREC == OKMI_PREC_F32) {
#pragma unroll
            for (int ks = 0; ks < KC / 4; ++ks) {
                const float b = src[(4 * ks + g) * LDS_LD];
#pragma unroll
                for (int mt = 0; mt < MT; ++mt)
                    acc[mt][s] = __builtin_amdgcn_mfma_f32_16x16x4f32(c.cf[mt][ks], b, acc[mt][s], 0, 0, 0);
            }
        } else {
            float bx[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) bx[j] = src[(8 * g + j) * LDS_LD];
            Split<PREC> bs;
            split8<PREC>(bx, bs);
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) acc[mt][s] = mfma_split<PREC>(c.ca[mt], bs, acc[mt][s]);
        }
    }
}

// Density terms of the 16-point column `col` of the tile from its MO values:
// lane holds Phi[m = 16 mt + 4 g + i][tile point 16 col + n16].
template <int PREC, int PASS, int MT>
__device__ __forceinline__ void epilogue(const FusedArgs& a, int64_t tile, int col, int lane,
                                         const f4 (&acc)[MT][PassCodes<PASS>::n])
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
            const double w = a.occ[mt * 16 + 4 * g + i];
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

// Whether an instance uses one LDS buffer and three workgroups per CU: rho
// and rho + grad rho with l <= 3 and up to 48 MOs (32 in the batch kernel,
// which needs a few more registers) fit in 168 VGPRs without spilling (33 KB
// AO tile plus up to 11.5 KB of C stage) and run at 3 waves/SIMD; the others
// keep two buffers (66.5 KB, 2 waves/SIMD), where holding C through the
// evaluation hides its latency.
template <int PASS, int LM, int MT, bool BATCH> struct FusedPlan {
    static constexpr bool single = PASS <= 1 && LM <= 3 && MT <= (BATCH ? 2 : 3);
    static constexpr int buffers = single ? 1 : 2;
    static constexpr int occupancy = single ? 3 : 2;
};

// One tile of 64 points by four waves that all evaluate and all contract:
// every thread of the workgroup calls this with the same tile.  `lds` holds
// FusedPlan::buffers AO buffers, `cst` the shared C operand (one-buffer
// schedule, split precisions), `next` two work counters, `live` two chunk
// tags.  The first barrier orders it after any previous tile's LDS reads.
// LM: highest angular momentum of the basis (3 or 4).
template <int PREC, int PASS, int MT, int LM, bool BATCH>
__device__ __forceinline__ void rho_tile(const FusedArgs& a, int64_t tile, float* lds, bf8* cst, int* next, int* live)
{
    constexpr int NS = PassCodes<PASS>::n;
    constexpr int BUF = FusedLds<PASS>::BUF;
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int g = lane >> 4, n16 = lane & 15;
    if (threadIdx.x < 2) next[threadIdx.x] = 0;
    __syncthreads();
    int64_t pidx;
    bool pact;
    double px, py, pz;
    tile_point(a, tile, lane, pidx, pact, px, py, pz);   // lane l of every wave: point l of the tile
    const Box box = tile_box(a, tile, px, py, pz);

    f4 acc[MT][NS];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int s = 0; s < NS; ++s) acc[mt][s] = f4{0.f, 0.f, 0.f, 0.f};

    if constexpr (FusedPlan<PASS, LM, MT, BATCH>::single) {
        // One buffer: evaluate (C fetched meanwhile), C to LDS, barrier,
        // contract, barrier.  The point coordinates are consumed here so
        // that the evaluation does not count their loads as outstanding
        // (which would make it wait for the C loads issued after them).
        asm volatile("" ::"v"(px), "v"(py), "v"(pz));
        for (int ch = 0; ch < a.nchunk; ++ch) {
            CRegs<PREC, MT> cr;
            if constexpr (CStage<PREC, MT>::on) fetch_c<PREC, MT>(a, ch * KC, cr);
            eval_chunk<PASS, LM>(a, ch, box, px, py, pz, lds, &next[0], &live[0], ch, wave == 0, lane);
            // the staging area is free: every wave read it before the
            // previous barrier
            if constexpr (CStage<PREC, MT>::on) put_c<PREC, MT>(cr, cst);
            __syncthreads();
            // every wave finished taking from the counter before the barrier
            // and takes from it again after the next one
            if (threadIdx.x == 0) next[0] = 0;
            if (live[0] == ch) {
                COperand<PREC, MT> c;
                if constexpr (CStage<PREC, MT>::on) read_c<PREC, MT>(cst, g, n16, c);
                else load_c<PREC, MT>(a, ch * KC, g, n16, c);
                contract<PREC, PASS, MT>(lds, wave, g, n16, c, acc);
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
    constexpr bool c_early = PASS <= 1 ? MT <= 3 : MT <= 2;
    int buf = 0;
    for (int ch = 0; ch < a.nchunk; ++ch) {
        float* L = lds + buf * BUF;
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
            contract<PREC, PASS, MT>(L, wave, g, n16, c, acc);
        }
        buf ^= 1;
    }
    epilogue<PREC, PASS, MT>(a, tile, wave, lane, acc);
}

template <int PREC, int PASS, int MT, int LM>
__global__ __launch_bounds__(256, (FusedPlan<PASS, LM, MT, false>::occupancy)) void k_rho_fused(FusedArgs a)
{
    using P = FusedPlan<PASS, LM, MT, false>;
    __shared__ float lds[P::buffers * FusedLds<PASS>::BUF];
    __shared__ bf8 cst[P::single && CStage<PREC, MT>::on ? CStage<PREC, MT>::slots : 1];
    __shared__ int next[2];   // per buffer: next work entry to evaluate
    __shared__ int live[2];   // per buffer: index of the last chunk with a live shell
    rho_tile<PREC, PASS, MT, LM, false>(a, blockIdx.x, lds, cst, next, live);
}

struct GatherCodes {
    int n;
    int c[16];
};

// out[r][p] = term[c] (+ term2[c] for second derivatives), or 2 rho for c = 0
__global__ void k_gather(const double* rho, const double* term, const double* term2, int64_t ld_term,
                         GatherCodes codes, double* out, int64_t ld_out, int64_t np)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    for (int r = 0; r < codes.n; ++r) {
        const int c = codes.c[r];
        double v = c == 0 ? 2.0 * rho[p] : term[c * ld_term + p];
        if (c >= 4) v += term2[c * ld_term + p];
        out[r * ld_out + p] = v;
    }
}

__global__ void k_pad_occ(const double* occ, int nmo, int mpad, double* out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < mpad) out[i] = i < nmo ? occ[i] : 0.0;
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
    OkmiProfScope prof("rho_fused", st);
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

}  // namespace

namespace {
size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

template <typename T>
int upload(T** d, const std::vector<T>& h, const char* what)
{
    hipError_t e = hipMalloc((void**)d, sizeof(T) * std::max<size_t>(1, h.size()));
    if (e != hipSuccess) return okmi_hip_fail(e, what);
    e = hipMemcpy(*d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) return okmi_hip_fail(e, what);
    return 0;
}

// Internal row layout of the fused kernels: shells in order, and a shell
// that would cross a 32-row chunk boundary starts the next chunk, so every
// chunk holds whole shells.  Returns the padded row count (-1 if a shell
// has more than 32 rows) and each shell's first internal row.
int fused_layout(const OkmiBasis* b, std::vector<int>* krow)
{
    int k = 0;
    if (krow) krow->assign(b->nshell, 0);
    for (int s = 0; s < b->nshell; ++s) {
        const int n = b->h_shells[s].nrow;
        if (n > KC) return -1;
        if (k % KC + n > KC) k = (k + KC - 1) / KC * KC;
        if (krow) (*krow)[s] = k;
        k += n;
    }
    return (k + KC - 1) / KC * KC;
}

size_t fused_fixed_bytes(const OkmiBasis* b, int nmo, int prec)
{
    const size_t mpad = ((nmo + 15) / 16) * 16, kpad = (size_t)std::max(KC, fused_layout(b, nullptr));
    const size_t elem = prec == OKMI_PREC_F32 ? 4 : prec == OKMI_PREC_BF16X3 ? 4 : 6;   // bytes per C entry
    return align256(elem * mpad * kpad) + align256(mpad * 8);
}

// Fused tables: per chunk of KC internal rows the shells inside it, each
// chunk's work list sorted by estimated evaluation cost (primitives,
// Cartesian components, output rows; costliest first, taken dynamically by
// the waves), copies of the shell and primitive tables with internal rows and
// the fused screening radii, and the internal-row -> output-row map.
int prepare_fused_tables(OkmiBasis* b)
{
    if (b->d_chunks) return 0;
    std::vector<int> krow;
    const int nkpad = fused_layout(b, &krow);
    if (nkpad < 0) {
        okmi_set_error("rho_fused: a contraction has more than %d output rows", KC);
        return 1;
    }
    const int nchunk = nkpad / KC;
    std::vector<int2> ch(nchunk);
    std::vector<int> work;
    int s = 0;
    for (int c = 0; c < nchunk; ++c) {
        std::vector<std::pair<int, int>> cost;
        int used = 0;
        while (s < b->nshell && krow[s] < (c + 1) * KC) {
            const ShellRec& sh = b->h_shells[s];
            cost.push_back({16 * sh.nprim + 8 * (ncart(sh.l) + sh.nrow), s});
            used = krow[s] + sh.nrow - c * KC;
            ++s;
        }
        std::sort(cost.begin(), cost.end(), [](auto& x, auto& y) { return x.first > y.first; });
        ch[c] = make_int2((int)work.size(), (int)cost.size() | (used << 16));
        for (auto& cs : cost) work.push_back(cs.second);
    }
    // shells in work-list order (entry j of the list is record j), each
    // with its primitives in groups of 4 (prim0: first group)
    std::vector<ShellRec> fs(work.size());
    std::vector<PrimGroup> groups;
    std::vector<int> colmap(nkpad, -1);
    for (size_t j = 0; j < work.size(); ++j) {
        const int i = work[j];
        fs[j] = b->h_shells[i];
        for (int r = 0; r < fs[j].nrow; ++r) colmap[krow[i] + r] = fs[j].row0 + r;
        fs[j].row0 = krow[i];
        fs[j].rcut2 = b->h_frcut2[i];
        fs[j].prim0 = (int)groups.size();
        const ShellRec& hs = b->h_shells[i];
        // primitives by screening radius, largest first (ties keep file order)
        std::vector<int> order(hs.nprim);
        for (int k = 0; k < hs.nprim; ++k) order[k] = hs.prim0 + k;
        std::stable_sort(order.begin(), order.end(),
                         [&](int x, int y) { return b->h_fprcut2[x] > b->h_fprcut2[y]; });
        for (int q = 0; q < hs.nprim; q += 4) {
            PrimGroup pg;
            for (int k = 0; k < 4; ++k) {
                const bool used = q + k < hs.nprim;
                const int pi = used ? order[q + k] : 0;
                pg.alpha[k] = used ? b->h_prims[pi].alpha : 0.0;
                pg.d[k] = used ? b->h_prims[pi].d : 0.0;
                pg.rcut2[k] = used ? b->h_fprcut2[pi] : -1.0;
            }
            groups.push_back(pg);
        }
    }
    std::vector<ShellRec> os(b->h_shells, b->h_shells + b->nshell);
    for (int i = 0; i < b->nshell; ++i) os[i].rcut2 = b->h_frcut2[i];
    std::vector<PrimRec> fp(b->h_prims, b->h_prims + b->nprim);
    for (int i = 0; i < b->nprim; ++i) fp[i].rcut2 = b->h_fprcut2[i];
    int rc = upload(&b->d_chunks, ch, "rho_fused: chunk map");
    if (!rc) rc = upload(&b->d_fshells, fs, "rho_fused: shells");
    if (!rc) rc = upload(&b->d_fprims, fp, "rho_fused: primitives");
    if (!rc) rc = upload(&b->d_fgroups, groups, "rho_fused: primitive groups");
    if (!rc) rc = upload(&b->d_oshells, os, "rho_fused: output-row shells");
    if (!rc) rc = upload(&b->d_colmap, colmap, "rho_fused: row map");
    if (rc) return rc;
    b->nchunk = nchunk;
    b->nkpad = nkpad;
    return 0;
}
}  // namespace

size_t okmi_fused_ws_bytes(const OkmiBasis* b, int nmo, int prec, int64_t np)
{
    if (nmo > 64) return okmi_big_ws_bytes(b, nmo, prec, 7, np);
    return fused_fixed_bytes(b, nmo, prec) + align256((size_t)20 * np * 8) + 256;
}

int okmi_launch_rho_fused(OkmiBasis* b, const GridDesc& g, int64_t p0, int64_t np, const double* d_c,
                          const double* d_occ, int nmo, int prec, const int* drv_out, int nout, double* rho,
                          double* delta, int64_t ld_delta, void* ws, size_t ws_bytes, hipStream_t st)
{
    if (np <= 0) return 0;
    const int mt = (nmo + 15) / 16;
    if (mt > 4)   // more MOs than the fused kernel keeps in registers
        return okmi_launch_rho_big(b, g, p0, np, d_c, d_occ, nmo, prec, drv_out, nout, rho, delta, ld_delta, ws,
                                   ws_bytes, st);
    if (prec != OKMI_PREC_F32 && prec != OKMI_PREC_BF16X3 && prec != OKMI_PREC_BF16X6) {
        okmi_set_error("rho_fused: unknown precision %d", prec);
        return 1;
    }
    if (nout > 16) {
        okmi_set_error("rho_fused: %d derivative outputs (up to 16)", nout);
        return 1;
    }
    int rc = prepare_fused_tables(b);
    if (rc) return rc;
    const size_t fixed = fused_fixed_bytes(b, nmo, prec);
    if (!ws || ws_bytes < fixed + 256 + 20 * 8 * 64) {
        okmi_set_error("rho_fused: workspace of %zu bytes is too small (need at least %zu)", ws_bytes,
                       fixed + 256 + 20 * 8 * 64);
        return 1;
    }
    int64_t chunk = (int64_t)((ws_bytes - fixed - 256) / (20 * 8));
    // a whole regular grid whose term buffers fit is tiled in 4x4x4 bricks
    const bool brick = g.regular && p0 == 0 && np == g.nx * g.ny * g.nz && chunk >= np;
    if (chunk >= np) chunk = np;
    else chunk &= ~(int64_t)63;
    const int64_t mpad = mt * 16, kpad = b->nkpad;
    char* w = (char*)(((uintptr_t)ws + 255) & ~(uintptr_t)255);
    void* cpl = w;
    const size_t elem = prec == OKMI_PREC_F32 ? 4 : prec == OKMI_PREC_BF16X3 ? 4 : 6;
    double* occp = (double*)(w + align256(elem * mpad * kpad));
    double* term = (double*)(w + fixed);
    double* term2 = term + 10 * chunk;

    hipLaunchKernelGGL(k_prep_c, dim3((unsigned)((mpad * kpad + 255) / 256)), dim3(256), 0, st, d_c, nmo, b->nao,
                       prec, mpad, kpad, (const int*)b->d_colmap, cpl);
    hipLaunchKernelGGL(k_pad_occ, dim3(1), dim3(256), 0, st, d_occ, nmo, (int)mpad, occp);

    int need = 0;
    GatherCodes gc{};
    gc.n = nout;
    for (int r = 0; r < nout; ++r) {
        if (drv_out[r] < 0 || drv_out[r] > 9) {
            okmi_set_error("The selection `drv=%d` is not valid!", drv_out[r]);
            return 1;
        }
        need |= 1 << drv_out[r];
        gc.c[r] = drv_out[r];
    }
    const bool p1 = (need & ~1) != 0, p2 = (need & (7 << 4)) != 0, p3 = (need & (7 << 7)) != 0;
    int lm = 0;   // highest angular momentum of the basis
    for (int i = 0; i < b->nshell; ++i) lm = std::max(lm, b->h_shells[i].l);

    FusedArgs a{};
    a.shells = b->d_fshells; a.groups = b->d_fgroups; a.T = b->d_T; a.chunks = b->d_chunks;
    a.grid = g; a.nao = b->nao; a.nchunk = b->nchunk;
    a.C = cpl; a.kpad = kpad; a.mpad = mpad; a.occ = occp;
    a.term = term; a.term2 = term2; a.ld_term = chunk; a.need_mask = need;
    a.brick = brick;
    a.tlx = a.tly = a.tlz = 2;
    a.nbx = (g.nx + 3) / 4; a.nby = (g.ny + 3) / 4; a.nbz = (g.nz + 3) / 4;
    for (int64_t c0 = 0; c0 < np; c0 += chunk) {
        const int64_t n = np - c0 < chunk ? np - c0 : chunk;
        a.p0 = p0 + c0;
        a.np = n;
        a.rho = rho + c0;
        int r;
        if (prec == OKMI_PREC_F32) {
            r = launch_pass<OKMI_PREC_F32>(a, p1 ? 1 : 0, mt, lm, st);
            if (!r && p2) r = launch_pass<OKMI_PREC_F32>(a, 2, mt, lm, st);
            if (!r && p3) r = launch_pass<OKMI_PREC_F32>(a, 3, mt, lm, st);
        } else if (prec == OKMI_PREC_BF16X3) {
            r = launch_pass<OKMI_PREC_BF16X3>(a, p1 ? 1 : 0, mt, lm, st);
            if (!r && p2) r = launch_pass<OKMI_PREC_BF16X3>(a, 2, mt, lm, st);
            if (!r && p3) r = launch_pass<OKMI_PREC_BF16X3>(a, 3, mt, lm, st);
        } else {
            r = launch_pass<OKMI_PREC_BF16X6>(a, p1 ? 1 : 0, mt, lm, st);
            if (!r && p2) r = launch_pass<OKMI_PREC_BF16X6>(a, 2, mt, lm, st);
            if (!r && p3) r = launch_pass<OKMI_PREC_BF16X6>(a, 3, mt, lm, st);
        }
        if (r) return r;
        if (nout > 0)
            hipLaunchKernelGGL(k_gather, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, rho + c0, term, term2,
                               chunk, gc, delta + c0, ld_delta, n);
    }
    return okmi_check_launch("rho_fused");
}

// ---------------------------------------------------------------------------
// Many molecules in one persistent launch.  Every molecule keeps its own
// tables (shells, primitives, transform, chunk map), MO coefficient planes,
// grid and outputs, described by one FusedArgs record; the tiles of all
// molecules form one index space (prefix sums tile_start, molecules in work
// order, costliest first by default).  A fixed grid of workgroups strides
// through it; each tile finds its molecule by a uniform binary search and
// runs the single-molecule tile body.  Launch sequence (graph-capturable:
// kernels only, no memsets, no allocation): C planes + padded occupations for
// every molecule, then the tile kernel, whose first-derivative terms are
// written straight into the caller's gradient rows.

namespace {

struct PrepItem {
    const int* colmap;    // [kpad] internal row -> C column (-1: zero)
    const double* C;      // [nmo][nao]
    const double* occ;    // [nmo]
    void* planes;         // [planes][mpad][kpad]
    double* occ_pad;      // [mpad]
    int nmo, nao;
    int64_t kpad;
};

struct BatchArgs {
    const FusedArgs* mols;       // per molecule, in work order
    const int64_t* tile_start;   // [nmol + 1]
    int nmol;
    int64_t ntile;
};

static_assert(sizeof(FusedArgs) % 8 == 0, "FusedArgs is loaded as 64-bit words");

// FusedArgs of molecule m through scalar (constant address space) loads
__device__ __forceinline__ FusedArgs load_fused_args(const FusedArgs* base, int m)
{
    constexpr int NW = sizeof(FusedArgs) / 8;
    struct Words { uint64_t w[NW]; } v;
    const OKMI_CONST uint64_t* src = (const OKMI_CONST uint64_t*)cptr(base + m);
#pragma unroll
    for (int i = 0; i < NW; ++i) v.w[i] = src[i];
    return __builtin_bit_cast(FusedArgs, v);
}

__device__ __forceinline__ int find_item(const int64_t* start, int n, int64_t t)
{
    const OKMI_CONST int64_t* s = cptr(start);
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (s[mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <int PREC, int PASS, int MT, int LM>
__global__ __launch_bounds__(256, (FusedPlan<PASS, LM, MT, true>::occupancy)) void k_rho_batch(BatchArgs b)
{
    using P = FusedPlan<PASS, LM, MT, true>;
    __shared__ float lds[P::buffers * FusedLds<PASS>::BUF];
    __shared__ bf8 cst[P::single && CStage<PREC, MT>::on ? CStage<PREC, MT>::slots : 1];
    __shared__ int next[2];
    __shared__ int live[2];
    for (int64_t t = blockIdx.x; t < b.ntile; t += gridDim.x) {
        const int m = find_item(b.tile_start, b.nmol, t);
        const FusedArgs a = load_fused_args(b.mols, m);
        rho_tile<PREC, PASS, MT, LM, true>(a, t - cptr(b.tile_start)[m], lds, cst, next, live);
    }
}

// element e of item m: C plane entries first (mpad * kpad), then occupations
__global__ void k_prep_batch(const PrepItem* items, const int64_t* start, int nitem, int64_t total, int prec,
                             int64_t mpad)
{
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        int lo = 0, hi = nitem - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (start[mid] <= e) lo = mid;
            else hi = mid - 1;
        }
        const PrepItem it = items[lo];
        const int64_t i = e - start[lo];
        const int64_t nc = mpad * it.kpad;
        if (i >= nc) {
            const int64_t j = i - nc;
            it.occ_pad[j] = j < it.nmo ? it.occ[j] : 0.0;
            continue;
        }
        const int64_t m = i / it.kpad, k = i - m * it.kpad;
        const int src = it.colmap[k];
        const double v = (m < it.nmo && src >= 0) ? it.C[m * it.nao + src] : 0.0;
        store_c_split(v, prec, i, nc, it.planes);
    }
}

template <int PREC, int PASS, int LM>
int launch_batch_mt(const BatchArgs& a, int mt, unsigned grid, hipStream_t st)
{
    switch (mt) {
    case 1: hipLaunchKernelGGL((k_rho_batch<PREC, PASS, 1, LM>), dim3(grid), dim3(256), 0, st, a); return 0;
    case 2: hipLaunchKernelGGL((k_rho_batch<PREC, PASS, 2, LM>), dim3(grid), dim3(256), 0, st, a); return 0;
    case 3: hipLaunchKernelGGL((k_rho_batch<PREC, PASS, 3, LM>), dim3(grid), dim3(256), 0, st, a); return 0;
    case 4: hipLaunchKernelGGL((k_rho_batch<PREC, PASS, 4, LM>), dim3(grid), dim3(256), 0, st, a); return 0;
    default: break;
    }
    okmi_set_error("batch: %d MO tiles (up to 4)", mt);
    return 1;
}

// lm: highest angular momentum over the batch's bases
template <int PREC>
int launch_batch(const BatchArgs& a, int pass, int mt, int lm, unsigned grid, hipStream_t st)
{
    if (lm <= 3)
        return pass == 0 ? launch_batch_mt<PREC, 0, 3>(a, mt, grid, st) : launch_batch_mt<PREC, 1, 3>(a, mt, grid, st);
    return pass == 0 ? launch_batch_mt<PREC, 0, 4>(a, mt, grid, st) : launch_batch_mt<PREC, 1, 4>(a, mt, grid, st);
}

}  // namespace

struct OkmiBatch {
    int nmol = 0, prec = 0, pass = 0, mt = 0, lm = 0;   // lm: highest angular momentum
    int64_t mpad = 0, ntile = 0, prep_total = 0;
    unsigned grid = 0;                    // persistent workgroups (0: one per tile)
    FusedArgs* d_mols = nullptr;
    int64_t* d_tile_start = nullptr;
    PrepItem* d_items = nullptr;
    int64_t* d_item_start = nullptr;
    void* d_planes = nullptr;
    double* d_occ = nullptr;
    std::vector<int> order;               // work order -> caller's molecule index
    std::vector<FusedArgs> h_mols;        // host copy of the descriptors (tile shape changes)
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t graph_stream = nullptr;
};

void okmi_batch_free(OkmiBatch* b)
{
    if (!b) return;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (void* p : {(void*)b->d_mols, (void*)b->d_tile_start, (void*)b->d_items, (void*)b->d_item_start,
                    b->d_planes, (void*)b->d_occ})
        if (p) (void)hipFree(p);
    delete b;
}

// brick extents 2^lx x 2^ly x 2^lz for the regular-grid molecules: tile
// counts, tile offsets and descriptors (re)uploaded; any captured graph is
// dropped (its kernels hold the old tile count)
static int batch_set_shape(OkmiBatch* B, int lx, int ly, int lz)
{
    if (lx < 0 || ly < 0 || lz < 0 || lx + ly + lz != 6) {
        okmi_set_error("batch: tile shape 2^%d x 2^%d x 2^%d is not 64 points", lx, ly, lz);
        return 1;
    }
    std::vector<int64_t> tstart(B->nmol + 1, 0);
    for (int j = 0; j < B->nmol; ++j) {
        FusedArgs& a = B->h_mols[j];
        int64_t tiles;
        if (a.brick) {
            a.tlx = lx; a.tly = ly; a.tlz = lz;
            a.nbx = (a.grid.nx + (1 << lx) - 1) >> lx;
            a.nby = (a.grid.ny + (1 << ly) - 1) >> ly;
            a.nbz = (a.grid.nz + (1 << lz) - 1) >> lz;
            tiles = a.nbx * a.nby * a.nbz;
            if (tiles >= ((int64_t)1 << 31)) {   // brick indices are 32-bit
                okmi_set_error("batch: molecule %d has %lld tiles (at most 2^31 - 1)", j, (long long)tiles);
                return 1;
            }
        } else {
            tiles = (a.np + PTS - 1) / PTS;
        }
        tstart[j + 1] = tstart[j] + tiles;
    }
    if (B->exec) {
        (void)hipGraphExecDestroy(B->exec);
        (void)hipGraphDestroy(B->graph);
        B->exec = nullptr;
        B->graph = nullptr;
    }
    if (B->d_mols) (void)hipFree(B->d_mols);
    if (B->d_tile_start) (void)hipFree(B->d_tile_start);
    B->d_mols = nullptr;
    B->d_tile_start = nullptr;
    int rc = upload(&B->d_mols, B->h_mols, "batch: descriptors");
    if (!rc) rc = upload(&B->d_tile_start, tstart, "batch: tile offsets");
    B->ntile = tstart[B->nmol];
    return rc;
}

int okmi_batch_set_shape_impl(OkmiBatch* B, int lx, int ly, int lz) { return batch_set_shape(B, lx, ly, lz); }

int okmi_batch_build(int nmol, OkmiBasis* const* bases, const GridDesc* grids, const int64_t* npts,
                     const double* const* d_c, const double* const* d_occ, const int* nmo, double* const* d_rho,
                     double* const* d_grad, int prec, int order_by_cost, OkmiBatch** out)
{
    if (prec != OKMI_PREC_F32 && prec != OKMI_PREC_BF16X3 && prec != OKMI_PREC_BF16X6) {
        okmi_set_error("batch: unknown precision %d", prec);
        return 1;
    }
    OkmiBatch* B = new OkmiBatch();
    B->nmol = nmol;
    B->prec = prec;
    int mt = 1;
    bool grad = false;
    std::vector<double> cost(nmol);
    for (int m = 0; m < nmol; ++m) {
        if (nmo[m] <= 0 || nmo[m] > 64) {
            okmi_set_error("batch: molecule %d has %d MOs (1..64 in a batch)", m, nmo[m]);
            okmi_batch_free(B);
            return 1;
        }
        if (npts[m] <= 0) {
            okmi_set_error("batch: molecule %d has an empty grid", m);
            okmi_batch_free(B);
            return 1;
        }
        int rc = prepare_fused_tables(bases[m]);
        if (rc) {
            okmi_batch_free(B);
            return rc;
        }
        mt = std::max(mt, (nmo[m] + 15) / 16);
        grad = grad || d_grad[m] != nullptr;
        for (int s = 0; s < bases[m]->nshell; ++s) B->lm = std::max(B->lm, bases[m]->h_shells[s].l);
        double prims = 0;
        for (int s = 0; s < bases[m]->nshell; ++s) prims += bases[m]->h_shells[s].nprim + bases[m]->h_shells[s].nrow;
        cost[m] = prims * (double)npts[m];
    }
    B->mt = mt;
    B->mpad = 16 * mt;
    B->pass = grad ? 1 : 0;
    B->order.resize(nmol);
    for (int m = 0; m < nmol; ++m) B->order[m] = m;
    if (order_by_cost)
        std::stable_sort(B->order.begin(), B->order.end(), [&](int x, int y) { return cost[x] > cost[y]; });
    const size_t elem = prec == OKMI_PREC_F32 ? 4 : prec == OKMI_PREC_BF16X3 ? 4 : 6;
    std::vector<size_t> poff(nmol);
    size_t pbytes = 0;
    for (int j = 0; j < nmol; ++j) {
        const int m = B->order[j];
        poff[j] = pbytes;
        pbytes += (elem * B->mpad * (int64_t)bases[m]->nkpad + 255) & ~(size_t)255;
    }
    hipError_t e = hipMalloc(&B->d_planes, std::max<size_t>(pbytes, 256));
    if (e == hipSuccess) e = hipMalloc((void**)&B->d_occ, sizeof(double) * B->mpad * nmol);
    if (e != hipSuccess) {
        okmi_batch_free(B);
        return okmi_hip_fail(e, "batch: coefficient planes");
    }
    std::vector<FusedArgs> args(nmol);
    std::vector<int64_t> istart(nmol + 1, 0);
    std::vector<PrepItem> items(nmol);
    for (int j = 0; j < nmol; ++j) {
        const int m = B->order[j];
        const OkmiBasis* bs = bases[m];
        const GridDesc& g = grids[m];
        FusedArgs a{};
        a.shells = bs->d_fshells; a.groups = bs->d_fgroups; a.T = bs->d_T;
        a.chunks = bs->d_chunks;
        a.grid = g; a.p0 = 0; a.np = npts[m];
        a.nao = bs->nao; a.nchunk = bs->nchunk;
        a.kpad = bs->nkpad; a.mpad = B->mpad;
        a.C = (char*)B->d_planes + poff[j];
        a.occ = B->d_occ + (int64_t)j * B->mpad;
        a.rho = d_rho[m];
        // first-derivative terms go straight to the gradient rows: term[c][p]
        // for c = 1..3 is row c-1 of d_grad[m]
        a.term = d_grad[m] ? (double*)((uintptr_t)d_grad[m] - sizeof(double) * (uintptr_t)npts[m]) : nullptr;
        a.term2 = nullptr;
        a.ld_term = npts[m];
        a.need_mask = d_grad[m] ? 0xE : 0;
        a.brick = g.regular;
        args[j] = a;
        PrepItem it{};
        it.colmap = bs->d_colmap; it.C = d_c[m]; it.occ = d_occ[m]; it.planes = (void*)a.C; it.occ_pad = (double*)a.occ;
        it.nmo = nmo[m]; it.nao = bs->nao; it.kpad = a.kpad;
        items[j] = it;
        istart[j + 1] = istart[j] + B->mpad * a.kpad + B->mpad;
    }
    B->prep_total = istart[nmol];
    B->h_mols = args;
    int rc = batch_set_shape(B, 2, 2, 2);
    if (!rc) rc = upload(&B->d_items, items, "batch: prep items");
    if (!rc) rc = upload(&B->d_item_start, istart, "batch: prep offsets");
    if (rc) {
        okmi_batch_free(B);
        return rc;
    }
    *out = B;
    return 0;
}

int okmi_batch_enqueue(OkmiBatch* B, hipStream_t st)
{
    const int64_t pblocks = std::min<int64_t>((B->prep_total + 255) / 256, 8192);
    hipLaunchKernelGGL(k_prep_batch, dim3((unsigned)pblocks), dim3(256), 0, st, B->d_items, B->d_item_start, B->nmol,
                       B->prep_total, B->prec, B->mpad);
    BatchArgs a{};
    a.mols = B->d_mols; a.tile_start = B->d_tile_start; a.nmol = B->nmol; a.ntile = B->ntile;
    const unsigned grid = B->grid ? (unsigned)std::min<int64_t>(B->grid, B->ntile) : (unsigned)B->ntile;
    switch (B->prec) {
    case OKMI_PREC_F32: return launch_batch<OKMI_PREC_F32>(a, B->pass, B->mt, B->lm, grid, st);
    case OKMI_PREC_BF16X3: return launch_batch<OKMI_PREC_BF16X3>(a, B->pass, B->mt, B->lm, grid, st);
    default: return launch_batch<OKMI_PREC_BF16X6>(a, B->pass, B->mt, B->lm, grid, st);
    }
}

int okmi_batch_run(OkmiBatch* B, int use_graph, hipStream_t st)
{
    if (!use_graph) {
        OkmiProfScope prof("rho_batch", st);
        int rc = okmi_batch_enqueue(B, st);
        if (rc) return rc;
        return okmi_check_launch("batch");
    }
    if (!st) {
        okmi_set_error("batch: graph launches need a non-default stream");
        return 1;
    }
    if (B->exec && B->graph_stream != st) {
        (void)hipGraphExecDestroy(B->exec);
        (void)hipGraphDestroy(B->graph);
        B->exec = nullptr;
        B->graph = nullptr;
    }
    if (!B->exec) {
        hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return okmi_hip_fail(e, "batch: begin capture");
        const int rc = okmi_batch_enqueue(B, st);
        hipGraph_t g = nullptr;
        e = hipStreamEndCapture(st, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return okmi_hip_fail(e, "batch: end capture");
        e = hipGraphInstantiate(&B->exec, g, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            (void)hipGraphDestroy(g);
            return okmi_hip_fail(e, "batch: instantiate graph");
        }
        B->graph = g;
        B->graph_stream = st;
    }
    OkmiProfScope prof("rho_batch", st);
    const hipError_t e = hipGraphLaunch(B->exec, st);
    if (e != hipSuccess) return okmi_hip_fail(e, "batch: graph launch");
    return 0;
}

int okmi_batch_set_grid_impl(OkmiBatch* B, int blocks)
{
    if (blocks < 0) {
        okmi_set_error("batch: negative workgroup count");
        return 1;
    }
    if ((unsigned)blocks != B->grid && B->exec) {
        (void)hipGraphExecDestroy(B->exec);
        (void)hipGraphDestroy(B->graph);
        B->exec = nullptr;
        B->graph = nullptr;
    }
    B->grid = (unsigned)blocks;
    return 0;
}

int64_t okmi_batch_ntile(const OkmiBatch* B) { return B->ntile; }

int okmi_prepare_fused(OkmiBasis* b) { return prepare_fused_tables(b); }
