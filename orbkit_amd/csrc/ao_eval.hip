// AO evaluation kernel: one lane per grid point, wave-uniform loop over
// shells, output rows written as coalesced SoA  out[set][row][ld].
//
// Replaces orbkit's c_lcreator (orbkit/c_grid-based.c:9-79) looped over
// contractions by cy_core.aocreator (orbkit/cy_core.pyx:49-78), the ORCA row
// renormalisation (orbkit/core.py:98-100) and cartesian2spherical
// (orbkit/core.py:135-176): all of these are one pass here.
#include "okmi_internal.h"
#include "ao_device.h"

namespace {

struct AoArgs {
    const ShellRec* shells;
    const PrimRec* prims;
    const double* T;
    GridDesc grid;
    int64_t p0, np;
    int64_t ld;          // row stride of out (elements)
    int64_t set_stride;  // stride between derivative sets (elements)
    int nshell;
    int need1;                 // first derivatives (radial sum S1) needed
    int dst[OKMI_MAX_SETS];    // output set of derivative code c, -1: not written by this launch
    // brick mode (regular grids): wave w of the launch takes brick bk0 + w
    // (4x4x4 points, z fastest) of the nbx x nby x nbz bricks and writes
    // columns 64 w + lane; lanes of a brick outside the grid repeat its
    // nearest grid point, so every column of the nbk bricks is written
    int brick;
    int64_t bk0, nbk, nbx, nby, nbz;
    // optional live-chunk mask (AoLiveMask): the wave walks only the shells
    // of its tile's live chunks (chunk table and work-ordered shells of the
    // fused layout); null: every shell of `shells`
    const unsigned long long* live_mask;
    int mask_words, mask_shift;
    const int2* chunks;        // per chunk: first work entry, shells | used rows << 16
    const ShellRec* wshells;   // work order, internal rows, fused radii, PrimRec primitives
};

// squared distance from a shell centre to the wave's points' bounding box
// (wave-uniform); lo/hi are wave minima/maxima of the point coordinates
__device__ __forceinline__ double box_dist2(const double (&lo)[3], const double (&hi)[3], const ShellRec& sh)
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

// FP64 outputs use the libm-accurate exponential (parity path); FP32
// outputs the reduced one (relative error < 2.6e-9, far below FP32's 6e-8)
template <typename OutT>
__device__ __forceinline__ double gauss_out(double t)
{
    if constexpr (sizeof(OutT) == 8) return gauss_f64(t);
    else return gauss_f64_fast(t);
}

// rows of one shell from its canonical Cartesian values c, stored into the
// AO set starting at o (row stride a.ld)
template <int L, typename OutT>
__device__ __forceinline__ void store_rows(const AoArgs& a, const ShellRec& sh, const double* c, OutT* o, bool act)
{
    constexpr int NC = ncart(L);
    // streaming (non-temporal) stores: the sets are read back by a later
    // kernel from HBM, so the rows are not kept in L2
    auto store = [&](int r, double v) {
        if (act) __builtin_nontemporal_store((OutT)v, &o[(int64_t)r * a.ld]);
    };
    // standard spherical / Cartesian shells: sparse compile-time rows;
    // other transforms (ORCA row factors, custom maps): the dense block
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

// One shell at the lane's point for the derivative codes of group GRP:
// 0 = value and first derivatives (codes 0-3, from shared monomials),
// 1 = second derivatives (codes 4-9, cart_all; the mixed ones with orbkit's
// formula).  Only codes with a destination set are written.
template <int L, int GRP, typename OutT>
__device__ __forceinline__ void shell_out(const AoArgs& a, const ShellRec& sh, double d2, double px, double py,
                                          double pz, int64_t i, bool act, OutT* out)
{
    const double X = px - sh.cx, Y = py - sh.cy, Z = pz - sh.cz;
    const double rr = X * X + Y * Y + Z * Z;
    const bool need1 = GRP == 1 || a.need1;
    double S0 = 0.0, S1 = 0.0, S2 = 0.0;
    for (int p = 0; p < sh.nprim; ++p) {
        const PrimRec pr = load_prim(a.prims, sh.prim0 + p);
        if (d2 >= pr.rcut2) continue;   // below 1e-40 for every point of the wave
        const double de = pr.d * gauss_out<OutT>(pr.alpha * rr);
        S0 += de;
        if (need1) S1 += (-2.0 * pr.alpha) * de;
        if constexpr (GRP == 1) S2 += (4.0 * pr.alpha * pr.alpha) * de;
    }
    ShellPows<L, double> P;
    P.init(X, Y, Z);
    constexpr int NC = ncart(L);
    OutT* o = out + (int64_t)sh.row0 * a.ld + i;
    double c[NC];
    if constexpr (GRP == 0) {
        double m[NC], u[L > 0 ? ncart(L - 1) : 1];
        first_mono<L, 0, double>(P, m);
        if (a.dst[0] >= 0) {
#pragma unroll
            for (int k = 0; k < NC; ++k) c[k] = m[k] * S0;
            store_rows<L, OutT>(a, sh, c, o + a.dst[0] * a.set_stride, act);
        }
        if (!need1) return;
        lower_mono<L, 0, double>(P, S0, u);
        if (a.dst[1] >= 0) {
            first_deriv<L, 0, 0, double>(m, u, X * S1, c);
            store_rows<L, OutT>(a, sh, c, o + a.dst[1] * a.set_stride, act);
        }
        if (a.dst[2] >= 0) {
            first_deriv<L, 1, 0, double>(m, u, Y * S1, c);
            store_rows<L, OutT>(a, sh, c, o + a.dst[2] * a.set_stride, act);
        }
        if (a.dst[3] >= 0) {
            first_deriv<L, 2, 0, double>(m, u, Z * S1, c);
            store_rows<L, OutT>(a, sh, c, o + a.dst[3] * a.set_stride, act);
        }
    } else {
#define OKMI_SECOND(D)                                                             \
        if (a.dst[D] >= 0) {                                                       \
            cart_all<L, D, double>(P, S0, S1, S2, c);                              \
            store_rows<L, OutT>(a, sh, c, o + a.dst[D] * a.set_stride, act);      \
        }
        OKMI_SECOND(4) OKMI_SECOND(5) OKMI_SECOND(6) OKMI_SECOND(7) OKMI_SECOND(8) OKMI_SECOND(9)
#undef OKMI_SECOND
    }
}

// occupancy (waves/SIMD) the register allocation targets without spilling:
// FP32 output 4 (group 0) and 3 (group 1), FP64 output 3 and 2
template <typename OutT, int GRP>
__global__ __launch_bounds__(256, sizeof(OutT) == 4 ? (GRP == 0 ? 4 : 3) : (GRP == 0 ? 3 : 2))
void k_ao_eval(AoArgs a, OutT* __restrict__ out)
{
    int64_t i;
    bool act;
    double px, py, pz;
    if (a.brick) {
        const GridDesc& g = a.grid;
        const int lane = threadIdx.x & 63;
        const int64_t wb = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
        int64_t t = a.bk0 + wb;
        const int64_t bz = t % a.nbz;
        t /= a.nbz;
        const int64_t by = t % a.nby, bx = t / a.nby;
        int64_t ix = 4 * bx + (lane >> 4), iy = 4 * by + ((lane >> 2) & 3), iz = 4 * bz + (lane & 3);
        ix = ix < g.nx ? ix : g.nx - 1;
        iy = iy < g.ny ? iy : g.ny - 1;
        iz = iz < g.nz ? iz : g.nz - 1;
        act = wb < a.nbk;
        i = wb * 64 + lane;
        px = g.x[ix]; py = g.y[iy]; pz = g.z[iz];
    } else {
        i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        act = i < a.np;
        grid_point(a.grid, a.p0 + (act ? i : 0), px, py, pz);
    }
    // bounding box of the wave's points (inactive lanes repeat point 0,
    // which only enlarges the box)
    double lo[3] = {px, py, pz}, hi[3] = {px, py, pz};
#pragma unroll
    for (int o = 1; o < 64; o <<= 1)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            lo[k] = fmin(lo[k], __shfl_xor(lo[k], o));
            hi[k] = fmax(hi[k], __shfl_xor(hi[k], o));
        }
    // one shell for the wave's points: its rows evaluated, or zero when it is
    // out of reach (every value and derivative below the screening bound)
    auto shell = [&](const ShellRec& sh) {
        const double d2 = __longlong_as_double(__builtin_amdgcn_readfirstlane(__double_as_longlong(box_dist2(lo, hi, sh))));
        if (d2 >= sh.rcut2) {
            if (act)
                for (int c = 0; c < OKMI_MAX_SETS; ++c) {
                    if (a.dst[c] < 0) continue;
                    OutT* o = out + a.dst[c] * a.set_stride + (int64_t)sh.row0 * a.ld + i;
                    for (int r = 0; r < sh.nrow; ++r) __builtin_nontemporal_store((OutT)0, &o[(int64_t)r * a.ld]);
                }
            return;
        }
        switch (sh.l) {
        case 0: shell_out<0, GRP, OutT>(a, sh, d2, px, py, pz, i, act, out); break;
        case 1: shell_out<1, GRP, OutT>(a, sh, d2, px, py, pz, i, act, out); break;
        case 2: shell_out<2, GRP, OutT>(a, sh, d2, px, py, pz, i, act, out); break;
        case 3: shell_out<3, GRP, OutT>(a, sh, d2, px, py, pz, i, act, out); break;
        default: shell_out<4, GRP, OutT>(a, sh, d2, px, py, pz, i, act, out); break;
        }
    };
    if (!a.live_mask) {
        for (int s = 0; s < a.nshell; ++s) shell(load_shell(a.shells, s));
        return;
    }
    // the shells of the live chunks of this wave's tile; the rows of the
    // other chunks are never read, and are left unwritten.  Masks exist for
    // the tiles holding the launch's points only: a wave past the last point
    // (the grid's rounding) has nothing to write and stops here.
    const int64_t wv = a.brick ? (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)
                               : ((int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63)) >> 6;
    if (a.brick ? wv >= a.nbk : wv * 64 >= a.np) return;
    const int64_t wt = wv >> a.mask_shift;
    const unsigned long long* mrow = a.live_mask + wt * a.mask_words;
    for (int w = 0; w < a.mask_words; ++w) {
        const unsigned long long m0 = mrow[w];
        unsigned long long m = ((unsigned long long)__builtin_amdgcn_readfirstlane((unsigned)(m0 >> 32)) << 32) |
                               (unsigned)__builtin_amdgcn_readfirstlane((unsigned)m0);
        while (m) {
            const int c = 64 * w + __builtin_ctzll(m);
            m &= m - 1;
            const int2 cw = load_int2(a.chunks, c);
            const int nsh = cw.y & 0xffff;
            for (int e = 0; e < nsh; ++e) shell(load_shell(a.wshells, cw.x + e));
        }
    }
}

}  // namespace

int okmi_launch_ao_eval(const OkmiBasis* b, const GridDesc& g, int64_t p0, int64_t np, const int* drv, int ndrv,
                        int out_dtype, void* out, int64_t ld, int64_t set_stride, hipStream_t st)
{
    return okmi_launch_ao_eval_ex(b, b->d_shells, b->d_prims, g, p0, np, nullptr, drv, ndrv, out_dtype, out, ld,
                                  set_stride, st);
}

int okmi_launch_ao_eval_ex(const OkmiBasis* b, const ShellRec* shells, const PrimRec* prims, const GridDesc& g,
                           int64_t p0, int64_t np, const AoBricks* bricks, const int* drv, int ndrv, int out_dtype,
                           void* out, int64_t ld, int64_t set_stride, hipStream_t st, const AoLiveMask* live)
{
    if (np <= 0 && !bricks) return 0;
    AoArgs a{};
    if (live) {
        a.live_mask = live->mask;
        a.mask_words = live->words;
        a.mask_shift = live->shift;
        a.chunks = b->d_chunks;
        a.wshells = live->wshells ? live->wshells : b->d_wshells;
    }
    a.shells = shells;
    a.prims = prims;
    a.T = b->d_T;
    a.grid = g;
    a.p0 = p0;
    a.np = np;
    a.ld = ld;
    a.set_stride = set_stride;
    a.nshell = b->nshell;
    // each code's first set; the groups write those, repeats are copies
    int first[OKMI_MAX_SETS];
    for (int c = 0; c < OKMI_MAX_SETS; ++c) first[c] = -1;
    for (int j = 0; j < ndrv; ++j) {
        if (drv[j] < 0 || drv[j] >= OKMI_MAX_SETS) {
            okmi_set_error("ao_eval: derivative code %d is not in 0..9", drv[j]);
            return 1;
        }
        if (first[drv[j]] < 0) first[drv[j]] = j;
    }
    dim3 blk(256), grd((unsigned)((np + 255) / 256));
    int64_t width = np;   // columns written per row
    if (bricks) {
        if (bricks->nbk <= 0) return 0;
        a.brick = 1;
        a.bk0 = bricks->bk0; a.nbk = bricks->nbk;
        a.nbx = bricks->nbx; a.nby = bricks->nby; a.nbz = bricks->nbz;
        grd = dim3((unsigned)((bricks->nbk + 3) / 4));
        width = bricks->nbk * 64;
    }
    OkmiProfScope prof("ao_eval", st);
    for (int grp = 0; grp < 2; ++grp) {
        const int c0 = grp == 0 ? 0 : 4, c1 = grp == 0 ? 4 : OKMI_MAX_SETS;
        bool any = false;
        for (int c = 0; c < OKMI_MAX_SETS; ++c) {
            a.dst[c] = c >= c0 && c < c1 ? first[c] : -1;
            any |= a.dst[c] >= 0;
        }
        if (!any) continue;
        a.need1 = first[1] >= 0 || first[2] >= 0 || first[3] >= 0;
        if (out_dtype == OKMI_F64) {
            if (grp == 0) hipLaunchKernelGGL((k_ao_eval<double, 0>), grd, blk, 0, st, a, (double*)out);
            else hipLaunchKernelGGL((k_ao_eval<double, 1>), grd, blk, 0, st, a, (double*)out);
        } else {
            if (grp == 0) hipLaunchKernelGGL((k_ao_eval<float, 0>), grd, blk, 0, st, a, (float*)out);
            else hipLaunchKernelGGL((k_ao_eval<float, 1>), grd, blk, 0, st, a, (float*)out);
        }
    }
    const size_t es = out_dtype == OKMI_F64 ? sizeof(double) : sizeof(float);
    for (int j = 0; j < ndrv; ++j) {
        const int f = first[drv[j]];
        if (f == j) continue;
        char* base = (char*)out;
        const hipError_t e = hipMemcpy2DAsync(base + (size_t)j * set_stride * es, (size_t)ld * es,
                                              base + (size_t)f * set_stride * es, (size_t)ld * es, (size_t)width * es,
                                              (size_t)b->nao, hipMemcpyDeviceToDevice, st);
        if (e != hipSuccess) return okmi_hip_fail(e, "ao_eval: repeated derivative set");
    }
    return okmi_check_launch("ao_eval");
}
