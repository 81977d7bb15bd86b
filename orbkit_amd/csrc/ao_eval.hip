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
    int ndrv;
    int need1, need2;
    int drv[OKMI_MAX_SETS];
    // brick mode (regular grids): wave w of the launch takes brick bk0 + w
    // (4x4x4 points, z fastest) of the nbx x nby x nbz bricks and writes
    // columns 64 w + lane; lanes of a brick outside the grid repeat its
    // nearest grid point, so every column of the nbk bricks is written
    int brick;
    int64_t bk0, nbk, nbx, nby, nbz;
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
// outputs the Cody-Waite one (relative error < 7e-12, far below FP32)
template <typename OutT>
__device__ __forceinline__ double gauss_out(double t)
{
    if constexpr (sizeof(OutT) == 8) return gauss_f64(t);
    else return gauss_f64_fast(t);
}

template <int L, typename OutT>
__device__ __forceinline__ void shell_out(const AoArgs& a, const ShellRec& sh, double d2, double px, double py,
                                          double pz, int64_t i, bool act, OutT* out)
{
    const double X = px - sh.cx, Y = py - sh.cy, Z = pz - sh.cz;
    const double rr = X * X + Y * Y + Z * Z;
    double S0 = 0.0, S1 = 0.0, S2 = 0.0;
    for (int p = 0; p < sh.nprim; ++p) {
        const PrimRec pr = load_prim(a.prims, sh.prim0 + p);
        if (d2 >= pr.rcut2) continue;   // below 1e-40 for every point of the wave
        const double de = pr.d * gauss_out<OutT>(pr.alpha * rr);
        S0 += de;
        if (a.need1) S1 += (-2.0 * pr.alpha) * de;
        if (a.need2) S2 += (4.0 * pr.alpha * pr.alpha) * de;
    }
    ShellPows<L, double> P;
    P.init(X, Y, Z);
    constexpr int NC = ncart(L);
    const OKMI_CONST double* T = cptr(a.T) + sh.t0;
    for (int j = 0; j < a.ndrv; ++j) {
        double c[NC];
        cart_drv<L, double>(P, a.drv[j], S0, S1, S2, c);
        OutT* o = out + j * a.set_stride + (int64_t)sh.row0 * a.ld + i;
        auto store = [&](int r, double v) {
            if (act) o[(int64_t)r * a.ld] = (OutT)v;
        };
        // standard spherical / Cartesian shells: sparse compile-time rows;
        // other transforms (ORCA row factors, custom maps): the dense block
        if (sh.kind == SHELL_T_SPHERICAL) {
            sph_rows_to<L, 0>(c, store);
        } else if (sh.kind == SHELL_T_CARTESIAN) {
            cart_rows_to<L, 0>(c, store);
        } else {
            for (int r = 0; r < sh.nrow; ++r) {
                double v = 0.0;
#pragma unroll
                for (int k = 0; k < NC; ++k) v = fma(T[r * NC + k], c[k], v);
                store(r, v);
            }
        }
    }
}

template <typename OutT>
__global__ __launch_bounds__(256) void k_ao_eval(AoArgs a, OutT* __restrict__ out)
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
    for (int s = 0; s < a.nshell; ++s) {
        const ShellRec sh = load_shell(a.shells, s);
        const double d2 = __longlong_as_double(__builtin_amdgcn_readfirstlane(__double_as_longlong(box_dist2(lo, hi, sh))));
        if (d2 >= sh.rcut2) {
            // every value and derivative of the shell is below 1e-40 here
            if (act)
                for (int j = 0; j < a.ndrv; ++j) {
                    OutT* o = out + j * a.set_stride + (int64_t)sh.row0 * a.ld + i;
                    for (int r = 0; r < sh.nrow; ++r) o[(int64_t)r * a.ld] = (OutT)0;
                }
            continue;
        }
        switch (sh.l) {
        case 0: shell_out<0, OutT>(a, sh, d2, px, py, pz, i, act, out); break;
        case 1: shell_out<1, OutT>(a, sh, d2, px, py, pz, i, act, out); break;
        case 2: shell_out<2, OutT>(a, sh, d2, px, py, pz, i, act, out); break;
        case 3: shell_out<3, OutT>(a, sh, d2, px, py, pz, i, act, out); break;
        default: shell_out<4, OutT>(a, sh, d2, px, py, pz, i, act, out); break;
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
                           void* out, int64_t ld, int64_t set_stride, hipStream_t st)
{
    if (np <= 0 && !bricks) return 0;
    AoArgs a{};
    a.shells = shells;
    a.prims = prims;
    a.T = b->d_T;
    a.grid = g;
    a.p0 = p0;
    a.np = np;
    a.ld = ld;
    a.set_stride = set_stride;
    a.nshell = b->nshell;
    a.ndrv = ndrv;
    a.need1 = a.need2 = 0;
    for (int j = 0; j < ndrv; ++j) {
        a.drv[j] = drv[j];
        if (drv[j] >= 1) a.need1 = 1;
        if (drv[j] >= 4) a.need2 = 1;
    }
    dim3 blk(256), grd((unsigned)((np + 255) / 256));
    if (bricks) {
        if (bricks->nbk <= 0) return 0;
        a.brick = 1;
        a.bk0 = bricks->bk0; a.nbk = bricks->nbk;
        a.nbx = bricks->nbx; a.nby = bricks->nby; a.nbz = bricks->nbz;
        grd = dim3((unsigned)((bricks->nbk + 3) / 4));
    }
    OkmiProfScope prof("ao_eval", st);
    if (out_dtype == OKMI_F64)
        hipLaunchKernelGGL(k_ao_eval<double>, grd, blk, 0, st, a, (double*)out);
    else
        hipLaunchKernelGGL(k_ao_eval<float>, grd, blk, 0, st, a, (float*)out);
    return okmi_check_launch("ao_eval");
}
