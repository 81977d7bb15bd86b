// Shared definitions for the gfx950 grid-density kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define OKMI_MAX_L 4          // s..g, the range orbkit's tables cover
#define OKMI_MAX_SETS 10      // derivative codes 0..9

// One contracted shell as the kernels see it (wave-uniform record, read
// through the scalar cache).  Cartesian components are always evaluated in
// the canonical (molden) order of orbkit/tools.py:117-135; the per-shell
// dense map `T` (nrow x ncart(l)) then produces the output rows: identity /
// permutation / renormalisation for Cartesian bases, the cart2sph
// combinations of orbkit/tools.py:148-191 for spherical ones, with the
// 1/sqrt((2lx-1)!!(2ly-1)!!(2lz-1)!!) part of orbkit's primitive norm folded
// in column-wise.
struct ShellRec {
    double cx, cy, cz;   // centre (bohr)
    double rcut2;        // beyond this squared distance every value and
                         // derivative of the shell is below 1e-40
    int l;               // angular momentum (all components share it)
    int nprim;           // primitives
    int prim0;           // first primitive in the primitive table
    int row0;            // first output row
    int nrow;            // output rows of this shell
    int t0;              // offset of this shell's T block
    int atom;            // nucleus index (for nuclear-position gradients)
    int kind;            // T is: 0 general, 1 standard spherical, 2 standard Cartesian
};
static_assert(sizeof(ShellRec) == 64, "ShellRec layout");

enum { SHELL_T_GENERAL = 0, SHELL_T_SPHERICAL = 1, SHELL_T_CARTESIAN = 2 };

// Primitive: exponent alpha and d = c * (2/pi)^(3/4) 2^l alpha^((2l+3)/4)
// (or c alone for pre-normalised bases, orbkit/c_support.c:182).
struct PrimRec {
    double alpha;
    double d;
    double rcut2;        // beyond this squared distance this primitive's terms are below 1e-40
};

// Four primitives of one shell, each field grouped so that a group is read
// by a few wide scalar loads and one wait.  Unused slots have alpha = d = 0
// and rcut2 = -1 (never within range).
struct PrimGroup {
    double alpha[4];
    double d[4];
    double rcut2[4];
};
static_assert(sizeof(PrimGroup) == 96, "PrimGroup layout");

// Points: either a vector grid (x[i], y[i], z[i]) or a regular grid given by
// its axes, flattened with x slowest and z fastest (orbkit/cy_grid.pyx:12-30).
struct GridDesc {
    const double* x;
    const double* y;
    const double* z;
    int64_t nx, ny, nz;   // regular: axis lengths; vector: nx = npts, ny = nz = 0
    int regular;
    int pad;
};

__device__ __forceinline__ void grid_point(const GridDesc& g, int64_t p, double& px, double& py, double& pz)
{
    if (g.regular) {
        int64_t nyz = g.ny * g.nz;
        int64_t i = p / nyz;
        int64_t r = p - i * nyz;
        int64_t j = r / g.nz;
        int64_t k = r - j * g.nz;
        px = g.x[i]; py = g.y[j]; pz = g.z[k];
    } else {
        px = g.x[p]; py = g.y[p]; pz = g.z[p];
    }
}

// Read-only tables are read through the constant address space: loads with
// wave-uniform addresses then become scalar loads (the compiler cannot prove
// on its own that the kernel's stores leave the tables untouched).
// Screening bounds: every AO value and derivative term of a shell or
// primitive is below the bound beyond its radius.  The unscreened FP64 AO
// kernel skips terms below 1e-40; the FP64 paths that screen per brick (the
// fused FP64 kernel, the staged FP64 path, the backward) below 1e-24; the
// paths whose AO values end in FP32 (FP32 and bf16-split fused kernels, the
// many-MO and batch paths) below 1e-12: at any point where the density
// exceeds 1e-10 of its maximum some AO value there is above about 1e-5, and
// terms below 1e-12 stay at the FP32 rounding of such a value (2^-24 * 1e-5).
#define OKMI_SCREEN 1e-40
#define OKMI_SCREEN_FUSED 1e-24
#define OKMI_SCREEN_SPLIT 1e-12

#define OKMI_CONST __attribute__((address_space(4)))
template <typename T>
__device__ __forceinline__ const OKMI_CONST T* cptr(const T* p)
{
    return (const OKMI_CONST T*)p;
}

__device__ __forceinline__ ShellRec load_shell(const ShellRec* base, int i)
{
    const OKMI_CONST ShellRec* p = cptr(base) + i;
    ShellRec s;
    s.cx = p->cx; s.cy = p->cy; s.cz = p->cz; s.rcut2 = p->rcut2;
    s.l = p->l; s.nprim = p->nprim; s.prim0 = p->prim0; s.row0 = p->row0;
    s.nrow = p->nrow; s.t0 = p->t0; s.atom = p->atom; s.kind = p->kind;
    return s;
}

__device__ __forceinline__ PrimRec load_prim(const PrimRec* base, int i)
{
    const OKMI_CONST PrimRec* p = cptr(base) + i;
    PrimRec r;
    r.alpha = p->alpha;
    r.d = p->d;
    r.rcut2 = p->rcut2;
    return r;
}

__device__ __forceinline__ PrimGroup load_group(const PrimGroup* base, int i)
{
    const OKMI_CONST PrimGroup* p = cptr(base) + i;
    PrimGroup g;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        g.alpha[j] = p->alpha[j];
        g.d[j] = p->d[j];
        g.rcut2[j] = p->rcut2[j];
    }
    return g;
}

__device__ __forceinline__ int2 load_int2(const int2* base, int i)
{
    const OKMI_CONST int* p = (const OKMI_CONST int*)cptr(base) + 2 * i;
    return make_int2(p[0], p[1]);
}

__host__ __device__ constexpr int ncart(int l) { return (l + 1) * (l + 2) / 2; }

// Canonical Cartesian exponents, molden order.
__host__ __device__ constexpr int cart_exp(int l, int c, int axis)
{
    // packed as lx*100 + ly*10 + lz
    constexpr int t0[1] = {0};
    constexpr int t1[3] = {100, 10, 1};
    constexpr int t2[6] = {200, 20, 2, 110, 101, 11};
    constexpr int t3[10] = {300, 30, 3, 120, 210, 201, 102, 12, 21, 111};
    constexpr int t4[15] = {400, 40, 4, 310, 301, 130, 31, 103, 13, 220, 202, 22, 211, 121, 112};
    int v = l == 0 ? t0[c] : l == 1 ? t1[c] : l == 2 ? t2[c] : l == 3 ? t3[c] : t4[c];
    return axis == 0 ? v / 100 : axis == 1 ? (v / 10) % 10 : v % 10;
}
