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
    int l;               // angular momentum (all components share it)
    int nprim;           // primitives
    int prim0;           // first primitive in the primitive table
    int row0;            // first output row
    int nrow;            // output rows of this shell
    int t0;              // offset of this shell's T block
    int atom;            // nucleus index (for nuclear-position gradients)
    int pad[3];
};
static_assert(sizeof(ShellRec) == 64, "ShellRec layout");

// Primitive: exponent alpha and d = c * (2/pi)^(3/4) 2^l alpha^((2l+3)/4)
// (or c alone for pre-normalised bases, orbkit/c_support.c:182).
struct PrimRec {
    double alpha;
    double d;
};

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
