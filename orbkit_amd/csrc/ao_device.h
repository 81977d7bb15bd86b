// Per-point evaluation of one contracted Cartesian shell and its
// derivatives, templated on the angular momentum so that every monomial
// index is a compile-time constant (no register indexing, no scratch).
//
// Math (orbkit/c_grid-based.c:38-71, orbkit/c_support.c:28-175): with
// X = x - A_x etc., r2 = X^2+Y^2+Z^2 and radial sums over primitives
//   S0 = sum_p d_p e_p,  S1 = sum_p (-2 a_p) d_p e_p,  S2 = sum_p 4 a_p^2 d_p e_p,
// (e_p = exp(-a_p r2)) a component x^lx y^ly z^lz gives
//   drv 0   : X^lx Y^ly Z^lz S0
//   d/dx    : (lx X^(lx-1) S0 + X^(lx+1) S1) Y^ly Z^lz
//   d2/dx2  : (lx(lx-1) X^(lx-2) S0 + (2lx+1) X^lx S1 + X^(lx+2) S2) Y^ly Z^lz
//   d2/dxdy : X^(lx+1) Y^(ly+1) Z^lz S2 + K S0, K as orbkit/c_support.c:121-168
//             (lowering term only; the reference's mixed-derivative formula,
//             reproduced for parity).
#pragma once
#include "okmi_common.h"

template <int L, typename R>
struct ShellPows {
    R p[3][L + 3];
    __device__ __forceinline__ void init(R X, R Y, R Z)
    {
        p[0][0] = R(1); p[1][0] = R(1); p[2][0] = R(1);
#pragma unroll
        for (int n = 1; n < L + 3; ++n) {
            p[0][n] = p[0][n - 1] * X;
            p[1][n] = p[1][n - 1] * Y;
            p[2][n] = p[2][n - 1] * Z;
        }
    }
};

// monomial with one exponent shifted; exponent < 0 never reaches here
template <int L, typename R>
__device__ __forceinline__ R mono(const ShellPows<L, R>& P, int lx, int ly, int lz)
{
    return P.p[0][lx] * P.p[1][ly] * P.p[2][lz];
}

// value of Cartesian component C (canonical order) for derivative code D
template <int L, int C, int D, typename R>
__device__ __forceinline__ R cart_comp(const ShellPows<L, R>& P, R S0, R S1, R S2)
{
    constexpr int e[3] = {cart_exp(L, C, 0), cart_exp(L, C, 1), cart_exp(L, C, 2)};
    if constexpr (D == 0) {
        return mono<L, R>(P, e[0], e[1], e[2]) * S0;
    } else if constexpr (D <= 3) {
        constexpr int a = D - 1;
        constexpr int u0 = e[0] + (a == 0), u1 = e[1] + (a == 1), u2 = e[2] + (a == 2);
        R v = mono<L, R>(P, u0, u1, u2) * S1;
        if constexpr (e[a] > 0) {
            constexpr int d0 = e[0] - (a == 0), d1 = e[1] - (a == 1), d2 = e[2] - (a == 2);
            v += R(e[a]) * mono<L, R>(P, d0, d1, d2) * S0;
        }
        return v;
    } else if constexpr (D <= 6) {
        constexpr int a = D - 4;
        constexpr int u0 = e[0] + 2 * (a == 0), u1 = e[1] + 2 * (a == 1), u2 = e[2] + 2 * (a == 2);
        R v = mono<L, R>(P, u0, u1, u2) * S2 + R(2 * e[a] + 1) * mono<L, R>(P, e[0], e[1], e[2]) * S1;
        if constexpr (e[a] >= 2) {
            constexpr int d0 = e[0] - 2 * (a == 0), d1 = e[1] - 2 * (a == 1), d2 = e[2] - 2 * (a == 2);
            v += R(e[a] * (e[a] - 1)) * mono<L, R>(P, d0, d1, d2) * S0;
        }
        return v;
    } else {   // 7: xy, 8: xz, 9: yz
        constexpr int i = D == 9 ? 1 : 0;
        constexpr int j = D == 7 ? 1 : 2;
        constexpr int u0 = e[0] + (i == 0) + (j == 0), u1 = e[1] + (i == 1) + (j == 1), u2 = e[2] + (i == 2) + (j == 2);
        R v = mono<L, R>(P, u0, u1, u2) * S2;
        if constexpr (e[i] == 0 && e[j] > 0) {
            constexpr int d0 = e[0] - (j == 0), d1 = e[1] - (j == 1), d2 = e[2] - (j == 2);
            v += R(e[j]) * mono<L, R>(P, d0, d1, d2) * S0;
        } else if constexpr (e[i] > 0 && e[j] == 0) {
            constexpr int d0 = e[0] - (i == 0), d1 = e[1] - (i == 1), d2 = e[2] - (i == 2);
            v += R(e[i]) * mono<L, R>(P, d0, d1, d2) * S0;
        } else if constexpr (e[i] > 0 && e[j] > 0) {
            constexpr int d0 = e[0] - (i == 0) - (j == 0), d1 = e[1] - (i == 1) - (j == 1),
                          d2 = e[2] - (i == 2) - (j == 2);
            v += R(e[i] * e[j]) * mono<L, R>(P, d0, d1, d2) * S0;
        }
        return v;
    }
}

template <int L, int D, typename R, int C = 0>
__device__ __forceinline__ void cart_all(const ShellPows<L, R>& P, R S0, R S1, R S2, R* out)
{
    if constexpr (C < ncart(L)) {
        out[C] = cart_comp<L, C, D, R>(P, S0, S1, S2);
        cart_all<L, D, R, C + 1>(P, S0, S1, S2, out);
    }
}

// canonical index of component (lx, ly, lz) of angular momentum l
__host__ __device__ constexpr int cart_index(int l, int lx, int ly, int lz)
{
    for (int c = 0; c < ncart(l); ++c)
        if (cart_exp(l, c, 0) == lx && cart_exp(l, c, 1) == ly && cart_exp(l, c, 2) == lz) return c;
    return -1;
}

// Value and first derivatives of all components from shared monomials:
//   chi_c = m_c S0,   d_a chi_c = (R_a S1) m_c + e_a (m_{c - 1_a} S0),
// m_c the degree-L monomial of component c, m_{c - 1_a} the degree-(L-1)
// one with exponent a lowered (the same terms as cart_comp, D = 0..3,
// evaluated once per shell and point instead of once per derivative).
template <int L, int C, typename R>
__device__ __forceinline__ void first_mono(const ShellPows<L, R>& P, R* m)
{
    if constexpr (C < ncart(L)) {
        m[C] = mono<L, R>(P, cart_exp(L, C, 0), cart_exp(L, C, 1), cart_exp(L, C, 2));
        first_mono<L, C + 1, R>(P, m);
    }
}

template <int L, int C, typename R>
__device__ __forceinline__ void lower_mono(const ShellPows<L, R>& P, R S0, R* u)
{
    if constexpr (L > 0 && C < ncart(L - 1)) {
        u[C] = mono<L, R>(P, cart_exp(L - 1, C, 0), cart_exp(L - 1, C, 1), cart_exp(L - 1, C, 2)) * S0;
        lower_mono<L, C + 1, R>(P, S0, u);
    }
}

// out[c] = d_a chi_c for axis A, with RS1 = R_a S1
template <int L, int A, int C, typename R>
__device__ __forceinline__ void first_deriv(const R* m, const R* u, R RS1, R* out)
{
    if constexpr (C < ncart(L)) {
        constexpr int e0 = cart_exp(L, C, 0), e1 = cart_exp(L, C, 1), e2 = cart_exp(L, C, 2);
        constexpr int ea = A == 0 ? e0 : A == 1 ? e1 : e2;
        if constexpr (ea > 0) {
            constexpr int lo = cart_index(L - 1, e0 - (A == 0), e1 - (A == 1), e2 - (A == 2));
            out[C] = fma(RS1, m[C], R(ea) * u[lo]);
        } else {
            out[C] = RS1 * m[C];
        }
        first_deriv<L, A, C + 1, R>(m, u, RS1, out);
    }
}

// sum_c w_c e_a(c) u_{c - 1_a} for axis A, u the degree-(L-1) monomials
// (times S0): the lowered-exponent half of sum_c w_c d_a chi_c
template <int L, int A, int C, typename R>
__device__ __forceinline__ R lowered_dot(const R* w, const R* u, R acc)
{
    if constexpr (C < ncart(L)) {
        constexpr int e0 = cart_exp(L, C, 0), e1 = cart_exp(L, C, 1), e2 = cart_exp(L, C, 2);
        constexpr int ea = A == 0 ? e0 : A == 1 ? e1 : e2;
        if constexpr (ea > 0) {
            constexpr int lo = cart_index(L - 1, e0 - (A == 0), e1 - (A == 1), e2 - (A == 2));
            acc = fma(w[C] * R(ea), u[lo], acc);
        }
        return lowered_dot<L, A, C + 1, R>(w, u, acc);
    } else {
        return acc;
    }
}

// runtime (wave-uniform) derivative code -> compile-time evaluation
template <int L, typename R>
__device__ __forceinline__ void cart_drv(const ShellPows<L, R>& P, int d, R S0, R S1, R S2, R* out)
{
    switch (d) {
    case 0: cart_all<L, 0, R>(P, S0, S1, S2, out); break;
    case 1: cart_all<L, 1, R>(P, S0, S1, S2, out); break;
    case 2: cart_all<L, 2, R>(P, S0, S1, S2, out); break;
    case 3: cart_all<L, 3, R>(P, S0, S1, S2, out); break;
    case 4: cart_all<L, 4, R>(P, S0, S1, S2, out); break;
    case 5: cart_all<L, 5, R>(P, S0, S1, S2, out); break;
    case 6: cart_all<L, 6, R>(P, S0, S1, S2, out); break;
    case 7: cart_all<L, 7, R>(P, S0, S1, S2, out); break;
    case 8: cart_all<L, 8, R>(P, S0, S1, S2, out); break;
    default: cart_all<L, 9, R>(P, S0, S1, S2, out); break;
    }
}

// exp(-t) in FP64 for t >= 0; t beyond 746 underflows to +0 exactly as libm.
__device__ __forceinline__ double gauss_f64(double t)
{
    return t < 746.0 ? exp(-t) : 0.0;
}

// exp(-t) to FP32 accuracy for any t >= 0 given in FP64: the range reduction
// t*log2(e) = n + f (n integer, 0 <= f < 1) is done in FP64 so that large
// arguments (tight exponents far from their centre) keep a full FP32
// mantissa; 2^-f uses the hardware exp2 (v_exp_f32) and 2^-n an exponent
// adjustment.  Results below FP32's range flush to +0.
__device__ __forceinline__ float gauss_f32(double t)
{
    const double u = t * 1.4426950408889634;   // log2(e)
    if (u >= 150.0) return 0.0f;
    const double n = floor(u);
    const float f = (float)(u - n);
    return __builtin_amdgcn_ldexpf(__builtin_amdgcn_exp2f(-f), -(int)n);
}

// exp(-t) for t >= 0 in FP64 with 2.6e-9 relative error (under 1/20 of an
// FP32 half-ulp), for the paths whose AO values end in FP32 or narrower.
// Cody-Waite reduction t = k ln2 + r, |r| <= ln2/2 (ln2 split in a high part
// exact in 32 bits and a low part), then exp(s), s = -r, by a degree-6
// polynomial: the Chebyshev interpolant of exp on [-ln2/2, ln2/2] (7
// Chebyshev nodes, converted to monomial coefficients), and 2^-k by exponent
// adjustment, which underflows to +0 for large t.  Branch-free: t is
// clamped to 1e4 (exp(-1e4) is +0 in FP64).
__device__ __forceinline__ double gauss_f64_fast(double t)
{
    t = fmin(t, 1.0e4);
    const double k = __builtin_rint(t * 1.4426950408889634);
    double r = fma(-k, 6.93147180369123816490e-01, t);
    r = fma(-k, 1.90821492927058770002e-10, r);
    const double s = -r;
    double p = 0.0013941118940759135;
    p = fma(p, s, 0.0083751347746787615);
    p = fma(p, s, 0.041666352770387056);
    p = fma(p, s, 0.16666415414041169);
    p = fma(p, s, 0.50000000471463202);
    p = fma(p, s, 1.0000000377388703);
    p = fma(p, s, 0.99999999999999989);
    return ldexp(p, -(int)k);
}

// Standard real-solid-harmonic rows in orbkit's (l, m) order
// (m = 0, 1, -1, 2, -2, ...; p: x, y, z), each as at most three canonical
// Cartesian components with orbkit's cart2sph coefficient (orbkit/tools.py:
// 148-191) times the 1/sqrt((2lx-1)!!(2ly-1)!!(2lz-1)!!) normalisation
// factor.  basis_create checks each spherical shell's transform against this
// table before it is evaluated through it.
struct SphTerm { int c; double v; };
struct SphRow { int n; SphTerm t[3]; };
__host__ __device__ constexpr SphRow sph_row(int l, int r)
{
    constexpr SphRow L0[1] = {{1, {{0, 1.0}, {0, 0.0}, {0, 0.0}}}};
    constexpr SphRow L1[3] = {{1, {{0, 1.0}, {0, 0.0}, {0, 0.0}}},
                              {1, {{1, 1.0}, {0, 0.0}, {0, 0.0}}},
                              {1, {{2, 1.0}, {0, 0.0}, {0, 0.0}}}};
    constexpr SphRow L2[5] = {{3, {{0, -0.28867513459481292}, {1, -0.28867513459481292}, {2, 0.57735026918962584}}},
                              {1, {{4, 1.0}, {0, 0.0}, {0, 0.0}}},
                              {1, {{5, 1.0}, {0, 0.0}, {0, 0.0}}},
                              {2, {{0, 0.5}, {1, -0.5}, {0, 0.0}}},
                              {1, {{3, 1.0}, {0, 0.0}, {0, 0.0}}}};
    constexpr SphRow L3[7] = {{3, {{2, 0.2581988897471611}, {5, -0.3872983346207417}, {8, -0.3872983346207417}}},
                              {3, {{0, -0.15811388300841897}, {3, -0.15811388300841897}, {6, 0.63245553203367599}}},
                              {3, {{1, -0.15811388300841897}, {4, -0.15811388300841897}, {7, 0.63245553203367599}}},
                              {2, {{5, 0.5}, {8, -0.5}, {0, 0.0}}},
                              {1, {{9, 1.0}, {0, 0.0}, {0, 0.0}}},
                              {2, {{0, 0.20412414523193148}, {3, -0.61237243569579447}, {0, 0.0}}},
                              {2, {{1, -0.20412414523193148}, {4, 0.61237243569579447}, {0, 0.0}}}};
    return l == 0 ? L0[r] : l == 1 ? L1[r] : l == 2 ? L2[r] : L3[r];
}

// Output rows of a standard spherical shell (sph_row table) or a standard
// Cartesian shell (component norm factors) from the canonical Cartesian
// values c, unrolled at compile time: store(r, value) for each row r.
template <int L, int R, typename Store>
__device__ __forceinline__ void sph_rows_to(const double* c, Store&& store)
{
    if constexpr (L <= 3 && R < 2 * L + 1) {
        constexpr SphRow row = sph_row(L, R);
        double v = row.t[0].v * c[row.t[0].c];
        if constexpr (row.n > 1) v = fma(row.t[1].v, c[row.t[1].c], v);
        if constexpr (row.n > 2) v = fma(row.t[2].v, c[row.t[2].c], v);
        store(R, v);
        sph_rows_to<L, R + 1>(c, store);
    }
}

// 1/sqrt((2lx-1)!!(2ly-1)!!(2lz-1)!!) of canonical component c
__host__ __device__ constexpr double cart_norm(int l, int c)
{
    int d = 1;
    for (int a = 0; a < 3; ++a) {
        int e = cart_exp(l, c, a);
        for (int n = 2 * e - 1; n > 0; n -= 2) d *= n;
    }
    return d == 1 ? 1.0 : d == 3 ? 0.57735026918962584 : d == 9 ? 0.33333333333333333
         : d == 15 ? 0.2581988897471611 : d == 105 ? 0.09759000729485333 : 0.0;
}

template <int L, int R, typename Store>
__device__ __forceinline__ void cart_rows_to(const double* c, Store&& store)
{
    if constexpr (R < ncart(L)) {
        store(R, c[R] * cart_norm(L, R));
        cart_rows_to<L, R + 1>(c, store);
    }
}
