// Backward pass of rho(r) = sum_i occ_i Phi_i(r)^2, Phi = C A(r), for an
// upstream gradient g(r) = dL/drho(r).  orbkit has no derivative code; this
// differentiates the forward path of orbkit/core.py:rho_compute
// (orbkit/core.py:529-607: AO values, MO values C A, occ-weighted squares)
// with respect to the MO coefficients, the primitive exponents and the
// nuclear positions.  Per chunk of points (AO values are recomputed, never
// kept across chunks):
//
//   A        AO values (okmi_launch_ao_eval, FP64)
//   Phi      = C A                                  (FP64 MFMA, k_gemm_f64)
//   B_ir     = 2 occ_i g_r Phi_ir                   (k_bmat)
//   dL/dC   += B A^T                                (k_gemm_tn, FP64 MFMA, split over points,
//                                                    partials summed in a fixed order)
//   G        = C^T B   (= dL/dA)                    (k_gemm_f64 with C^T)
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
#include "okmi_internal.h"
#include "ao_device.h"

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int MAXP = 16;     // primitives per shell handled by one block (more: extra windows)
constexpr int RANGE = 16384; // points per shell-gradient block

__global__ void k_transpose(const double* C, int nmo, int nao, double* CT)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)nmo * nao) return;
    const int64_t m = i / nao, k = i - m * nao;
    CT[k * nmo + m] = C[i];
}

// B[i][p] = 2 occ_i g(p) Phi[i][p] for p < np, 0 for np <= p < ld_used
__global__ void k_bmat(const double* Phi, const double* occ, const double* g, int nmo, int64_t np, int64_t nfill,
                       int64_t ld, double* B)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nfill) return;
    const double gp = p < np ? 2.0 * g[p] : 0.0;
    for (int i = 0; i < nmo; ++i) B[i * ld + p] = p < np ? occ[i] * gp * Phi[i * ld + p] : 0.0;
}

// dL/dC partials: part[s][m][n] = sum over the points k of range s of
// B[m][k] A[n][k].  One wave per (point range, tile of 16 NT AOs) computes
// every MO tile of that AO tile (groups of up to MG tiles), so each AO row is
// read from HBM once; the waves of one point range run on one XCD (blocks
// are dealt to the 8 XCDs round-robin: range s takes blocks with index
// % 8 == s % 8), where they share that range's B rows through the XCD's L2.
// v_mfma_f64_16x16x4f64 with the K index of lane (r = l%16, q = l/16) in
// MFMA j of a 16-point step set to 4q + j, so that each lane reads 4
// consecutive points (one 32-byte load) per row.  Rows are padded with
// zeros up to a multiple of 16 points (krange a multiple of 16, pointers
// 32-byte aligned, ld a multiple of 4).  Item i of lane l of the
// accumulator holds D[4i + l/16][l%16].  The partials are written (no
// atomics) and summed in a fixed order by k_reduce_dc.
constexpr int NT = 4;
template <int MG>
__global__ __launch_bounds__(64) void k_gemm_tn(const double* __restrict__ B, int nmo, const double* __restrict__ A,
                                                int nao, int64_t kend, int64_t ld, int64_t krange, int nsplit,
                                                double* __restrict__ part)
{
    const int ntile = (nao + 16 * NT - 1) / (16 * NT);
    const int xcd = blockIdx.x & 7, local = blockIdx.x >> 3;
    const int t = local % ntile, split = (local / ntile) * 8 + xcd;
    if (split >= nsplit) return;
    const int lane = threadIdx.x;
    const int r = lane & 15, q = lane >> 4;
    const int n0 = t * (16 * NT);
    const int64_t k0 = (int64_t)split * krange;
    const int64_t k1 = min(kend, k0 + krange);
    const d4* ap[NT];
    bool nrow[NT];
#pragma unroll
    for (int u = 0; u < NT; ++u) {
        nrow[u] = n0 + 16 * u + r < nao;
        ap[u] = (const d4*)(A + (int64_t)(nrow[u] ? n0 + 16 * u + r : 0) * ld + 4 * q);
    }
    double* out = part + (int64_t)split * nmo * nao;
    const int nmt = (nmo + 15) / 16;
    for (int g0 = 0; g0 < nmt; g0 += MG) {
        const d4* bp[MG];
        bool mrow[MG];
#pragma unroll
        for (int mt = 0; mt < MG; ++mt) {
            const int m = (g0 + mt) * 16 + r;
            mrow[mt] = m < nmo;
            bp[mt] = (const d4*)(B + (int64_t)(mrow[mt] ? m : 0) * ld + 4 * q);
        }
        d4 acc[MG][NT];
#pragma unroll
        for (int mt = 0; mt < MG; ++mt)
#pragma unroll
            for (int u = 0; u < NT; ++u) acc[mt][u] = d4{0.0, 0.0, 0.0, 0.0};
        for (int64_t k = k0; k < k1; k += 16) {
            d4 vb[MG], va[NT];
#pragma unroll
            for (int mt = 0; mt < MG; ++mt) {
                vb[mt] = bp[mt][k >> 2];
                if (!mrow[mt]) vb[mt] = d4{0.0, 0.0, 0.0, 0.0};
            }
#pragma unroll
            for (int u = 0; u < NT; ++u) {
                va[u] = ap[u][k >> 2];
                if (!nrow[u]) va[u] = d4{0.0, 0.0, 0.0, 0.0};
            }
#pragma unroll
            for (int j = 0; j < 4; ++j)
#pragma unroll
                for (int mt = 0; mt < MG; ++mt)
#pragma unroll
                    for (int u = 0; u < NT; ++u)
                        acc[mt][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(vb[mt][j], va[u][j], acc[mt][u], 0, 0, 0);
        }
#pragma unroll
        for (int mt = 0; mt < MG; ++mt)
#pragma unroll
            for (int u = 0; u < NT; ++u)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int m = (g0 + mt) * 16 + 4 * i + q, n = n0 + 16 * u + r;
                    if (m < nmo && n < nao) out[(int64_t)m * nao + n] = acc[mt][u][i];
                }
    }
}

// dC[e] += sum_s part[s][e] for e < n, s = 0 .. nsplit - 1 in order
__global__ void k_reduce_dc(const double* part, int nsplit, int64_t n, double* dC)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    double v = 0.0;
    for (int s = 0; s < nsplit; ++s) v += part[(int64_t)s * n + e];
    dC[e] += v;
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) v += __shfl_xor(v, o);
    return v;
}

struct ShellGradArgs {
    const ShellRec* shells;
    const PrimRec* prims;
    const double* T;
    GridDesc grid;
    int64_t p0, np, ld;
    int nprim_total;
    int is_normalized;
    const double* G;     // [nao][ld]
    double* partial;     // [nrange][nprim_total + 3 nshell]
    int64_t ncol;
};

// Per-lane contributions of one point to the shell's position gradient and
// to the exponent gradients of primitives [pw0, pw0 + MP), accumulated into
// gr / ga when `near` (the lane's point is within the shell's radius).
// Every primitive's d e is computed once (kept for those of the window).
template <int L, int MP>
__device__ __forceinline__ void shell_point(const ShellGradArgs& a, const ShellRec& sh, int pw0, int pend,
                                            double px, double py, double pz, int64_t i, double nfac, bool near,
                                            double (&gr)[3], double (&ga)[MP])
{
    constexpr int NC = ncart(L);
    const double X = px - sh.cx, Y = py - sh.cy, Z = pz - sh.cz;
    const double rr = X * X + Y * Y + Z * Z;
    double S0 = 0.0, S1 = 0.0;
    double dw[MP];
#pragma unroll
    for (int j = 0; j < MP; ++j) dw[j] = 0.0;
    for (int p = 0; p < sh.nprim; ++p) {
        const PrimRec pr = load_prim(a.prims, sh.prim0 + p);
        const double de = pr.d * gauss_f64(pr.alpha * rr);
        S0 += de;
        S1 = fma(-2.0 * pr.alpha, de, S1);
#pragma unroll
        for (int j = 0; j < MP; ++j)
            if (p == pw0 + j) dw[j] = de;
    }
    // w_c = sum_rows G_row T[row][c]
    const OKMI_CONST double* T = cptr(a.T) + sh.t0;
    double w[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) w[k] = 0.0;
    for (int rw = 0; rw < sh.nrow; ++rw) {
        const double g = a.G[(int64_t)(sh.row0 + rw) * a.ld + i];
#pragma unroll
        for (int k = 0; k < NC; ++k) w[k] = fma(g, T[rw * NC + k], w[k]);
    }
    // q = sum_c w_c m_c and, with the lowered monomials u (times S0),
    // sum_c w_c d_a chi_c = R_a S1 q + sum_c w_c e_a(c) u_{c - 1_a}
    ShellPows<L, double> P;
    P.init(X, Y, Z);
    double m[NC];
    first_mono<L, 0, double>(P, m);
    double q = 0.0;
#pragma unroll
    for (int k = 0; k < NC; ++k) q = fma(w[k], m[k], q);
    if (pw0 == 0 && near) {
        double u[L > 0 ? ncart(L - 1) : 1];
        lower_mono<L, 0, double>(P, S0, u);
        gr[0] -= fma(X * S1, q, lowered_dot<L, 0, 0, double>(w, u, 0.0));
        gr[1] -= fma(Y * S1, q, lowered_dot<L, 1, 0, double>(w, u, 0.0));
        gr[2] -= fma(Z * S1, q, lowered_dot<L, 2, 0, double>(w, u, 0.0));
    }
    if (!near) q = 0.0;
#pragma unroll
    for (int j = 0; j < MP; ++j) {
        if (pw0 + j < pend) {
            const double alpha = load_prim(a.prims, sh.prim0 + pw0 + j).alpha;
            ga[j] = fma(q * dw[j], nfac / alpha - rr, ga[j]);
        }
    }
}

template <int L, int MP>
__device__ void shell_grad(const ShellGradArgs& a, const ShellRec& sh, int pw0)
{
    __shared__ double red[4][3 + MP];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int pend = min(sh.nprim, pw0 + MP);
    const double nfac = a.is_normalized ? 0.0 : (2.0 * L + 3.0) / 4.0;
    double gr[3] = {0.0, 0.0, 0.0};
    double ga[MP];
#pragma unroll
    for (int j = 0; j < MP; ++j) ga[j] = 0.0;
    const int64_t r0 = (int64_t)blockIdx.x * RANGE;
    const int64_t r1 = min(a.np, r0 + RANGE);
    for (int64_t b = r0; b < r1; b += 256) {
        const int64_t i = b + threadIdx.x;
        const bool act = i < r1;
        double px, py, pz;
        grid_point(a.grid, a.p0 + (act ? i : r0), px, py, pz);
        const double X = px - sh.cx, Y = py - sh.cy, Z = pz - sh.cz;
        const bool near = act && X * X + Y * Y + Z * Z < sh.rcut2;
        if (!__builtin_amdgcn_ballot_w64(near)) continue;   // no point of this wave reaches the shell
        shell_point<L, MP>(a, sh, pw0, pend, px, py, pz, act ? i : r0, nfac, near, gr, ga);
    }
    // block reduction in a fixed order (deterministic)
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double v = wave_sum(gr[d]);
        if (lane == 0) red[wave][d] = v;
    }
#pragma unroll
    for (int j = 0; j < MP; ++j) {
        if (pw0 + j < pend) {
            const double v = wave_sum(ga[j]);
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
template <int MP>
__global__ __launch_bounds__(256) void k_shell_grad(ShellGradArgs a)
{
    const ShellRec sh = load_shell(a.shells, blockIdx.y);
    const int pw0 = blockIdx.z * MP;
    if (pw0 >= sh.nprim) return;
    switch (sh.l) {
    case 0: shell_grad<0, MP>(a, sh, pw0); break;
    case 1: shell_grad<1, MP>(a, sh, pw0); break;
    case 2: shell_grad<2, MP>(a, sh, pw0); break;
    case 3: shell_grad<3, MP>(a, sh, pw0); break;
    default: shell_grad<4, MP>(a, sh, pw0); break;
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

__global__ void k_scatter_grad(const double* colsum, const ShellRec* shells, int nshell, int nprim_total,
                               int natoms, double* dalpha, double* dR)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nprim_total) {
        dalpha[t] += colsum[t];
    } else if (t < nprim_total + 3 * natoms) {
        const int a = (t - nprim_total) / 3, d = (t - nprim_total) % 3;
        double s = 0.0;
        for (int sh = 0; sh < nshell; ++sh)
            if (load_shell(shells, sh).atom == a) s += colsum[nprim_total + 3 * sh + d];
        dR[3 * a + d] += s;
    }
}

// zero-filled padding of a [rows][ld] buffer's columns [c0, c1) (fill kernel,
// no memset, so the sequence can sit in a captured graph)
__global__ void k_zero_cols(double* X, int rows, int64_t ld, int64_t c0, int64_t c1)
{
    const int64_t n = c1 - c0;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * rows) return;
    X[(i / n) * ld + c0 + i % n] = 0.0;
}

inline size_t pad64(size_t n) { return (n + 63) & ~(size_t)63; }

struct BwdLayout {
    int64_t chunk = 0;
    int nrange = 0;
    int64_t ncol = 0;
    int nsplit_max = 0;   // point ranges of the dL/dC contraction
    size_t ct = 0, a = 0, phi = 0, bm = 0, g = 0, part = 0, cols = 0, dcp = 0, total = 0;   // offsets in doubles
};

// dL/dC partials: up to 512 point ranges, at most 64 MB of them
constexpr size_t DC_PART_BYTES = (size_t)64 << 20;
int dc_split_max(const OkmiBasis* b, int nmo)
{
    const size_t one = (size_t)nmo * b->nao * sizeof(double);
    return (int)std::max<size_t>(1, std::min<size_t>(512, DC_PART_BYTES / std::max<size_t>(one, 1)));
}

BwdLayout bwd_layout(const OkmiBasis* b, int nmo, int64_t chunk)
{
    BwdLayout L;
    L.chunk = (chunk + 63) & ~(int64_t)63;
    L.nrange = (int)((L.chunk + RANGE - 1) / RANGE);
    L.ncol = b->nprim + 3 * (int64_t)b->nshell;
    size_t o = 0;
    L.ct = o; o += pad64((size_t)b->nao * nmo);
    L.a = o; o += pad64((size_t)b->nao * L.chunk);
    L.phi = o; o += pad64((size_t)nmo * L.chunk);
    L.bm = o; o += pad64((size_t)nmo * L.chunk);
    L.g = o; o += pad64((size_t)b->nao * L.chunk);
    L.part = o; o += pad64((size_t)L.nrange * L.ncol);
    L.cols = o; o += pad64((size_t)L.ncol);
    // ranges of at least 256 points: a short chunk needs few partials
    L.nsplit_max = (int)std::max<int64_t>(1, std::min<int64_t>(dc_split_max(b, nmo), L.chunk / 256));
    L.dcp = o; o += pad64((size_t)L.nsplit_max * nmo * b->nao);
    L.total = o;
    return L;
}

}  // namespace

size_t okmi_backward_ws_bytes(const OkmiBasis* b, int nmo, int64_t chunk)
{
    return sizeof(double) * bwd_layout(b, nmo, chunk).total + 256;
}

int okmi_launch_rho_backward(const OkmiBasis* b, const GridDesc& g, int64_t p0, int64_t np, const double* d_c,
                             const double* d_occ, int nmo, const double* d_grad, double* d_dc, double* d_dalpha,
                             double* d_dr, void* ws, size_t ws_bytes, hipStream_t st)
{
    if (np <= 0 || nmo <= 0) return 0;
    // largest chunk (a multiple of 64 points, at most np rounded up) whose layout fits
    int64_t lo = 64, hi = (np + 63) & ~(int64_t)63;
    if (!ws || okmi_backward_ws_bytes(b, nmo, lo) > ws_bytes) {
        okmi_set_error("rho_backward: workspace of %zu bytes is too small (need at least %zu)", ws_bytes,
                       okmi_backward_ws_bytes(b, nmo, lo));
        return 1;
    }
    while (lo < hi) {
        const int64_t mid = ((lo + hi) / 2 + 64) & ~(int64_t)63;
        if (mid > hi) break;
        if (okmi_backward_ws_bytes(b, nmo, mid) <= ws_bytes) lo = mid;
        else hi = mid - 64;
    }
    const BwdLayout Lo = bwd_layout(b, nmo, lo);
    const int64_t chunk = Lo.chunk;
    for (int s = 0; s < b->nshell; ++s)
        if (b->h_shells[s].nprim > 256 * MAXP) {
            okmi_set_error("rho_backward: shell %d has %d primitives (up to %d)", s, b->h_shells[s].nprim, 256 * MAXP);
            return 1;
        }
    int maxp = 1;
    for (int s = 0; s < b->nshell; ++s) maxp = std::max(maxp, b->h_shells[s].nprim);
    double* w = (double*)(((uintptr_t)ws + 255) & ~(uintptr_t)255);
    double* CT = w + Lo.ct;
    double* A = w + Lo.a;
    double* Phi = w + Lo.phi;
    double* B = w + Lo.bm;
    double* G = w + Lo.g;
    double* part = w + Lo.part;
    double* cols = w + Lo.cols;
    double* dcp = w + Lo.dcp;
    OkmiProfScope prof("rho_backward", st);
    hipLaunchKernelGGL(k_transpose, dim3((unsigned)(((int64_t)nmo * b->nao + 255) / 256)), dim3(256), 0, st, d_c, nmo,
                       b->nao, CT);
    const int set0 = 0;
    for (int64_t c0 = 0; c0 < np; c0 += chunk) {
        const int64_t n = np - c0 < chunk ? np - c0 : chunk;
        const int64_t n16 = (n + 15) & ~(int64_t)15;
        int rc = okmi_launch_ao_eval(b, g, p0 + c0, n, &set0, 1, OKMI_F64, A, chunk, 0, st);
        if (rc) return rc;
        if (n16 > n)
            hipLaunchKernelGGL(k_zero_cols, dim3((unsigned)((b->nao * (n16 - n) + 255) / 256)), dim3(256), 0, st, A,
                               b->nao, chunk, n, n16);
        rc = okmi_launch_gemm_f64(A, chunk, 0, 1, b->nao, d_c, nmo, n, Phi, chunk, st);
        if (rc) return rc;
        hipLaunchKernelGGL(k_bmat, dim3((unsigned)((n16 + 255) / 256)), dim3(256), 0, st, Phi, d_occ, d_grad + c0,
                           nmo, n, n16, chunk, B);
        // point ranges: about 2048 waves over the chip, at least 256 points
        // per range, at most the partial buffer's ranges
        const int64_t tiles = (b->nao + 16 * NT - 1) / (16 * NT);
        int64_t nsplit = std::max<int64_t>(1, std::min<int64_t>({(2048 + tiles - 1) / tiles, n16 / 256,
                                                                 (int64_t)Lo.nsplit_max}));
        int64_t krange = ((n16 + nsplit - 1) / nsplit + 15) & ~(int64_t)15;
        nsplit = (n16 + krange - 1) / krange;
        // blocks: (range group of 8, AO tile, XCD); range = group * 8 + XCD
        const unsigned nblk = (unsigned)(((nsplit + 7) / 8) * tiles * 8);
        const int nmt = (nmo + 15) / 16;
        {
            OkmiProfScope p2("gemm_tn", st);
            if (nmt == 1)
                hipLaunchKernelGGL(k_gemm_tn<1>, dim3(nblk), dim3(64), 0, st, B, nmo, A, b->nao, n16, chunk, krange,
                                   (int)nsplit, dcp);
            else if (nmt == 2)
                hipLaunchKernelGGL(k_gemm_tn<2>, dim3(nblk), dim3(64), 0, st, B, nmo, A, b->nao, n16, chunk, krange,
                                   (int)nsplit, dcp);
            else
                hipLaunchKernelGGL(k_gemm_tn<3>, dim3(nblk), dim3(64), 0, st, B, nmo, A, b->nao, n16, chunk, krange,
                                   (int)nsplit, dcp);
        }
        const int64_t ndc = (int64_t)nmo * b->nao;
        hipLaunchKernelGGL(k_reduce_dc, dim3((unsigned)((ndc + 255) / 256)), dim3(256), 0, st, dcp, (int)nsplit, ndc,
                           d_dc);
        rc = okmi_launch_gemm_f64(B, chunk, 0, 1, nmo, CT, b->nao, n, G, chunk, st);
        if (rc) return rc;
        ShellGradArgs sa{};
        sa.shells = b->d_shells; sa.prims = b->d_prims; sa.T = b->d_T; sa.grid = g;
        sa.p0 = p0 + c0; sa.np = n; sa.ld = chunk; sa.nprim_total = b->nprim; sa.is_normalized = b->is_normalized;
        sa.G = G; sa.partial = part; sa.ncol = Lo.ncol;
        const int nrange = (int)((n + RANGE - 1) / RANGE);
        {
            OkmiProfScope p3("shell_grad", st);
            if (maxp <= 4)
                hipLaunchKernelGGL(k_shell_grad<4>, dim3((unsigned)nrange, (unsigned)b->nshell, 1u), dim3(256), 0, st,
                                   sa);
            else
                hipLaunchKernelGGL(k_shell_grad<MAXP>,
                                   dim3((unsigned)nrange, (unsigned)b->nshell, (unsigned)((maxp + MAXP - 1) / MAXP)),
                                   dim3(256), 0, st, sa);
        }
        hipLaunchKernelGGL(k_reduce_cols, dim3((unsigned)((Lo.ncol + 255) / 256)), dim3(256), 0, st, part, nrange,
                           Lo.ncol, cols);
        hipLaunchKernelGGL(k_scatter_grad, dim3((unsigned)((b->nprim + 3 * b->natoms + 255) / 256)), dim3(256), 0, st,
                           cols, b->d_shells, b->nshell, b->nprim, b->natoms, d_dalpha, d_dr);
    }
    return okmi_check_launch("rho_backward");
}
