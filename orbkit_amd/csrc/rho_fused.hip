// Fused density kernel: AO evaluation -> LDS -> MFMA contraction with the MO
// coefficients -> density terms, in one launch.  MOs never leave registers.
//
// One wave owns 64 grid points (lane = point while evaluating AOs).  The AO
// rows are processed in chunks of 32 (one K step of the bf16 MFMA); for each
// chunk the wave evaluates the shells overlapping it (FP64 coordinates and
// exponent range reduction, FP32 values), stores the rows in LDS [row][point],
// then multiplies them with the matching 32 columns of C:
//   F32     v_mfma_f32_16x16x4_f32, C and AO in FP32
//   BF16X3  v_mfma_f32_16x16x32_bf16 on (hi, lo) splits, 3 products
//   BF16X6  (hi, mid, lo) splits, 6 products (terms below 2^-24 dropped)
// accumulating in FP32.  The epilogue forms the density terms in FP64.
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

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

constexpr int KC = 32;        // AO rows per chunk
constexpr int LDS_LD = 65;    // padded row stride (floats) of the LDS AO tile
constexpr int PTS = 64;       // points per wave

struct FusedArgs {
    const ShellRec* shells;
    const PrimRec* prims;
    const float* T;           // FP32 copy of the transform table
    const int2* chunks;       // first/last shell overlapping each chunk
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
};

template <int PASS> struct PassCodes;
template <> struct PassCodes<0> { static constexpr int n = 1; static constexpr int c[4] = {0, 0, 0, 0}; };
template <> struct PassCodes<1> { static constexpr int n = 4; static constexpr int c[4] = {0, 1, 2, 3}; };
template <> struct PassCodes<2> { static constexpr int n = 4; static constexpr int c[4] = {0, 4, 5, 6}; };
template <> struct PassCodes<3> { static constexpr int n = 4; static constexpr int c[4] = {0, 7, 8, 9}; };

template <int L, int PASS>
__device__ __forceinline__ void eval_shell_chunk(const FusedArgs& a, const ShellRec& sh, double px, double py,
                                                 double pz, int k0, float* lds, int lane)
{
    constexpr int NS = PassCodes<PASS>::n;
    constexpr bool need1 = PASS == 1 || PASS == 2;
    constexpr bool need2 = PASS >= 2;
    const double Xd = px - sh.cx, Yd = py - sh.cy, Zd = pz - sh.cz;
    const double rr = Xd * Xd + Yd * Yd + Zd * Zd;
    float S0 = 0.f, S1 = 0.f, S2 = 0.f;
    for (int p = 0; p < sh.nprim; ++p) {
        const PrimRec pr = a.prims[sh.prim0 + p];
        const float de = (float)pr.d * gauss_f32(pr.alpha * rr);
        S0 += de;
        if constexpr (need1) S1 += (float)(-2.0 * pr.alpha) * de;
        if constexpr (need2) S2 += (float)(4.0 * pr.alpha * pr.alpha) * de;
    }
    ShellPows<L, float> P;
    P.init((float)Xd, (float)Yd, (float)Zd);
    constexpr int NC = ncart(L);
    const float* T = a.T + sh.t0;
    const int rlo = max(0, k0 - sh.row0), rhi = min(sh.nrow, k0 + KC - sh.row0);
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        float c[NC];
        switch (PassCodes<PASS>::c[s]) {
        case 0: cart_all<L, 0, float>(P, S0, S1, S2, c); break;
        case 1: cart_all<L, 1, float>(P, S0, S1, S2, c); break;
        case 2: cart_all<L, 2, float>(P, S0, S1, S2, c); break;
        case 3: cart_all<L, 3, float>(P, S0, S1, S2, c); break;
        case 4: cart_all<L, 4, float>(P, S0, S1, S2, c); break;
        case 5: cart_all<L, 5, float>(P, S0, S1, S2, c); break;
        case 6: cart_all<L, 6, float>(P, S0, S1, S2, c); break;
        case 7: cart_all<L, 7, float>(P, S0, S1, S2, c); break;
        case 8: cart_all<L, 8, float>(P, S0, S1, S2, c); break;
        default: cart_all<L, 9, float>(P, S0, S1, S2, c); break;
        }
        float* dst = lds + s * KC * LDS_LD + (sh.row0 - k0) * LDS_LD + lane;
        for (int r = rlo; r < rhi; ++r) {
            float v = 0.f;
#pragma unroll
            for (int k = 0; k < NC; ++k) v = fmaf(T[r * NC + k], c[k], v);
            dst[r * LDS_LD] = v;
        }
    }
}

template <int PREC>
struct Split {
    bf8 p[PREC == OKMI_PREC_BF16X6 ? 3 : 2];
};

template <int PREC>
__device__ __forceinline__ void split8(const float (&x)[8], Split<PREC>& o)
{
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const __bf16 h = (__bf16)x[j];
        const float r1 = x[j] - (float)h;
        const __bf16 m = (__bf16)r1;
        o.p[0][j] = h;
        o.p[1][j] = m;
        if constexpr (PREC == OKMI_PREC_BF16X6) {
            const float r2 = r1 - (float)m;
            o.p[2][j] = (__bf16)r2;
        }
    }
}

template <int PREC>
__device__ __forceinline__ f4 mfma_split(const Split<PREC>& A, const Split<PREC>& B, f4 acc)
{
    // smallest terms first
    if constexpr (PREC == OKMI_PREC_BF16X6) {
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A.p[2], B.p[0], acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A.p[1], B.p[1], acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A.p[0], B.p[2], acc, 0, 0, 0);
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A.p[1], B.p[0], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A.p[0], B.p[1], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A.p[0], B.p[0], acc, 0, 0, 0);
    return acc;
}

template <int PREC, int PASS, int MT>
__global__ __launch_bounds__(64) void k_rho_fused(FusedArgs a)
{
    constexpr int NS = PassCodes<PASS>::n;
    constexpr int NT = PTS / 16;
    __shared__ float lds[NS * KC * LDS_LD];
    const int lane = threadIdx.x;
    const int64_t tile0 = (int64_t)blockIdx.x * PTS;
    const int64_t pl = tile0 + lane;
    double px, py, pz;
    grid_point(a.grid, a.p0 + (pl < a.np ? pl : 0), px, py, pz);

    f4 acc[MT][NT][NS];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt)
#pragma unroll
            for (int s = 0; s < NS; ++s) acc[mt][nt][s] = f4{0.f, 0.f, 0.f, 0.f};

    const int g = lane >> 4, n16 = lane & 15;
    for (int ch = 0; ch < a.nchunk; ++ch) {
        const int k0 = ch * KC;
        // operand A (MO coefficients) for this chunk
        Split<PREC> ca[MT];
        float cf[MT][KC / 4];
        if constexpr (PREC == OKMI_PREC_F32) {
            const float* C = (const float*)a.C;
#pragma unroll
            for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                for (int ks = 0; ks < KC / 4; ++ks)
                    cf[mt][ks] = C[(int64_t)(mt * 16 + n16) * a.kpad + k0 + 4 * ks + g];
        } else {
            const __bf16* C = (const __bf16*)a.C;
            constexpr int NP = PREC == OKMI_PREC_BF16X6 ? 3 : 2;
#pragma unroll
            for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                for (int q = 0; q < NP; ++q)
                    ca[mt].p[q] = *(const bf8*)(C + ((int64_t)q * a.mpad + mt * 16 + n16) * a.kpad + k0 + 8 * g);
        }
        // AO rows [k0, k0+KC) for this wave's 64 points
        const int2 sr = a.chunks[ch];
        if (k0 + KC > a.nao) {
            for (int s = 0; s < NS; ++s)
                for (int r = a.nao - k0; r < KC; ++r) lds[s * KC * LDS_LD + r * LDS_LD + lane] = 0.f;
        }
        for (int si = sr.x; si <= sr.y; ++si) {
            const ShellRec sh = a.shells[si];
            switch (sh.l) {
            case 0: eval_shell_chunk<0, PASS>(a, sh, px, py, pz, k0, lds, lane); break;
            case 1: eval_shell_chunk<1, PASS>(a, sh, px, py, pz, k0, lds, lane); break;
            case 2: eval_shell_chunk<2, PASS>(a, sh, px, py, pz, k0, lds, lane); break;
            case 3: eval_shell_chunk<3, PASS>(a, sh, px, py, pz, k0, lds, lane); break;
            default: eval_shell_chunk<4, PASS>(a, sh, px, py, pz, k0, lds, lane); break;
            }
        }
        __syncthreads();   // the block is one wave: orders the LDS writes before the reads
        // contraction
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                const float* src = lds + s * KC * LDS_LD + nt * 16 + n16;
                if constexpr (PREC == OKMI_PREC_F32) {
#pragma unroll
                    for (int ks = 0; ks < KC / 4; ++ks) {
                        const float b = src[(4 * ks + g) * LDS_LD];
#pragma unroll
                        for (int mt = 0; mt < MT; ++mt)
                            acc[mt][nt][s] = __builtin_amdgcn_mfma_f32_16x16x4f32(cf[mt][ks], b, acc[mt][nt][s], 0, 0, 0);
                    }
                } else {
                    float bx[8];
#pragma unroll
                    for (int j = 0; j < 8; ++j) bx[j] = src[(8 * g + j) * LDS_LD];
                    Split<PREC> bs;
                    split8<PREC>(bx, bs);
#pragma unroll
                    for (int mt = 0; mt < MT; ++mt) acc[mt][nt][s] = mfma_split<PREC>(ca[mt], bs, acc[mt][nt][s]);
                }
            }
        }
        __syncthreads();
    }

    // epilogue: lane holds Phi[m = 16 mt + 4 g + i][point = 16 nt + n16]
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
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
                const double f0 = acc[mt][nt][0][i];
                if constexpr (PASS <= 1) r0 += w * (f0 * f0);
                if constexpr (NS > 1) {
#pragma unroll
                    for (int s = 1; s < NS; ++s) t[s - 1] += 2.0 * w * (f0 * (double)acc[mt][nt][s][i]);
                }
                if constexpr (PASS == 1) {
                    const double fx = acc[mt][nt][1][i], fy = acc[mt][nt][2][i], fz = acc[mt][nt][3][i];
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
        const int64_t p = tile0 + nt * 16 + n16;
        if (g == 0 && p < a.np) {
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

template <int PREC, int PASS>
int launch_mt(const FusedArgs& a, int mt, hipStream_t st)
{
    dim3 blk(PTS), grd((unsigned)((a.np + PTS - 1) / PTS));
    OkmiProfScope prof("rho_fused", st);
    switch (mt) {
    case 1: hipLaunchKernelGGL((k_rho_fused<PREC, PASS, 1>), grd, blk, 0, st, a); return 0;
    case 2: hipLaunchKernelGGL((k_rho_fused<PREC, PASS, 2>), grd, blk, 0, st, a); return 0;
    case 3: hipLaunchKernelGGL((k_rho_fused<PREC, PASS, 3>), grd, blk, 0, st, a); return 0;
    case 4: hipLaunchKernelGGL((k_rho_fused<PREC, PASS, 4>), grd, blk, 0, st, a); return 0;
    default: break;
    }
    okmi_set_error("rho_fused: %d MO tiles (up to 4 supported)", mt);
    return 1;
}

template <int PREC>
int launch_pass(const FusedArgs& a, int pass, int mt, hipStream_t st)
{
    switch (pass) {
    case 0: return launch_mt<PREC, 0>(a, mt, st);
    case 1: return launch_mt<PREC, 1>(a, mt, st);
    case 2: return launch_mt<PREC, 2>(a, mt, st);
    default: return launch_mt<PREC, 3>(a, mt, st);
    }
}

// C (FP64 [nmo][nao]) -> padded operand planes
__global__ void k_prep_c(const double* C, int nmo, int nao, int prec, int64_t mpad, int64_t kpad, void* out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= mpad * kpad) return;
    const int64_t m = i / kpad, k = i - m * kpad;
    const double v = (m < nmo && k < nao) ? C[m * nao + k] : 0.0;
    if (prec == OKMI_PREC_F32) {
        ((float*)out)[i] = (float)v;
    } else {
        __bf16* o = (__bf16*)out;
        const float x = (float)v;
        const __bf16 h = (__bf16)x;
        const float r1 = x - (float)h;
        const __bf16 md = (__bf16)r1;
        o[i] = h;
        o[mpad * kpad + i] = md;
        if (prec == OKMI_PREC_BF16X6) o[2 * mpad * kpad + i] = (__bf16)(r1 - (float)md);
    }
}

}  // namespace

namespace {
size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

size_t fused_fixed_bytes(const OkmiBasis* b, int nmo, int prec)
{
    const size_t mpad = ((nmo + 15) / 16) * 16, kpad = ((b->nao + KC - 1) / KC) * KC;
    const size_t elem = prec == OKMI_PREC_F32 ? 4 : prec == OKMI_PREC_BF16X3 ? 4 : 6;   // bytes per C entry
    return align256(elem * mpad * kpad) + align256(mpad * 8);
}

int prepare_fused_tables(OkmiBasis* b)
{
    if (b->d_T32) return 0;
    std::vector<float> t32(b->ntab);
    for (int i = 0; i < b->ntab; ++i) t32[i] = (float)b->h_T[i];
    const int nchunk = (b->nao + KC - 1) / KC;
    std::vector<int2> ch(nchunk);
    int s = 0;
    for (int c = 0; c < nchunk; ++c) {
        const int r0 = c * KC, r1 = std::min(b->nao, r0 + KC) - 1;
        while (s < b->nshell && b->h_shells[s].row0 + b->h_shells[s].nrow <= r0) ++s;
        int e = s;
        while (e + 1 < b->nshell && b->h_shells[e + 1].row0 <= r1) ++e;
        ch[c] = make_int2(s, e);
    }
    hipError_t e = hipMalloc((void**)&b->d_T32, sizeof(float) * std::max(1, b->ntab));
    if (e != hipSuccess) return okmi_hip_fail(e, "rho_fused: transform table");
    e = hipMemcpy(b->d_T32, t32.data(), sizeof(float) * b->ntab, hipMemcpyHostToDevice);
    if (e != hipSuccess) return okmi_hip_fail(e, "rho_fused: transform table");
    e = hipMalloc((void**)&b->d_chunks, sizeof(int2) * nchunk);
    if (e != hipSuccess) return okmi_hip_fail(e, "rho_fused: chunk map");
    e = hipMemcpy(b->d_chunks, ch.data(), sizeof(int2) * nchunk, hipMemcpyHostToDevice);
    if (e != hipSuccess) return okmi_hip_fail(e, "rho_fused: chunk map");
    b->nchunk = nchunk;
    return 0;
}
}  // namespace

size_t okmi_fused_ws_bytes(const OkmiBasis* b, int nmo, int prec, int64_t np)
{
    return fused_fixed_bytes(b, nmo, prec) + align256((size_t)20 * np * 8) + 256;
}

int okmi_launch_rho_fused(OkmiBasis* b, const GridDesc& g, int64_t p0, int64_t np, const double* d_c,
                          const double* d_occ, int nmo, int prec, const int* drv_out, int nout, double* rho,
                          double* delta, int64_t ld_delta, void* ws, size_t ws_bytes, hipStream_t st)
{
    if (np <= 0) return 0;
    const int mt = (nmo + 15) / 16;
    if (mt > 4) {
        okmi_set_error("rho_fused: %d MOs (the fused kernel holds up to 64)", nmo);
        return 1;
    }
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
    if (chunk >= np) chunk = np;
    else chunk &= ~(int64_t)63;

    const int64_t mpad = mt * 16, kpad = (int64_t)b->nchunk * KC;
    char* w = (char*)(((uintptr_t)ws + 255) & ~(uintptr_t)255);
    void* cpl = w;
    const size_t elem = prec == OKMI_PREC_F32 ? 4 : prec == OKMI_PREC_BF16X3 ? 4 : 6;
    double* occp = (double*)(w + align256(elem * mpad * kpad));
    double* term = (double*)(w + fixed);
    double* term2 = term + 10 * chunk;

    hipLaunchKernelGGL(k_prep_c, dim3((unsigned)((mpad * kpad + 255) / 256)), dim3(256), 0, st, d_c, nmo, b->nao,
                       prec, mpad, kpad, cpl);
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

    FusedArgs a{};
    a.shells = b->d_shells; a.prims = b->d_prims; a.T = b->d_T32; a.chunks = b->d_chunks;
    a.grid = g; a.nao = b->nao; a.nchunk = b->nchunk;
    a.C = cpl; a.kpad = kpad; a.mpad = mpad; a.occ = occp;
    a.term = term; a.term2 = term2; a.ld_term = chunk; a.need_mask = need;
    for (int64_t c0 = 0; c0 < np; c0 += chunk) {
        const int64_t n = np - c0 < chunk ? np - c0 : chunk;
        a.p0 = p0 + c0;
        a.np = n;
        a.rho = rho + c0;
        int r;
        if (prec == OKMI_PREC_F32) {
            r = launch_pass<OKMI_PREC_F32>(a, p1 ? 1 : 0, mt, st);
            if (!r && p2) r = launch_pass<OKMI_PREC_F32>(a, 2, mt, st);
            if (!r && p3) r = launch_pass<OKMI_PREC_F32>(a, 3, mt, st);
        } else if (prec == OKMI_PREC_BF16X3) {
            r = launch_pass<OKMI_PREC_BF16X3>(a, p1 ? 1 : 0, mt, st);
            if (!r && p2) r = launch_pass<OKMI_PREC_BF16X3>(a, 2, mt, st);
            if (!r && p3) r = launch_pass<OKMI_PREC_BF16X3>(a, 3, mt, st);
        } else {
            r = launch_pass<OKMI_PREC_BF16X6>(a, p1 ? 1 : 0, mt, st);
            if (!r && p2) r = launch_pass<OKMI_PREC_BF16X6>(a, 2, mt, st);
            if (!r && p3) r = launch_pass<OKMI_PREC_BF16X6>(a, 3, mt, st);
        }
        if (r) return r;
        if (nout > 0)
            hipLaunchKernelGGL(k_gather, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, rho + c0, term, term2,
                               chunk, gc, delta + c0, ld_delta, n);
    }
    return okmi_check_launch("rho_fused");
}
