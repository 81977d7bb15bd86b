// Fused density kernel (AO evaluation -> LDS -> MFMA contraction -> density
// terms) for the FP32 and bf16-split precisions, the gather of orbkit's output
// rows, the host-side tables and launch sequence for every precision, and
// the many-molecule batch kernel.  The tile body is in fused_tile.h; the FP64
// instances are compiled in rho_fused_f64.hip.
#include "fused_tile.h"

namespace {

using namespace okmi_fused;

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

}  // namespace

namespace {
size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// bytes per entry of the MO coefficient operand: FP64, FP32, two or three
// bf16 planes
size_t c_elem_bytes(int prec)
{
    return prec == OKMI_PREC_F64 ? 8 : prec == OKMI_PREC_F32 ? 4 : prec == OKMI_PREC_BF16X3 ? 4 : 6;
}

template <typename T>
int upload(T** d, const std::vector<T>& h, const char* what)
{
    hipError_t e = hipMalloc((void**)d, sizeof(T) * std::max<size_t>(1, h.size()));
    if (e != hipSuccess) return okmi_hip_fail(e, what);
    e = hipMemcpy(*d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) return okmi_hip_fail(e, what);
    return 0;
}

// Internal row layout of the fused kernels: every 32-row chunk holds whole
// shells, packed from its first row.  A tile skips the contraction of every
// chunk none of whose shells reaches it, so shells that reach the same tiles
// should share chunks.  Shells are taken by screening radius class (powers
// of two, largest first), within a class by the Morton code of their centre
// (neighbouring atoms next to each other), then by radius, and each is placed
// in the first chunk with room for it (first fit), so the chunks stay nearly
// full (every padding row still costs MFMA work).  The diffuse shells, which
// reach every tile, share the first chunks; the tight ones share chunks
// with their spatial neighbours.  Live chunks per 4x4x4 tile: 20 atoms x
// [s,s,p,p,d,f] (spherical, 400 rows in 13 chunks, exponents 0.05 .. 1e5)
// on a 100^3 grid 6.8 instead of 11.4 in basis order; 200 atoms (2,950 rows
// in 93 chunks) on a 128^3 grid 18.7 instead of 42.3.  Returns the padded
// row count (-1 if a shell has more than 32 rows) and each shell's first
// internal row.
int fused_layout(const OkmiBasis* b, std::vector<int>* krow)
{
    std::vector<int> fill;   // rows used per chunk
    if (krow) krow->assign(b->nshell, 0);
    const int ns = b->nshell;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int s = 0; s < ns; ++s) {
        const double c[3] = {b->h_shells[s].cx, b->h_shells[s].cy, b->h_shells[s].cz};
        for (int d = 0; d < 3; ++d) {
            lo[d] = std::min(lo[d], c[d]);
            hi[d] = std::max(hi[d], c[d]);
        }
    }
    std::vector<uint32_t> morton(ns);
    std::vector<int> cls(ns);
    for (int s = 0; s < ns; ++s) {
        const double c[3] = {b->h_shells[s].cx, b->h_shells[s].cy, b->h_shells[s].cz};
        uint32_t q[3];
        for (int d = 0; d < 3; ++d)
            q[d] = hi[d] > lo[d] ? (uint32_t)std::min(1023.0, (c[d] - lo[d]) / (hi[d] - lo[d]) * 1023.0) : 0u;
        uint32_t m = 0;
        for (int bit = 0; bit < 10; ++bit)
            for (int d = 0; d < 3; ++d) m |= ((q[d] >> bit) & 1u) << (3 * bit + d);
        morton[s] = m;
        cls[s] = (int)std::floor(0.5 * std::log2(std::max(b->h_frcut2[s], 1e-300)));
    }
    std::vector<int> order(ns);
    for (int s = 0; s < ns; ++s) order[s] = s;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        if (cls[x] != cls[y]) return cls[x] > cls[y];
        if (morton[x] != morton[y]) return morton[x] < morton[y];
        return b->h_frcut2[x] > b->h_frcut2[y];
    });
    for (int s : order) {
        const int n = b->h_shells[s].nrow;
        if (n > KC) return -1;
        size_t c = 0;
        while (c < fill.size() && fill[c] + n > KC) ++c;
        if (c == fill.size()) fill.push_back(0);
        if (krow) (*krow)[s] = (int)c * KC + fill[c];
        fill[c] += n;
    }
    return (int)fill.size() * KC;
}

size_t fused_fixed_bytes(const OkmiBasis* b, int nmo, int prec)
{
    const size_t mpad = ((nmo + 15) / 16) * 16, kpad = (size_t)std::max(KC, fused_layout(b, nullptr));
    const size_t elem = c_elem_bytes(prec);
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
    std::vector<std::vector<std::pair<int, int>>> cost(nchunk);   // per chunk: (cost, shell)
    std::vector<int> used(nchunk, 0);                              // rows used per chunk
    for (int s = 0; s < b->nshell; ++s) {
        const ShellRec& sh = b->h_shells[s];
        const int c = krow[s] / KC;
        cost[c].push_back({16 * sh.nprim + 8 * (ncart(sh.l) + sh.nrow), s});
        used[c] = std::max(used[c], krow[s] + sh.nrow - c * KC);
    }
    for (int c = 0; c < nchunk; ++c) {
        std::stable_sort(cost[c].begin(), cost[c].end(), [](auto& x, auto& y) { return x.first > y.first; });
        ch[c] = make_int2((int)work.size(), (int)cost[c].size() | (used[c] << 16));
        for (auto& cs : cost[c]) work.push_back(cs.second);
    }
    // shells in work-list order (entry j of the list is record j), each
    // with its primitives in groups of 4 (prim0: first group); one set of
    // tables per screening bound (radii frc2 / fprc2)
    std::vector<int> colmap(nkpad, -1);
    for (size_t j = 0; j < work.size(); ++j)
        for (int r = 0; r < b->h_shells[work[j]].nrow; ++r) colmap[krow[work[j]] + r] = b->h_shells[work[j]].row0 + r;
    auto tables = [&](const std::vector<double>& frc2, const std::vector<double>& fprc2, std::vector<ShellRec>& fs,
                      std::vector<PrimGroup>& groups, std::vector<ShellRec>& is, std::vector<ShellRec>& wsh,
                      std::vector<PrimRec>& fp) {
        fs.assign(work.size(), ShellRec{});
        groups.clear();
        for (size_t j = 0; j < work.size(); ++j) {
            const int i = work[j];
            fs[j] = b->h_shells[i];
            fs[j].row0 = krow[i];
            fs[j].rcut2 = frc2[i];
            fs[j].prim0 = (int)groups.size();
            const ShellRec& hs = b->h_shells[i];
            // primitives by screening radius, largest first (ties keep file order)
            std::vector<int> order(hs.nprim);
            for (int k = 0; k < hs.nprim; ++k) order[k] = hs.prim0 + k;
            std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return fprc2[x] > fprc2[y]; });
            for (int q = 0; q < hs.nprim; q += 4) {
                PrimGroup pg;
                for (int k = 0; k < 4; ++k) {
                    const bool used = q + k < hs.nprim;
                    const int pi = used ? order[q + k] : 0;
                    pg.alpha[k] = used ? b->h_prims[pi].alpha : 0.0;
                    pg.d[k] = used ? b->h_prims[pi].d : 0.0;
                    pg.rcut2[k] = used ? fprc2[pi] : -1.0;
                }
                groups.push_back(pg);
            }
        }
        // the same shells writing their internal rows (the many-MO and
        // staged paths stage the AO sets in the fused row order), also in
        // work order, and the primitives with these radii
        is.assign(b->h_shells, b->h_shells + b->nshell);
        for (int i = 0; i < b->nshell; ++i) {
            is[i].rcut2 = frc2[i];
            is[i].row0 = krow[i];
        }
        wsh.resize(work.size());
        for (size_t j = 0; j < work.size(); ++j) wsh[j] = is[work[j]];
        fp.assign(b->h_prims, b->h_prims + b->nprim);
        for (int i = 0; i < b->nprim; ++i) fp[i].rcut2 = fprc2[i];
    };
    std::vector<ShellRec> fs, is, wsh, fs_s, is_s, wsh_s;
    std::vector<PrimGroup> groups, groups_s;
    std::vector<PrimRec> fp, fp_s;
    tables(b->h_frcut2, b->h_fprcut2, fs, groups, is, wsh, fp);
    tables(b->h_frcut2s, b->h_fprcut2s, fs_s, groups_s, is_s, wsh_s, fp_s);
    std::vector<ShellRec> os(b->h_shells, b->h_shells + b->nshell);
    for (int i = 0; i < b->nshell; ++i) os[i].rcut2 = b->h_frcut2[i];
    std::vector<int> padrows;
    for (int r = 0; r < nkpad; ++r)
        if (colmap[r] < 0) padrows.push_back(r);
    int rc = upload(&b->d_chunks, ch, "rho_fused: chunk map");
    if (!rc) rc = upload(&b->d_fshells, fs, "rho_fused: shells");
    if (!rc) rc = upload(&b->d_fprims, fp, "rho_fused: primitives");
    if (!rc) rc = upload(&b->d_fgroups, groups, "rho_fused: primitive groups");
    if (!rc) rc = upload(&b->d_oshells, os, "rho_fused: output-row shells");
    if (!rc) rc = upload(&b->d_ishells, is, "rho_fused: internal-row shells");
    if (!rc) rc = upload(&b->d_wshells, wsh, "rho_fused: work-order shells");
    if (!rc) rc = upload(&b->d_padrows, padrows, "rho_fused: padding rows");
    if (!rc) rc = upload(&b->d_colmap, colmap, "rho_fused: row map");
    if (!rc) rc = upload(&b->d_fshells_s, fs_s, "rho_fused: shells (FP32 bound)");
    if (!rc) rc = upload(&b->d_fgroups_s, groups_s, "rho_fused: primitive groups (FP32 bound)");
    if (!rc) rc = upload(&b->d_ishells_s, is_s, "rho_fused: internal-row shells (FP32 bound)");
    if (!rc) rc = upload(&b->d_wshells_s, wsh_s, "rho_fused: work-order shells (FP32 bound)");
    if (!rc) rc = upload(&b->d_fprims_s, fp_s, "rho_fused: primitives (FP32 bound)");
    if (rc) return rc;
    b->nchunk = nchunk;
    b->nkpad = nkpad;
    b->npadrows = (int)padrows.size();
    b->h_chunks = ch;
    b->h_fshells = fs;
    b->h_fshells_s = fs_s;
    return 0;
}
}  // namespace

// Per point: up to 20 FP64 term rows; FP64 with MO norms also one row of
// per-tile norms per 64 points (at most 64 MOs: 8 bytes per point)
constexpr size_t FUSED_PT_BYTES = 20 * 8, NORM_PT_BYTES = 8;

size_t okmi_fused_ws_bytes(const OkmiBasis* b, int nmo, int prec, int64_t np)
{
    if (nmo > 64) return okmi_big_ws_bytes(b, nmo, prec, 7, np);
    size_t n = fused_fixed_bytes(b, nmo, prec) + align256(FUSED_PT_BYTES * np) + 256;
    if (prec == OKMI_PREC_F64) n += align256(NORM_PT_BYTES * (np + 64)) + 256;
    return n;
}

int okmi_launch_rho_fused(OkmiBasis* b, const GridDesc& g, int64_t p0, int64_t np, const double* d_c,
                          const double* d_occ, int nmo, int prec, const int* drv_out, int nout, double* rho,
                          double* delta, int64_t ld_delta, double* mo_norm, void* ws, size_t ws_bytes,
                          hipStream_t st)
{
    if (np <= 0) return 0;
    if (mo_norm && prec != OKMI_PREC_F64) {
        okmi_set_error("rho_fused: MO norms are computed by the FP64 kernel only");
        return 1;
    }
    if (prec != OKMI_PREC_F64 && prec != OKMI_PREC_F32 && prec != OKMI_PREC_BF16X3 && prec != OKMI_PREC_BF16X6) {
        okmi_set_error("rho_fused: unknown precision %d", prec);
        return 1;
    }
    const int mt = (nmo + 15) / 16;
    if (mt > 4 && prec == OKMI_PREC_F64) {
        okmi_set_error("rho_fused: FP64 keeps up to 64 MOs in registers (%d given; okmi_rho has no limit)", nmo);
        return 1;
    }
    if (mt > 4)   // more MOs than the fused kernel keeps in registers
        return okmi_launch_rho_big(b, g, p0, np, d_c, d_occ, nmo, prec, drv_out, nout, rho, delta, ld_delta, ws,
                                   ws_bytes, st);
    if (nout > 16) {
        okmi_set_error("rho_fused: %d derivative outputs (up to 16)", nout);
        return 1;
    }
    int rc = prepare_fused_tables(b);
    if (rc) return rc;
    // MOs padded to 16 (the FP32-accumulating MFMA tiles) or, for FP64, to 4
    // (16-MO tiles plus 4-MO groups)
    const int64_t mpad = prec == OKMI_PREC_F64 ? (nmo + 3) / 4 * 4 : mt * 16, kpad = b->nkpad;
    const size_t fixed = fused_fixed_bytes(b, nmo, prec);
    // per point: the term rows, plus the per-tile MO norms (mpad doubles per
    // 64-point tile) when wanted
    const size_t per_pt = FUSED_PT_BYTES + (mo_norm ? NORM_PT_BYTES : 0);
    const size_t slack = 256 + (mo_norm ? 256 + (size_t)mpad * 8 : 0);
    if (!ws || ws_bytes < fixed + slack + per_pt * 64) {
        okmi_set_error("rho_fused: workspace of %zu bytes is too small (need at least %zu)", ws_bytes,
                       fixed + slack + per_pt * 64);
        return 1;
    }
    const size_t avail = ws_bytes - fixed - slack;
    // a whole regular grid whose buffers fit is tiled in 4x4x4 bricks (the
    // faces may add tiles beyond np / 64)
    const int64_t nbrick = ((g.nx + 3) / 4) * ((g.ny + 3) / 4) * ((g.nz + 3) / 4);
    const bool brick = g.regular && p0 == 0 && np == g.nx * g.ny * g.nz &&
                       FUSED_PT_BYTES * (size_t)np + (mo_norm ? (size_t)nbrick * mpad * 8 : 0) <= avail;
    int64_t chunk = brick ? np : std::min<int64_t>(np, (int64_t)(avail / per_pt) & ~(int64_t)63);
    if (chunk <= 0) chunk = np;   // unreachable: avail holds 64 points
    char* w = (char*)(((uintptr_t)ws + 255) & ~(uintptr_t)255);
    void* cpl = w;
    double* occp = (double*)(w + align256(c_elem_bytes(prec) * mpad * kpad));
    double* term = (double*)(w + fixed);
    double* term2 = term + 10 * chunk;
    double* norm_part = mo_norm ? (double*)(w + fixed + align256(FUSED_PT_BYTES * (size_t)chunk)) : nullptr;

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
    // FP64 screens with the FP64 bound, the paths ending in FP32 with theirs
    const bool split = prec != OKMI_PREC_F64;
    a.shells = split ? b->d_fshells_s : b->d_fshells;
    a.groups = split ? b->d_fgroups_s : b->d_fgroups;
    a.T = b->d_T; a.chunks = b->d_chunks;
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
        auto pass = [&](int ps) {
            switch (prec) {
            case OKMI_PREC_F64: return launch_pass_f64(a, ps, (int)mpad, lm, st);
            case OKMI_PREC_F32: return launch_pass<OKMI_PREC_F32>(a, ps, mt, lm, st);
            case OKMI_PREC_BF16X3: return launch_pass<OKMI_PREC_BF16X3>(a, ps, mt, lm, st);
            default: return launch_pass<OKMI_PREC_BF16X6>(a, ps, mt, lm, st);
            }
        };
        a.norm_part = norm_part;   // read by the first pass (rho) only
        int r = pass(p1 ? 1 : 0);
        if (!r && p2) r = pass(2);
        if (!r && p3) r = pass(3);
        if (!r && mo_norm) r = launch_norm_reduce(norm_part, brick ? nbrick : (n + PTS - 1) / PTS, (int)mpad, nmo,
                                                  mo_norm, st);
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
__global__ __launch_bounds__(256, (FusedPlan<PREC, PASS, LM, MT, true>::occupancy)) void k_rho_batch(BatchArgs b)
{
    using P = FusedPlan<PREC, PASS, LM, MT, true>;
    __shared__ typename Tile<PREC>::T lds[P::buffers * FusedLds<PREC, PASS>::BUF];
    __shared__ bf8 cst[P::single && CStage<PREC, MT>::on ? CStage<PREC, MT>::slots : 1];
    __shared__ int next[2];
    __shared__ int live[2];
    __shared__ unsigned long long lmask[LIVE_WORDS];
    for (int64_t t = blockIdx.x; t < b.ntile; t += gridDim.x) {
        const int m = find_item(b.tile_start, b.nmol, t);
        const FusedArgs a = load_fused_args(b.mols, m);
        rho_tile<PREC, PASS, MT, LM, true>(a, t - cptr(b.tile_start)[m], lds, cst, next, live, lmask);
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
    const size_t elem = c_elem_bytes(prec);
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
        const bool split = B->prec != OKMI_PREC_F64;   // the screening bound of the arithmetic
        a.shells = split ? bs->d_fshells_s : bs->d_fshells;
        a.groups = split ? bs->d_fgroups_s : bs->d_fgroups;
        a.T = bs->d_T;
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
        a.mt_live = (nmo[m] + 15) / 16;
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

extern "C" int okmi_fused_live_rows(const okmi_basis* b, const double* x, int64_t nx, const double* y, int64_t ny,
                                    const double* z, int64_t nz, int prec, double* rows_points)
{
    if (!b || !x || !y || !z || nx <= 0 || ny <= 0 || nz <= 0 || !rows_points) {
        okmi_set_error("fused_live_rows: invalid arguments");
        return 1;
    }
    if (b->h_chunks.empty()) {
        okmi_set_error("fused_live_rows: the basis has no fused tables");
        return 1;
    }
    // per axis and brick index: the span of the (up to 4) axis values
    auto spans = [](const double* a, int64_t n) {
        std::vector<std::pair<double, double>> s((size_t)((n + 3) / 4));
        for (int64_t i = 0; i < n; ++i) {
            auto& p = s[(size_t)(i / 4)];
            if (i % 4 == 0) p = {a[i], a[i]};
            p.first = std::min(p.first, a[i]);
            p.second = std::max(p.second, a[i]);
        }
        return s;
    };
    const auto sx = spans(x, nx), sy = spans(y, ny), sz = spans(z, nz);
    const int nch = (int)b->h_chunks.size();
    std::vector<int> used(nch);
    for (int c = 0; c < nch; ++c) used[c] = b->h_chunks[c].y >> 16;
    double total = 0.0;
    for (size_t i = 0; i < sx.size(); ++i)
        for (size_t j = 0; j < sy.size(); ++j)
            for (size_t k = 0; k < sz.size(); ++k) {
                const double lo[3] = {sx[i].first, sy[j].first, sz[k].first};
                const double hi[3] = {sx[i].second, sy[j].second, sz[k].second};
                const double pts = (double)(std::min<int64_t>(4, nx - 4 * (int64_t)i) *
                                            std::min<int64_t>(4, ny - 4 * (int64_t)j) *
                                            std::min<int64_t>(4, nz - 4 * (int64_t)k));
                int rows = 0;
                for (int c = 0; c < nch; ++c) {
                    const int2 cw = b->h_chunks[c];
                    const int nsh = cw.y & 0xffff;
                    for (int e = 0; e < nsh; ++e) {
                        const ShellRec& sh = (prec == OKMI_PREC_F64 ? b->h_fshells : b->h_fshells_s)[cw.x + e];
                        const double cc[3] = {sh.cx, sh.cy, sh.cz};
                        double d2 = 0.0;
                        for (int d = 0; d < 3; ++d) {
                            const double dd = std::max(std::max(lo[d] - cc[d], cc[d] - hi[d]), 0.0);
                            d2 += dd * dd;
                        }
                        if (d2 < sh.rcut2) {
                            rows += used[c];
                            break;
                        }
                    }
                }
                total += rows * pts;
            }
    *rows_points = total;
    return 0;
}
