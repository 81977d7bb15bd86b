// extern "C" entry points of liborbkit_amd.so (declared in include/orbkit_amd.h).
#include "okmi_internal.h"
#include "ao_device.h"

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace {
thread_local std::string g_err;

int dfact(int n)
{
    int r = 1;
    for (; n > 0; n -= 2) r *= n;
    return r;
}

// Smallest R (conservative) beyond which every primitive term of a shell,
// with the polynomial and up to second-derivative factors, is below `bound`:
// max_p |d_p| (2l+2) (1+r)^(l+2) (1+2 a_p r)^2 exp(-a_p r^2) < bound.
double shell_rcut(int l, const double* alpha, const double* d, int n, double bound)
{
    const double cut = std::log(bound);
    auto logf = [&](double r) {
        double m = -1e300;
        for (int p = 0; p < n; ++p) {
            if (d[p] == 0.0) continue;
            const double v = std::log(std::fabs(d[p])) + std::log(2.0 * l + 2.0) + (l + 2) * std::log1p(r) +
                             2.0 * std::log1p(2.0 * alpha[p] * r) - alpha[p] * r * r;
            m = std::max(m, v);
        }
        return m;
    };
    double lo = 0.0;
    for (int p = 0; p < n; ++p) lo = std::max(lo, std::sqrt((l + 4.0) / (2.0 * alpha[p])) + 1.0);
    if (logf(lo) < cut) return lo;
    double hi = std::max(2.0 * lo, 1e4);
    if (logf(hi) >= cut) return 1e300;
    for (int it = 0; it < 200 && hi - lo > 1e-9 * hi; ++it) {
        const double mid = 0.5 * (lo + hi);
        (logf(mid) < cut ? hi : lo) = mid;
    }
    return hi;
}

// T block kind: standard spherical rows (sph_row table), standard Cartesian
// (identity with the component norm factors) or general.
int shell_kind(int l, int nrow, const double* blk, bool spherical)
{
    const int NC = ncart(l);
    auto close = [](double a, double b) { return std::fabs(a - b) <= 1e-14 * std::max(1.0, std::fabs(b)); };
    if (spherical) {
        if (l > 3 || nrow != 2 * l + 1) return SHELL_T_GENERAL;
        for (int r = 0; r < nrow; ++r) {
            double ref[15] = {0};
            const SphRow row = sph_row(l, r);
            for (int t = 0; t < row.n; ++t) ref[row.t[t].c] += row.t[t].v;
            for (int c = 0; c < NC; ++c)
                if (!close(blk[r * NC + c], ref[c])) return SHELL_T_GENERAL;
        }
        return SHELL_T_SPHERICAL;
    }
    if (nrow != NC) return SHELL_T_GENERAL;
    for (int r = 0; r < nrow; ++r)
        for (int c = 0; c < NC; ++c)
            if (!close(blk[r * NC + c], r == c ? cart_norm(l, c) : 0.0)) return SHELL_T_GENERAL;
    return SHELL_T_CARTESIAN;
}

int canon_index(int l, int lx, int ly, int lz)
{
    for (int c = 0; c < ncart(l); ++c)
        if (cart_exp(l, c, 0) == lx && cart_exp(l, c, 1) == ly && cart_exp(l, c, 2) == lz) return c;
    return -1;
}

template <typename T>
int dev_upload(T** dst, const T* src, size_t n, const char* what)
{
    *dst = nullptr;
    if (n == 0) return 0;
    hipError_t e = hipMalloc((void**)dst, n * sizeof(T));
    if (e != hipSuccess) return okmi_hip_fail(e, what);
    e = hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice);
    if (e != hipSuccess) return okmi_hip_fail(e, what);
    return 0;
}

GridDesc to_desc(const okmi_grid* g)
{
    GridDesc d{};
    d.x = g->d_x; d.y = g->d_y; d.z = g->d_z;
    d.nx = g->nx; d.ny = g->ny; d.nz = g->nz;
    d.regular = g->regular;
    return d;
}

int check_grid(const okmi_grid* g, int64_t p0, int64_t np)
{
    if (!g || !g->d_x || !g->d_y || !g->d_z) {
        okmi_set_error("grid: null coordinate pointer");
        return 1;
    }
    int64_t total = g->regular ? g->nx * g->ny * g->nz : g->nx;
    if (g->nx < 0 || (g->regular && (g->ny <= 0 || g->nz <= 0))) {
        okmi_set_error("grid: invalid axis lengths");
        return 1;
    }
    if (p0 < 0 || np < 0 || p0 + np > total) {
        okmi_set_error("grid: point range [%lld, %lld) outside the %lld grid points", (long long)p0,
                       (long long)(p0 + np), (long long)total);
        return 1;
    }
    return 0;
}

int check_drv(const int* drv, int ndrv)
{
    if (ndrv < 1 || ndrv > OKMI_MAX_SETS || !drv) {
        okmi_set_error("drv: %d derivative sets requested (1..%d supported)", ndrv, OKMI_MAX_SETS);
        return 1;
    }
    for (int j = 0; j < ndrv; ++j)
        if (drv[j] < 0 || drv[j] > 9) {
            okmi_set_error("The selection `drv=%d` is not valid!", drv[j]);
            return 1;
        }
    return 0;
}
}  // namespace

void okmi_set_error(const char* fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
}

int okmi_hip_fail(hipError_t e, const char* what)
{
    okmi_set_error("%s: %s", what, hipGetErrorString(e));
    return 2;
}

int okmi_check_launch(const char* what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return okmi_hip_fail(e, what);
    return 0;
}

extern "C" const char* okmi_last_error(void) { return g_err.c_str(); }

namespace {
struct ProfEvent {
    std::string name;
    hipEvent_t start, stop;
};
bool g_prof = false;
std::vector<ProfEvent> g_prof_events;
thread_local bool g_no_alloc = false;   // inside okmi_rho_noalloc: no event creation either
}  // namespace

OkmiProfScope::OkmiProfScope(const char* name, hipStream_t s) : st(s)
{
    if (!g_prof || g_no_alloc) return;
    ProfEvent e{name, nullptr, nullptr};
    if (hipEventCreate(&e.start) != hipSuccess || hipEventCreate(&e.stop) != hipSuccess) return;
    (void)hipEventRecord(e.start, st);
    idx = (int)g_prof_events.size();
    g_prof_events.push_back(e);
}

OkmiProfScope::~OkmiProfScope()
{
    if (idx >= 0) (void)hipEventRecord(g_prof_events[idx].stop, st);
}

extern "C" int okmi_profile_enable(int on)
{
    for (auto& e : g_prof_events) {
        (void)hipEventSynchronize(e.stop);
        (void)hipEventDestroy(e.start);
        (void)hipEventDestroy(e.stop);
    }
    g_prof_events.clear();
    g_prof = on != 0;
    return 0;
}

extern "C" int okmi_profile_query(const char* name, int* count, double* total_ms)
{
    int n = 0;
    double t = 0.0;
    for (auto& e : g_prof_events) {
        if (e.name != name) continue;
        hipError_t r = hipEventSynchronize(e.stop);
        if (r != hipSuccess) return okmi_hip_fail(r, "profile_query");
        float ms = 0.f;
        r = hipEventElapsedTime(&ms, e.start, e.stop);
        if (r != hipSuccess) return okmi_hip_fail(r, "profile_query");
        t += ms;
        ++n;
    }
    if (count) *count = n;
    if (total_ms) *total_ms = t;
    return 0;
}

extern "C" int okmi_basis_create(int ncont, const int* ncart_per_cont, const int* lxlylz, const int* nprim_per_cont,
                                 const double* prim_coeffs, int natoms, const double* geo, const int* cont_atom,
                                 int is_normalized, const double* cart_scale, const int* nout_per_cont,
                                 const double* transform, okmi_basis** out)
{
    g_err.clear();
    if (!out || ncont <= 0 || !ncart_per_cont || !lxlylz || !nprim_per_cont || !prim_coeffs || !geo || !cont_atom) {
        okmi_set_error("basis_create: missing input");
        return 1;
    }
    *out = nullptr;
    std::vector<ShellRec> shells(ncont);
    std::vector<PrimRec> prims;
    std::vector<double> T, craw, frc2, fprc2, frc2s, fprc2s;
    std::vector<int> pl;
    const double pref = std::pow(2.0 / M_PI, 0.75);
    int crow = 0, prow = 0, orow = 0, tmat = 0, max_l = 0;
    for (int c = 0; c < ncont; ++c) {
        const int nc = ncart_per_cont[c];
        const int np = nprim_per_cont[c];
        if (nc <= 0 || np <= 0) {
            okmi_set_error("basis_create: contraction %d has %d components and %d primitives", c, nc, np);
            return 1;
        }
        const int* lc = lxlylz + 3 * crow;
        const int l = lc[0] + lc[1] + lc[2];
        if (l > OKMI_MAX_L) {
            okmi_set_error("basis_create: angular momentum %d in contraction %d (up to %d supported)", l, c,
                           OKMI_MAX_L);
            return 1;
        }
        std::vector<int> ci(nc);
        std::vector<double> colf(nc);
        for (int j = 0; j < nc; ++j) {
            const int* e = lc + 3 * j;
            if (e[0] < 0 || e[1] < 0 || e[2] < 0 || e[0] + e[1] + e[2] != l) {
                okmi_set_error("basis_create: contraction %d mixes angular momenta (component %d)", c, j);
                return 1;
            }
            ci[j] = canon_index(l, e[0], e[1], e[2]);
            double f = is_normalized ? 1.0
                                     : 1.0 / std::sqrt((double)(dfact(2 * e[0] - 1) * dfact(2 * e[1] - 1) *
                                                                dfact(2 * e[2] - 1)));
            colf[j] = f * (cart_scale ? cart_scale[crow + j] : 1.0);
        }
        const int NC = ncart(l);
        const int nrow = nout_per_cont ? nout_per_cont[c] : nc;
        if (nrow <= 0) {
            okmi_set_error("basis_create: contraction %d has no output rows", c);
            return 1;
        }
        std::vector<double> blk((size_t)nrow * NC, 0.0);
        if (!nout_per_cont) {
            for (int j = 0; j < nc; ++j) blk[(size_t)j * NC + ci[j]] += colf[j];
        } else {
            for (int r = 0; r < nrow; ++r)
                for (int j = 0; j < nc; ++j) blk[(size_t)r * NC + ci[j]] += transform[tmat + (size_t)r * nc + j] * colf[j];
            tmat += nrow * nc;
        }
        const int at = cont_atom[c];
        if (at < 0 || at >= natoms) {
            okmi_set_error("basis_create: contraction %d refers to atom %d of %d", c, at, natoms);
            return 1;
        }
        ShellRec& s = shells[c];
        s.cx = geo[3 * at]; s.cy = geo[3 * at + 1]; s.cz = geo[3 * at + 2];
        s.l = l; s.nprim = np; s.prim0 = prow; s.row0 = orow; s.nrow = nrow;
        s.t0 = (int)T.size(); s.atom = at;
        s.kind = shell_kind(l, nrow, blk.data(), nout_per_cont != nullptr);
        T.insert(T.end(), blk.begin(), blk.end());
        std::vector<double> al(np), dd(np);
        for (int p = 0; p < np; ++p) {
            const double a = prim_coeffs[2 * (prow + p)], cc = prim_coeffs[2 * (prow + p) + 1];
            const double nh = is_normalized ? 1.0 : pref * std::pow(2.0, l) * std::pow(a, (2.0 * l + 3.0) / 4.0);
            prims.push_back(PrimRec{a, cc * nh, 0.0});
            craw.push_back(cc);
            pl.push_back(l);
            al[p] = a;
            // the largest T entry bounds the output rows' magnitude
            double tmax = 0.0;
            for (double t : blk) tmax = std::max(tmax, std::fabs(t));
            dd[p] = cc * nh * std::max(1.0, tmax) * nc;
        }
        const double rc = shell_rcut(l, al.data(), dd.data(), np, OKMI_SCREEN);
        s.rcut2 = rc >= 1e150 ? 1e300 : rc * rc;
        const double rf = shell_rcut(l, al.data(), dd.data(), np, OKMI_SCREEN_FUSED);
        frc2.push_back(rf >= 1e150 ? 1e300 : rf * rf);
        const double rs = shell_rcut(l, al.data(), dd.data(), np, OKMI_SCREEN_SPLIT);
        frc2s.push_back(rs >= 1e150 ? 1e300 : rs * rs);
        for (int p = 0; p < np; ++p) {
            const double rp = shell_rcut(l, &al[p], &dd[p], 1, OKMI_SCREEN);
            prims[prims.size() - np + p].rcut2 = rp >= 1e150 ? 1e300 : rp * rp;
            const double rq = shell_rcut(l, &al[p], &dd[p], 1, OKMI_SCREEN_FUSED);
            fprc2.push_back(rq >= 1e150 ? 1e300 : rq * rq);
            const double rr = shell_rcut(l, &al[p], &dd[p], 1, OKMI_SCREEN_SPLIT);
            fprc2s.push_back(rr >= 1e150 ? 1e300 : rr * rr);
        }
        if (l > max_l) max_l = l;
        crow += nc; prow += np; orow += nrow;
    }
    OkmiBasis* b = new OkmiBasis();
    b->ncont = ncont; b->nshell = ncont; b->nao_cart = crow; b->nao = orow; b->nprim = prow;
    b->ntab = (int)T.size(); b->max_l = max_l; b->natoms = natoms; b->is_normalized = is_normalized;
    (void)hipGetDevice(&b->device);
    b->h_shells = new ShellRec[ncont];
    std::memcpy(b->h_shells, shells.data(), sizeof(ShellRec) * ncont);
    b->h_prims = new PrimRec[prow];
    std::memcpy(b->h_prims, prims.data(), sizeof(PrimRec) * prow);
    b->h_T = new double[T.size()];
    std::memcpy(b->h_T, T.data(), sizeof(double) * T.size());
    b->h_c = new double[prow];
    std::memcpy(b->h_c, craw.data(), sizeof(double) * prow);
    b->h_prim_l = new int[prow];
    std::memcpy(b->h_prim_l, pl.data(), sizeof(int) * prow);
    b->h_frcut2.assign(frc2.begin(), frc2.end());
    b->h_fprcut2.assign(fprc2.begin(), fprc2.end());
    b->h_frcut2s.assign(frc2s.begin(), frc2s.end());
    b->h_fprcut2s.assign(fprc2s.begin(), fprc2s.end());
    int rc = dev_upload(&b->d_shells, shells.data(), shells.size(), "basis_create: shells");
    if (!rc) rc = dev_upload(&b->d_prims, prims.data(), prims.size(), "basis_create: primitives");
    if (!rc) rc = dev_upload(&b->d_T, T.data(), T.size(), "basis_create: transform");
    // the fused kernels' tables (chunk layout, work lists), built here so
    // that no later call has to allocate device memory
    for (int i = 0; i < ncont; ++i) b->max_nrow = std::max(b->max_nrow, shells[i].nrow);
    if (!rc && b->max_nrow <= 32) rc = okmi_prepare_fused(b);
    if (rc) {
        okmi_basis_destroy(b);
        return rc;
    }
    *out = b;
    return 0;
}

extern "C" int okmi_basis_destroy(okmi_basis* b)
{
    if (!b) return 0;
    if (b->d_shells) (void)hipFree(b->d_shells);
    if (b->d_prims) (void)hipFree(b->d_prims);
    if (b->d_T) (void)hipFree(b->d_T);
    if (b->d_chunks) (void)hipFree(b->d_chunks);
    if (b->d_fshells) (void)hipFree(b->d_fshells);
    if (b->d_fprims) (void)hipFree(b->d_fprims);
    if (b->d_fgroups) (void)hipFree(b->d_fgroups);
    if (b->d_oshells) (void)hipFree(b->d_oshells);
    if (b->d_ishells) (void)hipFree(b->d_ishells);
    if (b->d_wshells) (void)hipFree(b->d_wshells);
    for (void* p : {(void*)b->d_fshells_s, (void*)b->d_ishells_s, (void*)b->d_wshells_s, (void*)b->d_fprims_s,
                    (void*)b->d_fgroups_s})
        if (p) (void)hipFree(p);
    if (b->d_padrows) (void)hipFree(b->d_padrows);
    if (b->d_colmap) (void)hipFree(b->d_colmap);
    delete[] b->h_shells;
    delete[] b->h_prims;
    delete[] b->h_T;
    delete[] b->h_c;
    delete[] b->h_prim_l;
    delete b;
    return 0;
}

extern "C" int okmi_basis_nao(const okmi_basis* b) { return b ? b->nao : -1; }

extern "C" int okmi_ao_eval(const okmi_basis* b, const okmi_grid* g, int64_t p0, int64_t np, const int* drv,
                            int ndrv, int out_dtype, void* d_out, int64_t ld, void* stream)
{
    g_err.clear();
    if (!b || !d_out) {
        okmi_set_error("ao_eval: null basis or output");
        return 1;
    }
    if (check_grid(g, p0, np) || check_drv(drv, ndrv)) return 1;
    if (ld < np) {
        okmi_set_error("ao_eval: row stride %lld < %lld points", (long long)ld, (long long)np);
        return 1;
    }
    if (out_dtype != OKMI_F64 && out_dtype != OKMI_F32) {
        okmi_set_error("ao_eval: unknown output type %d", out_dtype);
        return 1;
    }
    return okmi_launch_ao_eval(b, to_desc(g), p0, np, drv, ndrv, out_dtype, d_out, ld, (int64_t)b->nao * ld,
                               (hipStream_t)stream);
}

extern "C" int okmi_mocreator(const double* d_ao, int nao, int64_t npts, int64_t ld_ao, const double* d_c, int nmo,
                              double* d_mo, int64_t ld_mo, void* stream)
{
    g_err.clear();
    if (!d_ao || !d_c || !d_mo || nao <= 0 || nmo <= 0 || npts < 0 || ld_ao < npts || ld_mo < npts) {
        okmi_set_error("mocreator: invalid arguments");
        return 1;
    }
    return okmi_launch_gemm_f64(d_ao, ld_ao, 0, 1, nao, d_c, nmo, npts, d_mo, ld_mo, (hipStream_t)stream);
}

namespace {
// Workspace of the staged paths per point: the AO sets (FP64; in the fused
// row order, nkpad >= nao rows, for the screened density) and the
// live-chunk lists of 64-point tiles; plus alignment slack
size_t staged_per_point(const okmi_basis* b, int nsets)
{
    const size_t rows = std::max(b->nao, b->nkpad);
    return (size_t)nsets * rows * sizeof(double) + ((size_t)b->nchunk + 1) * sizeof(int) / 64 +
           (size_t)live_mask_words(b) * 8 / 64 + 2;
}
constexpr size_t STAGED_SLACK = 1024;

int64_t chunk_points(const okmi_basis* b, int nsets, int64_t np, size_t ws_bytes)
{
    if (ws_bytes <= STAGED_SLACK) return 0;
    int64_t c = (int64_t)((ws_bytes - STAGED_SLACK) / staged_per_point(b, nsets));
    if (c >= np) return np;
    c &= ~(int64_t)63;
    return c;
}
}  // namespace

extern "C" size_t okmi_workspace_bytes(const okmi_basis* b, int nsets, int64_t chunk)
{
    if (!b || nsets <= 0 || chunk <= 0) return 0;
    return staged_per_point(b, nsets) * (size_t)chunk + STAGED_SLACK;
}

extern "C" int okmi_mo_eval(const okmi_basis* b, const okmi_grid* g, int64_t p0, int64_t np, const int* drv,
                            int ndrv, const double* d_c, int nmo, double* d_mo, int64_t ld_mo, void* d_ws,
                            size_t ws_bytes, void* stream)
{
    g_err.clear();
    if (!b || !d_c || !d_mo || nmo <= 0 || ld_mo < np) {
        okmi_set_error("mo_eval: invalid arguments");
        return 1;
    }
    if (check_grid(g, p0, np) || check_drv(drv, ndrv)) return 1;
    if (np == 0) return 0;
    const int64_t chunk = chunk_points(b, ndrv, np, ws_bytes);
    if (chunk <= 0 || !d_ws) {
        okmi_set_error("mo_eval: workspace of %zu bytes holds no chunk (%zu bytes per 64 points)", ws_bytes,
                       okmi_workspace_bytes(b, ndrv, 64));
        return 1;
    }
    hipStream_t st = (hipStream_t)stream;
    const GridDesc gd = to_desc(g);
    for (int64_t c0 = 0; c0 < np; c0 += chunk) {
        const int64_t n = np - c0 < chunk ? np - c0 : chunk;
        int rc = okmi_launch_ao_eval(b, gd, p0 + c0, n, drv, ndrv, OKMI_F64, d_ws, chunk, (int64_t)b->nao * chunk, st);
        if (rc) return rc;
        rc = okmi_launch_gemm_f64((const double*)d_ws, chunk, (int64_t)b->nao * chunk, ndrv, b->nao, d_c, nmo, n,
                                  d_mo + c0, ld_mo, st);
        if (rc) return rc;
    }
    return 0;
}

int okmi_rho_sets(const int* drv_out, int nout, int* codes, int* ncodes)
{
    if (nout < 0 || nout > 16) {
        okmi_set_error("rho: %d derivative outputs (up to 16 supported)", nout);
        return 1;
    }
    int mask = 1;
    for (int r = 0; r < nout; ++r) {
        const int d = drv_out[r];
        if (d < 0 || d > 9) {
            okmi_set_error("The selection `drv=%d` is not valid!", d);
            return 1;
        }
        mask |= 1 << d;
        if (d >= 4 && d <= 6) mask |= 1 << (d - 3);
        if (d == 7) mask |= (1 << 1) | (1 << 2);
        if (d == 8) mask |= (1 << 1) | (1 << 3);
        if (d == 9) mask |= (1 << 2) | (1 << 3);
    }
    int n = 0;
    for (int c = 0; c < OKMI_MAX_SETS; ++c)
        if (mask >> c & 1) codes[n++] = c;
    *ncodes = n;
    return 0;
}

extern "C" int okmi_rho_staged(const okmi_basis* b, const okmi_grid* g, int64_t p0, int64_t np, const double* d_c,
                               const double* d_occ, int nmo, const int* drv_out, int nout, double* d_rho,
                               double* d_delta, int64_t ld_delta, double* d_mo_norm, void* d_ws, size_t ws_bytes,
                               void* stream)
{
    g_err.clear();
    if (!b || !d_c || !d_occ || !d_rho || nmo <= 0 || (nout > 0 && (!d_delta || ld_delta < np))) {
        okmi_set_error("rho: invalid arguments");
        return 1;
    }
    if (check_grid(g, p0, np)) return 1;
    int codes[OKMI_MAX_SETS];
    int nsets = 0;
    if (okmi_rho_sets(drv_out, nout, codes, &nsets)) return 1;
    if (np == 0) return 0;
    const int64_t chunk = chunk_points(b, nsets, np, ws_bytes);
    if (chunk <= 0 || !d_ws) {
        okmi_set_error("rho: workspace of %zu bytes holds no chunk (%zu bytes per 64 points)", ws_bytes,
                       okmi_workspace_bytes(b, nsets, 64));
        return 1;
    }
    int64_t set_off[OKMI_MAX_SETS] = {0};
    int set_mask = 0;
    for (int s = 0; s < nsets; ++s) {
        set_off[codes[s]] = (int64_t)s * b->nao * chunk;
        set_mask |= 1 << codes[s];
    }
    hipStream_t st = (hipStream_t)stream;
    const GridDesc gd = to_desc(g);
    if (!(b->nkpad > 0 && b->d_ishells)) {   // no fused tables: every AO row of every point
        for (int64_t c0 = 0; c0 < np; c0 += chunk) {
            const int64_t n = np - c0 < chunk ? np - c0 : chunk;
            int rc = okmi_launch_ao_eval(b, gd, p0 + c0, n, codes, nsets, OKMI_F64, d_ws, chunk,
                                         (int64_t)b->nao * chunk, st);
            if (rc) return rc;
            rc = okmi_launch_rho_f64((const double*)d_ws, chunk, set_off, set_mask, b->nao, d_c, nmo, d_occ, n,
                                     drv_out, nout, d_rho + c0, d_delta ? d_delta + c0 : nullptr, ld_delta, d_mo_norm,
                                     st);
            if (rc) return rc;
        }
        return 0;
    }
    // Screened: the AO sets in the fused row order (shells packed into
    // 32-row chunks by radius class and Morton code), per 64-point tile the
    // chunks with a shell in reach, and the GEMM contracts only those.  A
    // whole regular grid runs in 4x4x4 bricks (compact tiles), other grids
    // in point order.
    // C in the fused row order at the front of the workspace, then the AO
    // sets and the live-chunk lists
    const int64_t rows = b->nkpad;
    const bool brick = g->regular && p0 == 0 && np == g->nx * g->ny * g->nz;
    const int64_t nbx = (g->nx + 3) / 4, nby = (g->ny + 3) / 4, nbz = (g->nz + 3) / 4;
    const int64_t total = brick ? nbx * nby * nbz * 64 : np;
    const size_t cbytes = (((size_t)nmo * rows * sizeof(double)) + 255) & ~(size_t)255;
    const int64_t ch = ws_bytes > cbytes + 256 ? chunk_points(b, nsets, total, ws_bytes - cbytes - 256) : 0;
    if (ch <= 0 || (brick && ch % 64)) {
        okmi_set_error("rho: workspace of %zu bytes holds no chunk (%zu bytes for C and 64 points)", ws_bytes,
                       cbytes + 256 + okmi_workspace_bytes(b, nsets, 64));
        return 1;
    }
    double* Cp = (double*)(((uintptr_t)d_ws + 255) & ~(uintptr_t)255);
    double* A = Cp + cbytes / sizeof(double);
    int* live = (int*)(A + (size_t)nsets * rows * ch);
    const int64_t ntile = (ch + 63) / 64;
    int* live_count = live + ntile * b->nchunk;
    const int nw = live_mask_words(b);
    unsigned long long* lmask =
        (unsigned long long*)(((uintptr_t)(live_count + ntile) + 7) & ~(uintptr_t)7);
    if ((size_t)((char*)(lmask + ntile * nw) - (char*)d_ws) > ws_bytes) {
        okmi_set_error("rho: workspace of %zu bytes too small for the screened layout", ws_bytes);
        return 1;
    }
    for (int s = 0; s < nsets; ++s) set_off[codes[s]] = (int64_t)s * rows * ch;
    int rc = okmi_launch_permute_c(d_c, nmo, b->nao, rows, b->d_colmap, Cp, st);
    if (!rc) rc = okmi_launch_zero_rows(A, nsets, rows * ch, ch, b->d_padrows, b->npadrows, ch, st);
    if (rc) return rc;
    F64Screen scr{};
    scr.live = live; scr.live_count = live_count; scr.nchunk = b->nchunk;
    scr.ldc = (int)rows; scr.brick = brick ? 1 : 0;
    scr.nbx = nbx; scr.nby = nby; scr.nbz = nbz; scr.nx = g->nx; scr.ny = g->ny; scr.nz = g->nz;
    const AoLiveMask am{lmask, nw, 0, nullptr};
    for (int64_t c0 = 0; c0 < total; c0 += ch) {
        const int64_t n = total - c0 < ch ? total - c0 : ch;
        // live chunks first: the AO kernel leaves the rows of dead chunks unwritten
        rc = okmi_launch_live_chunks(b, gd, brick ? 1 : 0, c0 / 64, p0 + c0, n, 64, live, live_count, st, lmask);
        if (rc) return rc;
        if (brick) {
            AoBricks bk{c0 / 64, n / 64, nbx, nby, nbz};
            rc = okmi_launch_ao_eval_ex(b, b->d_ishells, b->d_fprims, gd, 0, n, &bk, codes, nsets, OKMI_F64, A, ch,
                                        rows * ch, st, &am);
        } else {
            rc = okmi_launch_ao_eval_ex(b, b->d_ishells, b->d_fprims, gd, p0 + c0, n, nullptr, codes, nsets, OKMI_F64,
                                        A, ch, rows * ch, st, &am);
        }
        if (rc) return rc;
        scr.bk0 = c0 / 64;
        rc = okmi_launch_rho_f64(A, ch, set_off, set_mask, (int)rows, Cp, nmo, d_occ, n, drv_out, nout,
                                 brick ? d_rho : d_rho + c0, d_delta ? (brick ? d_delta : d_delta + c0) : nullptr,
                                 ld_delta, d_mo_norm, st, &scr);
        if (rc) return rc;
    }
    return okmi_check_launch("rho");
}

extern "C" int okmi_rho(const okmi_basis* b, const okmi_grid* g, int64_t p0, int64_t np, const double* d_c,
                        const double* d_occ, int nmo, const int* drv_out, int nout, double* d_rho, double* d_delta,
                        int64_t ld_delta, double* d_mo_norm, void* d_ws, size_t ws_bytes, void* stream)
{
    g_err.clear();
    if (!b || !d_c || !d_occ || !d_rho || nmo <= 0 || (nout > 0 && (!d_delta || ld_delta < np))) {
        okmi_set_error("rho: invalid arguments");
        return 1;
    }
    if (check_grid(g, p0, np)) return 1;
    int codes[OKMI_MAX_SETS];
    int nsets = 0;
    if (okmi_rho_sets(drv_out, nout, codes, &nsets)) return 1;
    if (np == 0) return 0;
    // Up to 64 MOs: the fused FP64 kernel (same arithmetic type, no HBM round
    // trip of the AO sets) when the workspace holds its buffers for at least
    // one 64-point tile; otherwise the staged path.  The fused tables are a
    // cache built once per basis (the basis is not otherwise changed).
    okmi_basis* mb = const_cast<okmi_basis*>(b);
    if (nmo <= 64 && b->max_nrow <= 32 && d_ws && ws_bytes >= okmi_fused_ws_bytes(mb, nmo, OKMI_PREC_F64, 64))
        return okmi_launch_rho_fused(mb, to_desc(g), p0, np, d_c, d_occ, nmo, OKMI_PREC_F64, drv_out, nout, d_rho,
                                     d_delta, ld_delta, d_mo_norm, d_ws, ws_bytes, (hipStream_t)stream);
    return okmi_rho_staged(b, g, p0, np, d_c, d_occ, nmo, drv_out, nout, d_rho, d_delta, ld_delta, d_mo_norm, d_ws,
                           ws_bytes, stream);
}

// ---------------------------------------------------------------------------
// okmi_rho for callers that cannot allocate device memory.  Its workspace is
// `scratch`, a file-scope device array of OKMI_NOALLOC_SCRATCH_BYTES (each GPU
// the library runs on holds its own copy until the process ends).  In host
// code the array's name is not a device address: the address of the current
// GPU's copy comes from hipGetSymbolAddress, taken on the first call on that
// GPU (with it current) and kept per device.  Every buffer the kernels use in
// it is written before it is read (C planes, occupations, term rows, per-tile
// norms, or the staged path's AO sets), so nothing is cleared; calls that
// overlap on one GPU share the array and race.
constexpr size_t OKMI_NOALLOC_SCRATCH_BYTES = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[OKMI_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int OKMI_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[OKMI_MAX_DEVICES] = {};

int scratch_address(void** out)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return okmi_hip_fail(e, "rho_noalloc: current device");
    if (dev < 0 || dev >= OKMI_MAX_DEVICES) {
        okmi_set_error("rho_noalloc: device %d (up to %d devices)", dev, OKMI_MAX_DEVICES);
        return 1;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return okmi_hip_fail(e, "rho_noalloc: scratch address");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}
}  // namespace

extern "C" size_t okmi_rho_noalloc_scratch_bytes(void) { return OKMI_NOALLOC_SCRATCH_BYTES; }

extern "C" int okmi_rho_noalloc(const okmi_basis* b, const okmi_grid* g, int64_t p0, int64_t np, const double* d_c,
                                const double* d_occ, int nmo, const int* drv_out, int nout, double* d_rho,
                                double* d_delta, int64_t ld_delta, double* d_mo_norm, void* stream)
{
    g_err.clear();
    if (!b || !d_c || !d_occ || !d_rho || nmo <= 0 || (nout > 0 && (!d_delta || ld_delta < np))) {
        okmi_set_error("rho_noalloc: invalid arguments");
        return 1;
    }
    int codes[OKMI_MAX_SETS];
    int nsets = 0;
    if (okmi_rho_sets(drv_out, nout, codes, &nsets)) return 1;
    // the scratch must hold the fixed buffers and one 64-point pass of the
    // path okmi_rho takes
    const bool fused = nmo <= 64 && b->max_nrow <= 32;
    const size_t need = fused ? okmi_fused_ws_bytes(const_cast<okmi_basis*>(b), nmo, OKMI_PREC_F64, 64)
                              : okmi_workspace_bytes(b, nsets, 64) + (size_t)nmo * std::max(b->nao, b->nkpad) * 8 + 512;
    if (need > OKMI_NOALLOC_SCRATCH_BYTES) {
        okmi_set_error("rho_noalloc: %d MOs x %d AOs need %zu bytes of scratch for one 64-point pass "
                       "(the scratch holds %zu); use okmi_rho with a workspace",
                       nmo, b->nao, need, OKMI_NOALLOC_SCRATCH_BYTES);
        return 1;
    }
    void* ws = nullptr;
    if (scratch_address(&ws)) return 1;
    g_no_alloc = true;
    const int rc = okmi_rho(b, g, p0, np, d_c, d_occ, nmo, drv_out, nout, d_rho, d_delta, ld_delta, d_mo_norm, ws,
                            OKMI_NOALLOC_SCRATCH_BYTES, stream);
    g_no_alloc = false;
    return rc;
}

extern "C" size_t okmi_rho_fused_workspace_bytes(const okmi_basis* b, int nmo, int precision, int64_t chunk)
{
    if (!b || nmo <= 0 || chunk <= 0) return 0;
    return okmi_fused_ws_bytes(b, nmo, precision, chunk);
}

extern "C" int okmi_rho_fused(okmi_basis* b, const okmi_grid* g, int64_t p0, int64_t np, const double* d_c,
                              const double* d_occ, int nmo, const int* drv_out, int nout, int precision,
                              double* d_rho, double* d_delta, int64_t ld_delta, void* d_ws, size_t ws_bytes,
                              void* stream)
{
    g_err.clear();
    if (!b || !d_c || !d_occ || !d_rho || nmo <= 0 || (nout > 0 && (!d_delta || ld_delta < np))) {
        okmi_set_error("rho_fused: invalid arguments");
        return 1;
    }
    if (check_grid(g, p0, np)) return 1;
    return okmi_launch_rho_fused(b, to_desc(g), p0, np, d_c, d_occ, nmo, precision, drv_out, nout, d_rho, d_delta,
                                 ld_delta, nullptr, d_ws, ws_bytes, (hipStream_t)stream);
}

extern "C" size_t okmi_rho_backward_workspace_bytes(const okmi_basis* b, int nmo, int64_t chunk)
{
    if (!b || nmo <= 0 || chunk <= 0) return 0;
    return okmi_backward_ws_bytes(b, nmo, chunk);
}

extern "C" int okmi_rho_backward(const okmi_basis* b, const okmi_grid* g, int64_t p0, int64_t np, const double* d_c,
                                 const double* d_occ, int nmo, const double* d_grad, double* d_dc, double* d_dalpha,
                                 double* d_dgeo, void* d_ws, size_t ws_bytes, void* stream)
{
    g_err.clear();
    if (!b || !d_c || !d_occ || !d_grad || !d_dc || !d_dalpha || !d_dgeo || nmo <= 0) {
        okmi_set_error("rho_backward: invalid arguments");
        return 1;
    }
    if (check_grid(g, p0, np)) return 1;
    return okmi_launch_rho_backward(b, to_desc(g), p0, np, d_c, d_occ, nmo, d_grad, d_dc, d_dalpha, d_dgeo, d_ws,
                                    ws_bytes, (hipStream_t)stream);
}

extern "C" size_t okmi_rho_backward_f32_workspace_bytes(const okmi_basis* b, int nmo, int64_t chunk)
{
    if (!b || nmo <= 0 || chunk <= 0) return 0;
    return okmi_backward_f32_ws_bytes(b, nmo, chunk);
}

extern "C" int okmi_rho_backward_f32(const okmi_basis* b, const okmi_grid* g, int64_t p0, int64_t np,
                                     const float* d_c, const float* d_occ, int nmo, const float* d_grad,
                                     float* d_dc, float* d_dalpha, float* d_dgeo, void* d_ws, size_t ws_bytes,
                                     void* stream)
{
    g_err.clear();
    if (!b || !d_c || !d_occ || !d_grad || !d_dc || !d_dalpha || !d_dgeo || nmo <= 0) {
        okmi_set_error("rho_backward_f32: invalid arguments");
        return 1;
    }
    if (check_grid(g, p0, np)) return 1;
    return okmi_launch_rho_backward_f32(b, to_desc(g), p0, np, d_c, d_occ, nmo, d_grad, d_dc, d_dalpha, d_dgeo,
                                        d_ws, ws_bytes, (hipStream_t)stream);
}

extern "C" int okmi_batch_create(int nmol, okmi_basis* const* bases, const okmi_grid* grids, const double* const* d_c,
                                 const double* const* d_occ, const int* nmo, double* const* d_rho,
                                 double* const* d_grad, int precision, int order_by_cost, okmi_batch** out)
{
    g_err.clear();
    if (!out || nmol <= 0 || !bases || !grids || !d_c || !d_occ || !nmo || !d_rho) {
        okmi_set_error("batch_create: invalid arguments");
        return 1;
    }
    *out = nullptr;
    std::vector<GridDesc> gd(nmol);
    std::vector<int64_t> np(nmol);
    std::vector<double*> grad(nmol, nullptr);
    for (int m = 0; m < nmol; ++m) {
        const okmi_grid* g = grids + m;
        if (!bases[m] || !d_c[m] || !d_occ[m] || !d_rho[m]) {
            okmi_set_error("batch_create: molecule %d has a null input or output", m);
            return 1;
        }
        const int64_t total = g->regular ? g->nx * g->ny * g->nz : g->nx;
        if (check_grid(g, 0, total)) return 1;
        gd[m] = to_desc(g);
        np[m] = total;
        if (d_grad) grad[m] = d_grad[m];
    }
    return okmi_batch_build(nmol, bases, gd.data(), np.data(), d_c, d_occ, nmo, d_rho, grad.data(), precision,
                            order_by_cost, out);
}

extern "C" int64_t okmi_batch_tiles(const okmi_batch* b) { return b ? okmi_batch_ntile(b) : 0; }

extern "C" int okmi_batch_set_workgroups(okmi_batch* b, int workgroups)
{
    g_err.clear();
    if (!b) {
        okmi_set_error("batch_set_workgroups: null batch");
        return 1;
    }
    return okmi_batch_set_grid_impl(b, workgroups);
}

extern "C" int okmi_batch_set_tile_shape(okmi_batch* b, int log2_x, int log2_y, int log2_z)
{
    g_err.clear();
    if (!b) {
        okmi_set_error("batch_set_tile_shape: null batch");
        return 1;
    }
    return okmi_batch_set_shape_impl(b, log2_x, log2_y, log2_z);
}

extern "C" int okmi_batch_launch(okmi_batch* b, int use_graph, void* stream)
{
    g_err.clear();
    if (!b) {
        okmi_set_error("batch_launch: null batch");
        return 1;
    }
    return okmi_batch_run(b, use_graph, (hipStream_t)stream);
}

extern "C" int okmi_batch_destroy(okmi_batch* b)
{
    okmi_batch_free(b);
    return 0;
}

extern "C" size_t okmi_electron_count_workspace_bytes(void) { return okmi_electrons_ws_bytes(); }

extern "C" int okmi_electron_count(const double* d_rho, int64_t n, double d3r, void* d_ws, size_t ws_bytes,
                                   double* d_out, void* stream)
{
    g_err.clear();
    if ((!d_rho && n > 0) || !d_out || n < 0) {
        okmi_set_error("electron_count: invalid arguments");
        return 1;
    }
    return okmi_launch_electrons(d_rho, n, d3r, d_ws, ws_bytes, d_out, (hipStream_t)stream);
}

extern "C" size_t okmi_gross_atomic_workspace_bytes(const okmi_basis* b, int64_t chunk, int nrows_total, int nsel)
{
    if (!b || chunk <= 0 || nrows_total < 0 || nsel < 0) return 0;
    return okmi_gross_ws_bytes(b, chunk, nrows_total, nsel);
}

extern "C" int okmi_gross_atomic_density(const okmi_basis* b, const okmi_grid* g, int64_t p0, int64_t np, int drv,
                                         const double* d_c, const double* d_occ, int nmo, const int* sel_start,
                                         const int* sel_rows, int nsel, double* d_out, int64_t ld_out, void* d_ws,
                                         size_t ws_bytes, void* stream)
{
    g_err.clear();
    if (!b || !d_c || !d_occ || nmo <= 0 || nsel < 0 || (nsel > 0 && (!sel_start || !d_out || ld_out < np))) {
        okmi_set_error("gross_atomic_density: invalid arguments");
        return 1;
    }
    if (check_grid(g, p0, np) || check_drv(&drv, 1)) return 1;
    if (nsel > 0 && sel_start[nsel] > 0 && !sel_rows) {
        okmi_set_error("gross_atomic_density: missing AO row list");
        return 1;
    }
    return okmi_launch_gross_atomic(b, to_desc(g), p0, np, drv, d_c, d_occ, nmo, sel_start, sel_rows, nsel, d_out,
                                    ld_out, d_ws, ws_bytes, (hipStream_t)stream);
}

extern "C" int okmi_masked_mo_product(const double* d_c, const double* d_occ, int nmo, int nao, const double* d_ao,
                                      const double* d_mo, int64_t np, int64_t ld, const int* d_rows, int nrows,
                                      double* d_out, void* stream)
{
    g_err.clear();
    if (!d_c || !d_occ || !d_ao || !d_mo || !d_out || nmo <= 0 || nao <= 0 || ld < np || nrows < 0 ||
        (nrows > 0 && !d_rows)) {
        okmi_set_error("masked_mo_product: invalid arguments");
        return 1;
    }
    return okmi_launch_masked_mo_product(d_c, d_occ, nmo, nao, d_ao, d_mo, np, ld, d_rows, nrows, d_out,
                                         (hipStream_t)stream);
}

extern "C" int okmi_mo_pair_products(const double* d_bra, int na, const double* d_ket, int nb, int nd, int64_t np,
                                     int64_t ld_in, double* d_out, int64_t ld_out, void* stream)
{
    g_err.clear();
    if (!d_bra || !d_ket || !d_out || na < 0 || nb < 0 || nd < 0 || ld_in < np || ld_out < np) {
        okmi_set_error("mo_pair_products: invalid arguments");
        return 1;
    }
    return okmi_launch_mo_pairs(d_bra, na, d_ket, nb, nd, np, ld_in, d_out, ld_out, (hipStream_t)stream);
}

extern "C" int okmi_transition_flux(const double* d_mo, int nmo, int nd, int64_t np, int64_t ld_in,
                                    const int* d_pairs, int npair, double* d_out, int64_t ld_out, void* stream)
{
    g_err.clear();
    if (!d_mo || !d_out || nmo <= 0 || nd < 0 || npair < 0 || (npair > 0 && !d_pairs) || ld_in < np ||
        ld_out < np) {
        okmi_set_error("transition_flux: invalid arguments");
        return 1;
    }
    return okmi_launch_jmo(d_mo, nmo, nd, np, ld_in, d_pairs, npair, d_out, ld_out, (hipStream_t)stream);
}

extern "C" size_t okmi_ci_workspace_bytes(int nterm) { return nterm < 0 ? 0 : okmi_ci_ws_bytes(nterm); }

extern "C" int okmi_ci_reduce(const double* d_mo, const double* d_dmo, int nmo, int nd, int64_t npts, int64_t ld,
                              const double* coef, const int* ia, const int* ib, int nterm, int mode, double* d_out,
                              int64_t ld_out, void* d_ws, size_t ws_bytes, void* stream)
{
    g_err.clear();
    if (!d_mo || !d_out || nmo <= 0 || npts < 0 || ld < npts || nterm < 0 || mode < -1 || mode > 1 ||
        (nterm > 0 && (!coef || !ia || !ib)) || (mode >= 0 && (!d_dmo || nd <= 0 || ld_out < npts))) {
        okmi_set_error("ci_reduce: invalid arguments");
        return 1;
    }
    return okmi_launch_ci(d_mo, d_dmo, nmo, nd, npts, ld, coef, ia, ib, nterm, mode, d_out, ld_out, d_ws, ws_bytes,
                          (hipStream_t)stream);
}
