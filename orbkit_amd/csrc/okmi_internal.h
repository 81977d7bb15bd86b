// Library-internal state and helpers (not part of the C ABI).
#pragma once
#include "okmi_common.h"
#include "../../include/orbkit_amd.h"

#include <algorithm>
#include <vector>

struct OkmiBasis {
    int ncont = 0;
    int nshell = 0;
    int nao_cart = 0;      // Cartesian rows (before transform)
    int nao = 0;           // output rows (Cartesian or spherical)
    int nprim = 0;
    int ntab = 0;          // doubles in T
    int max_l = 0;
    int natoms = 0;
    int device = -1;
    ShellRec* d_shells = nullptr;
    PrimRec* d_prims = nullptr;
    double* d_T = nullptr;
    // host copies (kept for rebuilds and for the fused kernels' chunk plans)
    ShellRec* h_shells = nullptr;
    PrimRec* h_prims = nullptr;
    double* h_T = nullptr;
    double* h_c = nullptr;      // raw contraction coefficients per primitive
    int* h_prim_l = nullptr;    // l per primitive
    // fused-kernel tables (built on first use).  The fused kernels use their
    // own row order: each shell's rows lie inside one 32-row chunk (zero rows
    // pad the chunks), colmap maps internal rows to output rows (-1: pad),
    // and screening radii for the bound OKMI_SCREEN_FUSED.
    int2* d_chunks = nullptr;   // per chunk: first work entry, shells | used rows << 16
    ShellRec* d_fshells = nullptr;   // row0 = internal row, rcut2 = fused radius
    ShellRec* d_oshells = nullptr;   // output rows, rcut2 = fused radius
    ShellRec* d_ishells = nullptr;   // internal rows (prim0 into d_fprims), rcut2 = fused radius
    ShellRec* d_wshells = nullptr;   // the same in work order (entry j of a chunk's list is record j)
    int* d_padrows = nullptr;        // internal rows that hold no AO (chunk padding)
    int npadrows = 0;
    PrimRec* d_fprims = nullptr;     // rcut2 = fused radius
    PrimGroup* d_fgroups = nullptr;  // fused-kernel primitives, groups of 4 per shell in work order
    // the same tables with the radii of the bound OKMI_SCREEN_SPLIT (the
    // paths whose AO values end in FP32; same chunks, rows and work order)
    ShellRec* d_fshells_s = nullptr;
    ShellRec* d_ishells_s = nullptr;
    ShellRec* d_wshells_s = nullptr;
    PrimRec* d_fprims_s = nullptr;
    PrimGroup* d_fgroups_s = nullptr;
    int* d_colmap = nullptr;         // [nkpad]
    int nchunk = 0;
    int nkpad = 0;
    std::vector<double> h_frcut2, h_fprcut2;     // radii (squared) for OKMI_SCREEN_FUSED
    std::vector<double> h_frcut2s, h_fprcut2s;   // and for OKMI_SCREEN_SPLIT
    std::vector<int2> h_chunks;          // host copies of d_chunks, d_fshells and d_fshells_s
    std::vector<ShellRec> h_fshells, h_fshells_s;
    int is_normalized = 0;
    int max_nrow = 0;      // most output rows of one shell (the fused kernels take up to 32)
};

void okmi_set_error(const char* fmt, ...);
int okmi_check_launch(const char* what);
int okmi_hip_fail(hipError_t e, const char* what);

int okmi_launch_ao_eval(const OkmiBasis* b, const GridDesc& g, int64_t p0, int64_t np, const int* drv, int ndrv,
                        int out_dtype, void* out, int64_t ld, int64_t set_stride, hipStream_t st);

// the same with explicit shell/primitive tables (e.g. the fused screening
// radii) and, for regular grids, brick order: bricks [bk0, bk0 + nbk) of
// nbx x nby x nbz 4x4x4 bricks, column 64 (b - bk0) + lane
struct AoBricks {
    int64_t bk0, nbk, nbx, nby, nbz;
};
// With the shells in the fused row order (d_ishells): per tile of 64 << shift
// columns, bit c of mask[tile * words + c / 64] marks chunk c live; a shell
// out of reach of a wave's points whose chunk is dead for the wave's tile
// writes nothing (the contraction never reads those rows)
struct AoLiveMask {
    const unsigned long long* mask;
    int words;
    int shift;
    const ShellRec* wshells;   // work-order shells of the screening bound in use (null: d_wshells)
};
int okmi_launch_ao_eval_ex(const OkmiBasis* b, const ShellRec* shells, const PrimRec* prims, const GridDesc& g,
                           int64_t p0, int64_t np, const AoBricks* bricks, const int* drv, int ndrv, int out_dtype,
                           void* out, int64_t ld, int64_t set_stride, hipStream_t st,
                           const AoLiveMask* live = nullptr);

// fused-path tables (internal row order, fused screening radii) and the
// output-row shell table with the fused radii (d_oshells); built once
int okmi_prepare_fused(OkmiBasis* b);

// MO values Phi_s = C A_s for NS sets stored [set][mo][point] (FP64 MFMA)
int okmi_launch_gemm_f64(const double* A, int64_t ld, int64_t set_stride, int nsets, int nao, const double* C,
                         int nmo, int64_t np, double* mo_out, int64_t ld_mo, hipStream_t st);

// Screening of the staged FP64 density: AO rows in the fused row order
// (C permuted to the same order), per 64-column tile the chunks with a shell in
// reach, and optionally brick-ordered columns (column c is lane c % 64 of
// brick bk0 + c / 64 of an nbx x nby x nbz brick grid over nx x ny x nz).
struct F64Screen {
    const int* live;         // [tile][nchunk]
    const int* live_count;   // [tile]
    int nchunk;
    int ldc;                 // row stride of C (the basis's output rows)
    int brick;
    int64_t bk0, nbx, nby, nbz, nx, ny, nz;
};

// Density and derivative outputs from AO sets indexed by derivative code:
// set_off[c] is the offset of code c's AO set in A (codes in set_mask),
// out_code[r] the code of output row r.  With `scr` the AO rows (nao of
// them) are in the fused row order and each 64-column tile contracts only
// its live chunks.
int okmi_launch_rho_f64(const double* A, int64_t ld, const int64_t* set_off, int set_mask, int nao,
                        const double* C, int nmo, const double* occ, int64_t np, const int* out_code, int nout,
                        double* rho, double* delta, int64_t ld_delta, double* mo_norm, hipStream_t st,
                        const F64Screen* scr = nullptr);

// per tile of `np` columns (brick lanes from brick bk0 when `brick`, else
// points p0 + column) the 32-row chunks of the fused layout with a shell in
// reach, in order: list[tile][..], count[tile]
int okmi_launch_live_chunks(const OkmiBasis* b, const GridDesc& g, int brick, int64_t bk0, int64_t p0, int64_t n,
                            int np, int* list, int* count, hipStream_t st, unsigned long long* mask = nullptr, bool split = false);
// 64-bit words per tile of a live-chunk mask
inline int live_mask_words(const OkmiBasis* b) { return (b->nchunk + 63) / 64; }
// C [nmo][nao] -> [nmo][kpad] in the fused row order (colmap[k]: C column of row k, -1: zero)
int okmi_launch_permute_c(const double* C, int nmo, int nao, int64_t kpad, const int* colmap, double* out,
                          hipStream_t st);
// zero rows rows[0 .. nrows) of S sets [S][set_stride] (row stride ld) over columns [0, ncol)
int okmi_launch_zero_rows(double* a, int S, int64_t set_stride, int64_t ld, const int* rows, int nrows, int64_t ncol,
                          hipStream_t st);

// codes whose AO sets a density request needs: 0, the requested codes and
// the first derivatives entering second-derivative product terms
int okmi_rho_sets(const int* drv_out, int nout, int* codes, int* ncodes);

// fused AO -> MFMA -> density (reduced-precision contraction)
size_t okmi_fused_ws_bytes(const OkmiBasis* b, int nmo, int prec, int64_t np);
// (mo_norm, FP64 only: += sum_p phi_i(p)^2, or null)
int okmi_launch_rho_fused(OkmiBasis* b, const GridDesc& g, int64_t p0, int64_t np, const double* d_c,
                          const double* d_occ, int nmo, int prec, const int* drv_out, int nout, double* rho,
                          double* delta, int64_t ld_delta, double* mo_norm, void* ws, size_t ws_bytes,
                          hipStream_t st);

// many-MO density: FP32 AO sets per chunk + MFMA GEMM with density epilogue
// (S AO sets: 1, 4, 7 or 10, the smallest holding the requested codes)
size_t okmi_big_ws_bytes(const OkmiBasis* b, int nmo, int prec, int S, int64_t np);
int okmi_launch_rho_big(OkmiBasis* b, const GridDesc& g, int64_t p0, int64_t np, const double* d_c,
                        const double* d_occ, int nmo, int prec, const int* drv_out, int nout, double* rho,
                        double* delta, int64_t ld_delta, void* ws, size_t ws_bytes, hipStream_t st);

// backward of the FP64 density (accumulating dL/dC, dL/dalpha, dL/dR)
size_t okmi_backward_ws_bytes(const OkmiBasis* b, int nmo, int64_t chunk);
int okmi_launch_rho_backward(const OkmiBasis* b, const GridDesc& g, int64_t p0, int64_t np, const double* d_c,
                             const double* d_occ, int nmo, const double* d_grad, double* d_dc, double* d_dalpha,
                             double* d_dr, void* ws, size_t ws_bytes, hipStream_t st);
// the same in FP32 arithmetic (FP64 grid geometry, exp arguments and reductions)
size_t okmi_backward_f32_ws_bytes(const OkmiBasis* b, int nmo, int64_t chunk);
int okmi_launch_rho_backward_f32(const OkmiBasis* b, const GridDesc& g, int64_t p0, int64_t np, const float* d_c,
                                 const float* d_occ, int nmo, const float* d_grad, float* d_dc, float* d_dalpha,
                                 float* d_dr, void* ws, size_t ws_bytes, hipStream_t st);

// grid analysis helpers (extras.hip)
size_t okmi_gross_ws_bytes(const OkmiBasis* b, int64_t chunk, int nrows_total, int nsel);
int okmi_launch_gross_atomic(const OkmiBasis* b, const GridDesc& g, int64_t p0, int64_t np, int drv, const double* d_c,
                             const double* d_occ, int nmo, const int* sel_start, const int* sel_rows, int nsel,
                             double* d_out, int64_t ld_out, void* ws, size_t ws_bytes, hipStream_t st);
int okmi_launch_masked_mo_product(const double* d_c, const double* d_occ, int nmo, int nao, const double* d_ao,
                                  const double* d_mo, int64_t np, int64_t ld, const int* d_rows, int nrows,
                                  double* d_out, hipStream_t st);
int okmi_launch_mo_pairs(const double* d_bra, int na, const double* d_ket, int nb, int nd, int64_t np, int64_t ld_in,
                         double* d_out, int64_t ld_out, hipStream_t st);
int okmi_launch_jmo(const double* d_mo, int nmo, int nd, int64_t np, int64_t ld_in, const int* d_pairs, int npair,
                    double* d_out, int64_t ld_out, hipStream_t st);

// detCI grid reductions (detci.hip); mode -1 density, 0 jab, 1 a_nabla_b
size_t okmi_ci_ws_bytes(int nterm);
int okmi_launch_ci(const double* d_mo, const double* d_dmo, int nmo, int nd, int64_t np, int64_t ld,
                   const double* coef, const int* ia, const int* ib, int nterm, int mode, double* d_out, int64_t ld_out,
                   void* ws, size_t ws_bytes, hipStream_t st);

// compensated FP64 electron count (electrons.hip)
size_t okmi_electrons_ws_bytes();
int okmi_launch_electrons(const double* d_rho, int64_t n, double d3r, void* ws, size_t ws_bytes, double* d_out,
                          hipStream_t st);

// many molecules in one persistent launch (rho_fused.hip)
struct OkmiBatch;
int okmi_batch_build(int nmol, OkmiBasis* const* bases, const GridDesc* grids, const int64_t* npts,
                     const double* const* d_c, const double* const* d_occ, const int* nmo, double* const* d_rho,
                     double* const* d_grad, int prec, int order_by_cost, OkmiBatch** out);
int okmi_batch_run(OkmiBatch* b, int use_graph, hipStream_t st);
int okmi_batch_set_grid_impl(OkmiBatch* b, int blocks);
int okmi_batch_set_shape_impl(OkmiBatch* b, int lx, int ly, int lz);
int64_t okmi_batch_ntile(const OkmiBatch* b);
void okmi_batch_free(OkmiBatch* b);

// Optional per-kernel timing with HIP events recorded on the launch stream
// (okmi_profile_enable / okmi_profile_query).  Scope object around a launch.
struct OkmiProfScope {
    int idx = -1;
    hipStream_t st = nullptr;
    OkmiProfScope(const char* name, hipStream_t s);
    ~OkmiProfScope();
};
