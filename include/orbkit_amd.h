/*
 * C ABI of liborbkit_amd.so: orbkit's grid-density path on MI355X (gfx950).
 *
 * Conventions
 *   - Every function returns 0 on success and a nonzero status otherwise;
 *     okmi_last_error() then gives a message (thread-local).
 *   - Pointers named d_* are device pointers (HIP global memory); all other
 *     array pointers are host pointers.  Arrays are C-contiguous.
 *   - `stream` is a hipStream_t passed as void* (NULL = default stream).
 *     Work is enqueued on it; functions do not synchronise unless stated.
 *   - Grid points: a vector grid (x[i], y[i], z[i]) or a regular grid given
 *     by its three axes, flattened with x slowest and z fastest
 *     (orbkit/cy_grid.pyx:12-30).  [p0, p0+np) selects a contiguous range
 *     of flattened points, which is how orbkit's slices are expressed
 *     (orbkit/core.py:504-514).
 *   - Derivative codes: 0 none, 1 x, 2 y, 3 z, 4 xx, 5 yy, 6 zz, 7 xy, 8 xz,
 *     9 yz (orbkit/tools.py:225-239; codes 7-9 use orbkit's own formula,
 *     orbkit/c_support.c:121-168).
 */
#ifndef ORBKIT_AMD_H
#define ORBKIT_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OKMI_F64 = 0, OKMI_F32 = 1 };

/* arithmetic of the MO contraction in okmi_rho_fused */
enum { OKMI_PREC_F64 = 0, OKMI_PREC_F32 = 1, OKMI_PREC_BF16X3 = 2, OKMI_PREC_BF16X6 = 3 };

typedef struct okmi_grid {
    const double* d_x;
    const double* d_y;
    const double* d_z;
    int64_t nx, ny, nz;   /* regular: axis lengths; vector: nx = number of points, ny = nz = 0 */
    int regular;
    int pad;
} okmi_grid;

typedef struct OkmiBasis okmi_basis;

/* Message for the last failing call on this thread ("" if none). */
const char* okmi_last_error(void);

/*
 * Builds the device basis table.  The first arrays are exactly those orbkit's
 * AO kernel receives from AOClass (orbkit/cy_core.pyx:49-78, arguments of
 * cy_core.aocreator): Cartesian exponents lxlylz[nao_cart][3], components
 * per contraction, primitive (exponent, coefficient) pairs prim_coeffs[nprim][2],
 * primitives per contraction, atom of each contraction, atom positions
 * geo[natoms][3] (bohr) and the pre-normalised flag (orbkit/c_support.c:182).
 *   cart_scale   NULL or one factor per Cartesian row (orbkit's ao_spec[0]['N'],
 *                orbkit/core.py:98-100)
 *   nout_per_cont, transform
 *                NULL for Cartesian output rows; otherwise a dense map per
 *                contraction, nout x ncart (row-major, columns in that
 *                contraction's lxlylz order), concatenated: the
 *                Cartesian-to-spherical combinations of orbkit/core.py:135-176
 *                with orbkit/tools.py:148-220's coefficients.
 * Replaces: cy_core.aocreator's argument flattening (orbkit/cy_core.pyx:49-78)
 * and the per-call norm table of c_lcreator (orbkit/c_grid-based.c:27-36).
 */
int okmi_basis_create(int ncont, const int* ncart_per_cont, const int* lxlylz,
                      const int* nprim_per_cont, const double* prim_coeffs,
                      int natoms, const double* geo, const int* cont_atom,
                      int is_normalized, const double* cart_scale,
                      const int* nout_per_cont, const double* transform,
                      okmi_basis** out);
int okmi_basis_destroy(okmi_basis* b);
/* number of output AO rows (Cartesian or spherical) */
int okmi_basis_nao(const okmi_basis* b);

/*
 * AO values (or derivatives) on points [p0, p0+np):
 *   d_out[j][row][i] = derivative drv[j] of AO row `row` at point p0+i,
 *   row stride ld (>= np), set stride nao*ld, out_dtype OKMI_F64/OKMI_F32.
 * Replaces c_lcreator (orbkit/c_grid-based.h:3-5) over all contractions as
 * done by cy_core.aocreator (orbkit/cy_core.pyx:49-78), plus the row
 * renormalisation and cartesian2spherical of ao_creator (orbkit/core.py:98-102).
 */
int okmi_ao_eval(const okmi_basis* b, const okmi_grid* g, int64_t p0, int64_t np,
                 const int* drv, int ndrv, int out_dtype, void* d_out, int64_t ld, void* stream);

/*
 * MO[i][j] = sum_k C[i][k] AO[k][j] in FP64 on the device (MFMA).
 * d_ao[nao][ld_ao], d_c[nmo][nao], d_mo[nmo][ld_mo]; npts columns.
 * Replaces cy_core.mocreator (orbkit/cy_core.pyx:80-101).
 */
int okmi_mocreator(const double* d_ao, int nao, int64_t npts, int64_t ld_ao,
                   const double* d_c, int nmo, double* d_mo, int64_t ld_mo, void* stream);

/*
 * Workspace bytes okmi_rho / okmi_mo_eval need to process `chunk` points per
 * pass with `nsets` AO derivative sets (FP64 AO staging, with the density's
 * live-chunk lists).  The staged density (okmi_rho_staged, okmi_rho beyond
 * 64 MOs) also keeps a copy of C in its row order: nmo x (nao rounded up to
 * whole 32-row chunks) doubles more.
 */
size_t okmi_workspace_bytes(const okmi_basis* b, int nsets, int64_t chunk);

/*
 * MOs (and their derivatives) on points [p0, p0+np), FP64:
 *   d_mo[j][i][p] = derivative drv[j] of MO i at point p0+p, row stride ld_mo.
 * Replaces slice_rho's calc_mo branch (orbkit/core.py:250-263).
 */
int okmi_mo_eval(const okmi_basis* b, const okmi_grid* g, int64_t p0, int64_t np,
                 const int* drv, int ndrv, const double* d_c, int nmo,
                 double* d_mo, int64_t ld_mo, void* d_ws, size_t ws_bytes, void* stream);

/*
 * Density and derivatives on points [p0, p0+np), FP64 throughout:
 *   d_rho[p]        = sum_i occ_i phi_i(p)^2
 *   d_delta[r][p]   for each label r of drv_out (codes 1-9):
 *                     first derivatives 2 sum_i occ_i phi_i d_phi_i,
 *                     second ones add 2 sum_i occ_i d_a phi_i d_b phi_i
 *                     (orbkit/core.py:280-304); row stride ld_delta
 *   d_mo_norm[i]   += sum_p phi_i(p)^2 when non-NULL (orbkit/core.py:266-274)
 * Replaces slice_rho's density branch (orbkit/core.py:179-308).
 * Up to 64 MOs, and when the workspace holds at least
 * okmi_rho_fused_workspace_bytes(b, nmo, OKMI_PREC_F64, 64) bytes, it runs
 * the fused FP64 kernel (AO values in FP64 in LDS, FP64 MFMA; points in
 * passes as large as the workspace allows); otherwise okmi_rho_staged.  The
 * basis's fused-kernel tables are built on first use.
 */
int okmi_rho(const okmi_basis* b, const okmi_grid* g, int64_t p0, int64_t np,
             const double* d_c, const double* d_occ, int nmo,
             const int* drv_out, int nout, double* d_rho, double* d_delta, int64_t ld_delta,
             double* d_mo_norm, void* d_ws, size_t ws_bytes, void* stream);

/*
 * okmi_rho without a workspace argument, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  It
 * calls no allocating routine: its workspace is a device array of
 * okmi_rho_noalloc_scratch_bytes() (64 MiB) held by each GPU the library
 * runs on until the process ends, and the basis's fused tables are built by
 * okmi_basis_create.  Requests whose fixed buffers plus one 64-point pass
 * exceed the scratch (MO coefficients of very large bases) return an error;
 * larger grids run in several passes.  Calls that overlap on one GPU share
 * the scratch and race: serialise them (one stream per GPU, or wait).
 * Kernel timing (okmi_profile_enable) is not recorded inside this call.
 */
size_t okmi_rho_noalloc_scratch_bytes(void);
int okmi_rho_noalloc(const okmi_basis* b, const okmi_grid* g, int64_t p0, int64_t np,
                     const double* d_c, const double* d_occ, int nmo,
                     const int* drv_out, int nout, double* d_rho, double* d_delta, int64_t ld_delta,
                     double* d_mo_norm, void* stream);

/*
 * The same as okmi_rho in orbkit's own two stages, any number of MOs: the
 * FP64 AO sets of a chunk of points written to the workspace (okmi_ao_eval's
 * kernel, the workspace sized by okmi_workspace_bytes plus the copy of C),
 * then the FP64 MFMA contraction with the density assembled in its epilogue.
 * The AO rows are staged in the fused kernels' row order (shells in 32-row
 * chunks) and each 64-point tile (a 4x4x4 brick of a whole regular grid,
 * else 64 consecutive points) contracts only the chunks with a shell whose
 * screening radius reaches it.  Replaces the same code as okmi_rho
 * (orbkit/core.py:179-308, orbkit/cy_core.pyx:49-101).
 */
int okmi_rho_staged(const okmi_basis* b, const okmi_grid* g, int64_t p0, int64_t np,
                    const double* d_c, const double* d_occ, int nmo,
                    const int* drv_out, int nout, double* d_rho, double* d_delta, int64_t ld_delta,
                    double* d_mo_norm, void* d_ws, size_t ws_bytes, void* stream);

/*
 * Workspace bytes okmi_rho_fused needs to process `chunk` points per pass
 * (more than 64 MOs: sized for 7 AO sets; requests with mixed second
 * derivatives then run in shorter passes).
 */
size_t okmi_rho_fused_workspace_bytes(const okmi_basis* b, int nmo, int precision, int64_t chunk);

/*
 * Density and derivatives in one fused launch per pass: AO values are
 * evaluated per 64-point tile into LDS (FP64 coordinates and exponent range
 * reduction, FP32 values) and contracted with C on the matrix cores
 * (precision OKMI_PREC_F32: FP32 MFMA; OKMI_PREC_BF16X3 / _BF16X6: C and AO
 * split into two / three bf16 parts, 3 / 6 bf16 MFMA products, FP32
 * accumulation); the density terms are assembled in FP64.  Up to 64 MOs
 * are handled in one fused launch per pass (MOs never written to memory);
 * with more MOs the AO sets of each chunk of points are staged in HBM
 * (FP32) and contracted by an MFMA GEMM whose epilogue forms the density
 * terms of 128 MOs at a time (same precisions).  Outputs as okmi_rho.
 * Replaces slice_rho's density branch (orbkit/core.py:179-308) with the
 * reduced-precision contraction; okmi_rho is the FP64 form.
 */
int okmi_rho_fused(okmi_basis* b, const okmi_grid* g, int64_t p0, int64_t np,
                   const double* d_c, const double* d_occ, int nmo,
                   const int* drv_out, int nout, int precision,
                   double* d_rho, double* d_delta, int64_t ld_delta,
                   void* d_ws, size_t ws_bytes, void* stream);

/*
 * Workspace bytes okmi_rho_backward needs to process `chunk` points per pass.
 */
size_t okmi_rho_backward_workspace_bytes(const okmi_basis* b, int nmo, int64_t chunk);

/*
 * Gradient of a scalar loss L(rho) through the FP64 density of okmi_rho
 * (no derivative outputs) on points [p0, p0+np), given d_grad[p] = dL/drho
 * at point p0+p.  ACCUMULATES (callers zero the outputs first, so that the
 * slices of one grid can be processed by successive calls):
 *   d_dc[i][mu]   += dL/dC[i][mu]            (nmo x nao, the layout of d_c)
 *   d_dalpha[k]   += dL/dalpha_k             (one per primitive, in the order
 *                                             of prim_coeffs at basis_create)
 *   d_dgeo[a][d]  += dL/dR[a][d]             (natoms x 3, nuclear positions)
 * Row factors (cart_scale, the transform, contraction coefficients) are
 * held fixed.  FP64 throughout; AO values are recomputed per chunk.
 * orbkit has no derivative code: this differentiates the forward path of
 * rho_compute (orbkit/core.py:529-607) as evaluated by okmi_rho.
 */
int okmi_rho_backward(const okmi_basis* b, const okmi_grid* g, int64_t p0, int64_t np,
                      const double* d_c, const double* d_occ, int nmo, const double* d_grad,
                      double* d_dc, double* d_dalpha, double* d_dgeo,
                      void* d_ws, size_t ws_bytes, void* stream);

/*
 * Workspace bytes okmi_rho_backward_f32 needs to process `chunk` points per
 * pass (about half of the FP64 figure).
 */
size_t okmi_rho_backward_f32_workspace_bytes(const okmi_basis* b, int nmo, int64_t chunk);

/*
 * okmi_rho_backward in FP32 arithmetic: C, occupations, dL/drho and the
 * accumulated outputs are FP32; AO values, MO values, dL/dA and the MFMA
 * products (v_mfma_f32_16x16x4f32) are FP32.  Kept in FP64: the grid
 * points, r - R and each exp argument alpha |r - R|^2, the basis constants,
 * and every sum across lanes, point ranges and chunks (FP32 partials added
 * in FP64 in a fixed order, then rounded once into the FP32 outputs).
 * The same semantics as okmi_rho_backward (accumulating outputs, fixed row
 * factors); orbkit has no derivative code (see okmi_rho_backward).
 */
int okmi_rho_backward_f32(const okmi_basis* b, const okmi_grid* g, int64_t p0, int64_t np,
                          const float* d_c, const float* d_occ, int nmo, const float* d_grad,
                          float* d_dc, float* d_dalpha, float* d_dgeo,
                          void* d_ws, size_t ws_bytes, void* stream);

/*
 * Integrated electron count d_out[0] = d3r * sum_p d_rho[p] in FP64 with
 * compensated summation (Neumaier per thread, TwoSum merges in a fixed
 * order; deterministic).  orbkit prints numpy.sum(rho) * d3r
 * (orbkit/core.py:590-592).
 */
size_t okmi_electron_count_workspace_bytes(void);
int okmi_electron_count(const double* d_rho, int64_t n, double d3r, void* d_ws, size_t ws_bytes,
                        double* d_out, void* stream);

/*
 * Gross atomic densities (orbkit/extras.py:306-385) on points [p0, p0+np):
 *   d_out[j][p] = sum_i occ_i phi_i(p) phi_i^(j)(p),
 *   phi_i^(j) = sum_{mu in rows of selection j} C[i][mu] chi_mu,
 * with chi the AOs of derivative code drv (orbkit's drv argument); the AO
 * rows of selection j are sel_rows[sel_start[j] .. sel_start[j+1]) (host
 * arrays; usually the rows of one atom).  Evaluated as sum_{mu in j}
 * chi_mu (D chi)_mu with the density matrix D = C^T diag(occ) C (FP64).
 */
size_t okmi_gross_atomic_workspace_bytes(const okmi_basis* b, int64_t chunk, int nrows_total, int nsel);
int okmi_gross_atomic_density(const okmi_basis* b, const okmi_grid* g, int64_t p0, int64_t np, int drv,
                              const double* d_c, const double* d_occ, int nmo,
                              const int* sel_start, const int* sel_rows, int nsel,
                              double* d_out, int64_t ld_out, void* d_ws, size_t ws_bytes, void* stream);
/*
 * The same from caller-supplied AO values d_ao [nao][ld] and MO values
 * d_mo [nmo][ld] (orbkit's ao_list / mo_list arguments) for one selection
 * of AO rows d_rows (device array): d_out[p] = sum_i occ_i mo_i(p)
 * sum_{r in rows} C[i][r] ao_r(p).  Every row must lie in [0, nao): the
 * rows are read on the device only, so the caller checks them (the Python
 * wrapper raises ValueError otherwise).
 */
int okmi_masked_mo_product(const double* d_c, const double* d_occ, int nmo, int nao, const double* d_ao,
                           const double* d_mo, int64_t np, int64_t ld, const int* d_rows, int nrows,
                           double* d_out, void* stream);
/*
 * MO pair products (orbkit/core.py:935-939, calc_mo_matrix):
 *   d_out[d][n][m][p] = d_bra[n][p] * d_ket[d][m][p]   (row strides ld_in, ld_out)
 */
int okmi_mo_pair_products(const double* d_bra, int na, const double* d_ket, int nb, int nd, int64_t np,
                          int64_t ld_in, double* d_out, int64_t ld_out, void* stream);
/*
 * Transition electronic flux densities (orbkit/extras.py:441-492, calc_jmo):
 *   d_out[d][k][p] = -1/2 (mo[0][i][p] mo[1+d][j][p] - mo[0][j][p] mo[1+d][i][p])
 * for pairs (i, j) = d_pairs[2k], d_pairs[2k+1] (device array, every index
 * in [0, nmo): checked by the caller, as the Python wrapper does); d_mo is
 * [1 + nd][nmo][ld_in]: MO values, then nd derivative sets.
 */
int okmi_transition_flux(const double* d_mo, int nmo, int nd, int64_t np, int64_t ld_in,
                         const int* d_pairs, int npair, double* d_out, int64_t ld_out, void* stream);
/*
 * Text body of a Gaussian cube file as orbkit writes it
 * (orbkit/output/cube.py:88-96); host function.  data [nsets][nx][ny][nz].
 * With out == NULL returns the buffer size needed; else the number of bytes
 * written (0 if cap is too small); the text is NUL-terminated.
 */
size_t okmi_cube_format(const double* data, int nsets, int64_t nx, int64_t ny, int64_t nz, char* out, size_t cap);

/*
 * detCI grid reductions (orbkit/detci/cy_ci.pyx) over MO values
 * d_mo [nmo][ld] (and d_dmo [nd][nmo][ld] for the fluxes), for terms
 * (coef[k], ia[k], ib[k]) given in host arrays (indices checked against nmo):
 *   mode -1: d_out[p]    = sum_k (coef_k mo_a) mo_b                   (get_rho)
 *   mode  0: d_out[d][p] = -sum_k 1/2 coef_k (mo_a dmo_b - mo_b dmo_a) (get_jab)
 *   mode  1: d_out[d][p] = sum_k coef_k mo_a dmo_b                      (get_a_nabla_b)
 * Terms are summed in the given order (orbkit: identical determinants,
 * then single excitations).
 */
size_t okmi_ci_workspace_bytes(int nterm);
int okmi_ci_reduce(const double* d_mo, const double* d_dmo, int nmo, int nd, int64_t npts, int64_t ld,
                   const double* coef, const int* ia, const int* ib, int nterm, int mode,
                   double* d_out, int64_t ld_out, void* d_ws, size_t ws_bytes, void* stream);

/*
 * Many molecules in one launch: density (and gradient) of every molecule on
 * its own grid with the fused reduced-precision kernel (okmi_rho_fused's
 * arithmetic; up to 64 MOs per molecule).  Inputs are bound at creation:
 * per molecule a basis, a whole grid (regular or vector), device pointers
 * to C (nmo x nao, FP64) and occupations, and the outputs d_rho[m] (npts)
 * and, if d_grad is non-NULL and d_grad[m] non-NULL, d_grad[m] (3 x npts,
 * rows x, y, z).  okmi_batch_launch re-reads C and the occupations from
 * those pointers on every launch.  order_by_cost: process molecules
 * costliest first (1) or in the given order (0).
 * orbkit evaluates one molecule per rho_compute call (orbkit/core.py:314);
 * a batch is that call for many molecules at once.
 */
typedef struct OkmiBatch okmi_batch;
int okmi_batch_create(int nmol, okmi_basis* const* bases, const okmi_grid* grids,
                      const double* const* d_c, const double* const* d_occ, const int* nmo,
                      double* const* d_rho, double* const* d_grad, int precision, int order_by_cost,
                      okmi_batch** out);
/* tiles of 64 points over all molecules */
int64_t okmi_batch_tiles(const okmi_batch* b);
/* persistent workgroup count (0: one workgroup per tile; default 0) */
int okmi_batch_set_workgroups(okmi_batch* b, int workgroups);
/* brick of 2^log2_x x 2^log2_y x 2^log2_z regular-grid points per tile
 * (sum 6; default 2, 2, 2; vector grids always take 64 consecutive points) */
int okmi_batch_set_tile_shape(okmi_batch* b, int log2_x, int log2_y, int log2_z);
/* enqueue the whole batch on `stream`; use_graph = 1 captures the launch
 * sequence into a HIP graph on first use (per stream; `stream` must not be
 * the NULL stream) and replays it afterwards */
int okmi_batch_launch(okmi_batch* b, int use_graph, void* stream);
int okmi_batch_destroy(okmi_batch* b);

/*
 * Kernel timing for measurement tools: while enabled, every kernel launch
 * records a pair of HIP events on its stream.  okmi_profile_query sums the
 * elapsed times of the launches of one kernel family ("ao_eval", "gemm_f64",
 * "rho_f64", "rho_fused", "rho_fused_f64", "rho_big", "rho_batch"; the
 * backward's "rho_backward", "gemm_nn", "gemm_tn", "shell_grad" and their
 * FP32 forms "rho_backward_f32", "gemm_nn_f32", "gemm_tn_f32",
 * "shell_grad_f32") since the last okmi_profile_enable call.
 * No counterpart in the reference (orbkit has no kernel timers).
 */
int okmi_profile_enable(int on);
int okmi_profile_query(const char* kernel, int* count, double* total_ms);

/*
 * Work measure for measurement tools: the fused kernels (and the many-MO and
 * staged paths) contract, for each 4x4x4 brick of a regular grid, only the
 * 32-row AO chunks with a shell whose screening radius reaches the brick.
 * For the regular grid with host axes x, y, z this gives *rows_points = sum
 * over bricks of (AO rows of its live chunks) x (grid points in the brick);
 * the screened contraction's flops are 2 * NMO * sets * *rows_points.
 * prec: OKMI_PREC_F64 for the FP64 paths' radii (terms below 1e-24), any
 * other precision for the radii of the paths whose AO values end in FP32
 * (terms below 1e-12).
 * No counterpart in the reference (orbkit evaluates every AO everywhere).
 */
int okmi_fused_live_rows(const okmi_basis* b, const double* x, int64_t nx, const double* y, int64_t ny,
                         const double* z, int64_t nz, int prec, double* rows_points);

#ifdef __cplusplus
}
#endif
#endif
