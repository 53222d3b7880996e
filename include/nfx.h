/*
 * nfx.h - C ABI of libnfx, the MI355X (gfx950) assembly-and-solve library behind
 * the nfempy_amd drop-in for nfempy's weak-form / product-space / assembler path.
 *
 * Conventions
 *  - Every entry point returns an int status: 0 on success, a negative NFX_E_*
 *    code otherwise; nfx_last_error() returns a thread-local message. Nothing
 *    throws across this boundary.
 *  - Pointer arguments named d_* are device pointers (HBM); h_* are host pointers.
 *    Sizes are plain integers. Work is enqueued on the stream passed in (a
 *    hipStream_t, NULL = default stream) and is asynchronous unless the function
 *    says otherwise.
 *  - Indices are int32; values are FP32 (float) or FP64 (double) as selected by
 *    an fp64 flag.
 *  - Element kernels compute element matrices in the REFERENCE basis order
 *    (field by field; inside a field basis function a, component c -> a*n_comp+c).
 *    The reference's facet orientation handling (src/basis/permute_and_transform.py:8-149:
 *    a signed permutation of facet DoFs per element) is folded by the host into the
 *    destination codes, so kernels never branch on orientation.
 *
 * Slot / row codes (int32), one per element-local matrix entry or residual entry:
 *    code < 0        dropped
 *    bits 0..27      CSR value index (or stage index, or row)
 *    bit 28          contribute 0.0 (explicit structural zero)
 *    bit 29          index is into the stage array (multi-contributor slot)
 *    bit 30          contribute the negated value (orientation sign)
 *
 * Each entry point cites the reference interface it replaces (paths relative to
 * the OmarDuran/nfempy repository).
 */
#ifndef NFX_H
#define NFX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NFX_OK 0
#define NFX_E_INVALID -1       /* bad argument: size, pointer, shape */
#define NFX_E_HIP -2           /* a HIP runtime call failed */
#define NFX_E_UNSUPPORTED -3   /* element type / size not compiled in */
#define NFX_E_NOT_CONVERGED -4 /* solver stopped above the requested tolerance */

/* ---- library ------------------------------------------------------------ */

/* Version as 10000*major + 100*minor + patch (0.2.0: per-element metric argument of the
 * Darcy row assembly, slot_pairs of the elasticity assembly, the block-LU tail of
 * nfx_twogrid_pc, block operator of the MINRES solver; 0.3.0: tile launch map of the Darcy
 * row assembly, packed positions and entry signs of the lowest-order rows; 0.4.0: in-tile
 * row offsets in the lowest-order position words, compact RT2 batched launch, degree 3;
 * 0.5.0: flux-row operator fields of nfx_twogrid_pc (flux_width, flux_col, flux_coef, appended
 * after tail_coef) and nfx_flux_sweep_coef). */
int nfx_version(void);
/* Thread-local description of the last error on this thread. */
const char* nfx_last_error(void);
/* Number of visible devices. */
int nfx_device_count(int* h_count);

/* ---- assembly: mixed Darcy (RT_k - P_{k-1}) ----------------------------------
 * Replaces LaplaceDualWeakForm.evaluate_form
 * (src/weak_forms/laplace_dual_weak_form.py:11-87) evaluated for every element,
 * together with the Darcy scatter of examples/md_elliptic_scalar/md_elliptic_scalar.py:250-265
 * (src/assembly/ScatterFormData.py:5-25), i.e. the element loop of
 * SequentialAssembler.scatter_forms (src/assembly/SequentialAssembler.py:69-86).
 *
 *  d_cells  (E, dim+1) int32 vertex ids
 *  d_coords (V, 3) vertex coordinates (float or double as fp64)
 *  d_tables reference tables (float or double, matching fp64):
 *           w[nq] | phi[nq][nb][dim] | div[nq][nb] | psi[nq][np]
 *  d_kinv, d_f  inverse permeability and source at quadrature points (float or
 *           double as fp64), element e point q at index e*stride_e + q*stride_q
 *           (stride_q = 0: one value per element; both 0: one value)
 *  d_slot   (nl*nl, E) slot codes, entry-major, nl = nb + np
 *  d_rhs_row (np, E) row codes of the pressure residual
 *  outputs  d_vals (CSR values), d_stage (staged contributions), d_rhs (rg),
 *           float or double as fp64.
 * Supported: (dim, k) in {2, 3} x {1, 2}: RT1-P0 and RT2-P1.
 */
int nfx_assemble_darcy(int dim, int k, int n_elem, const int32_t* d_cells,
                       const void* d_coords, const void* d_tables,
                       const void* d_kinv, int kinv_stride_e, int kinv_stride_q,
                       const void* d_f, int f_stride_e, int f_stride_q,
                       const int32_t* d_slot, const int32_t* d_rhs_row, int fp64,
                       void* d_vals, void* d_stage, void* d_rhs, void* stream);

/* Row-owner variant of nfx_assemble_darcy (same reference interfaces): one thread
 * per CSR row gathers row a of each incident element matrix (at most two
 * incidences, element order) and the block stores its contiguous slice of CSR
 * values; the residual of every row is written (pressure rows: -sum w detJ f psi,
 * flux rows: 0, before the Dirichlet facet term). No slot map, no staging.
 *  d_inc    (2, n_rows) int32 incidences e*32 + a (reference DoF a), negative = none
 *  d_pos    column position of reference DoF b inside the row, per (row, incidence):
 *           nl <= 6 (lowest-order elements): one uint32, bits 4b..4b+3 the position
 *           (15 = not present), bit 4 nl + b set when entry (a, b) is negated (sign bit a
 *           xor sign bit b of d_signs); for nl <= 5 bits 25..31 of the row's two words hold
 *           its CSR offset rowptr[r] - rowptr[128 (r / 128)] (low 7 bits in the first word,
 *           high 7 in the second), read instead of d_rowptr[r]; other bits zero; otherwise nlp uint8 (255 = not
 *           present), nlp = nl, or nl rounded up to a multiple of 4 when nl > 8 (RT2
 *           elements: word loads)
 *  d_rowptr CSR row pointer; d_signs (E) uint32, bit a = reference DoF a negated (read
 *           for nl > 6 only)
 *  d_geo    optional per-element metric from nfx_element_geometry (same precision as
 *           the values): the kernel then reads it instead of deriving it from d_cells /
 *           d_coords (same values); NULL derives it; 16-byte aligned
 *  d_tile_map optional (n_tile_map) int32 launch order of the row tiles (tile t = CSR
 *           rows [128 t, 128 t + 128)): slot i holds tile d_tile_map[i], -1 (or any value
 *           outside the tile range) an empty slot; every tile must appear once. Workgroups
 *           go to the 8 XCDs round-robin, so slots x, x + 8, ... run on XCD x: a map that
 *           gives each XCD the tiles of one element range keeps each element's data in one
 *           L2. NULL = tiles in row order. Results do not depend on the map.
 * Other arguments as nfx_assemble_darcy. */
int nfx_assemble_darcy_rows(int dim, int k, int n_rows, const int32_t* d_inc,
                            const uint8_t* d_pos, const int32_t* d_rowptr,
                            const uint32_t* d_signs, const int32_t* d_cells,
                            const void* d_coords, const void* d_tables,
                            const void* d_kinv, int kinv_stride_e, int kinv_stride_q,
                            const void* d_f, int f_stride_e, int f_stride_q, const void* d_geo,
                            const int32_t* d_tile_map, int n_tile_map, int fp64, void* d_vals,
                            void* d_rhs, void* stream);

/* Per-element metric of the row kernels, computed once per mesh (the mapping of
 * src/geometry/mapping.py evaluated on the vertices): for element e, geo[e * gs + s] =
 * (J^T J)[d][d2] / |det J| over d <= d2 (row-major upper order), then |det J| (for a
 * 2-D element in 3-D, sqrt(det J^T J)), zero padding; gs = 8 (dim 3) or 4 (dim 2).
 * float or double by fp64. */
int nfx_element_geometry(int dim, int n_elem, const int32_t* d_cells, const void* d_coords, int fp64,
                         void* d_geo, void* stream);

/* Batched persistent variant for many independent 3-D systems (RT1-P0 and RT2-P1
 * mixed in one launch): d_items is an array of n_items descriptors
 *   struct { int k, n_rows, ks_e, ks_q, fs_e, fs_q;
 *            const int32_t* inc; const uint8_t* pos; const int32_t* rowptr;
 *            const uint32_t* signs; const int32_t* cells; const float* coords;
 *            const float* tables; const float* kinv; const float* f;
 *            float* vals; float* rhs;
 *            const float* geo; }      (FP32, nfx_darcy_batch_item_size bytes; geo, the
 *                                      nfx_element_geometry metric, is required)
 * and d_tiles (n_tiles, 2) int32 (item, first row) pairs, heaviest first; a fixed
 * grid of `grid` workgroups of tile_rows (64, 128, 256) threads walks the tiles
 * in a static order (block b: tiles b, b + grid, ...). Same per-row arithmetic as
 * nfx_assemble_darcy_rows. A workgroup keeps the reference table in LDS while
 * consecutive tiles carry the same tables pointer (give all instances of one degree
 * one table to benefit; any pointers are correct). degree = 0: tiles of both
 * degrees; degree = 1 or 2: every tile is of that degree (tiles of the other degree
 * are skipped) and the kernel is specialised for it (RT1: 8 waves per SIMD).
 * degree = 3: RT2 tiles whose items have per-element coefficients (kinv and f with
 * stride_q = 0), run with only the reference tensors in LDS and a staging slice of 33
 * entries per tile row (4 waves per SIMD): a tile's item field is item | rows << 16
 * (rows <= tile_rows; 0 = tile_rows) and its CSR span must be at most 33 * tile_rows
 * (other tiles are skipped). */
int nfx_assemble_darcy_batched(int n_items, const void* d_items, int n_tiles,
                               const int32_t* d_tiles, int tile_rows, int grid, int degree,
                               void* stream);
int nfx_darcy_batch_item_size(void);

/* Dirichlet pressure facet term, LaplaceDualWeakFormBCDirichlet.evaluate_form
 * (src/weak_forms/laplace_dual_weak_form.py:91-154) for every boundary facet:
 * rg[row] = sum_q w |detJ_f| p_D(x_q) psi(x_q) . n with the flux basis traced from
 * the neighbour cell and n the unit outward normal (src/geometry/compute_normal.py:7-24).
 *  d_facet (F, dim) facet vertex ids in stored order; d_neigh (F) neighbour element;
 *  d_combo (F) index of (local facet, facet vertex order) into d_trace;
 *  d_trace (ncombo, nqf, fd, dim) reference flux basis of the facet DoFs at the facet
 *  points mapped into the neighbour (src/basis/parametric_transformation.py:30-98);
 *  d_fw (nqf) facet weights; d_pD (F, nqf) Dirichlet data (float or double as
 *  fp64); d_rows (F, fd) row codes.
 * Each boundary facet owns its rows (element kernels never write them).
 */
int nfx_assemble_darcy_bc(int dim, int k, int n_facets, const int32_t* d_facet,
                          const int32_t* d_neigh, const int32_t* d_combo,
                          const int32_t* d_cells, const void* d_coords,
                          const void* d_trace, const void* d_fw, int nqf,
                          const void* d_pD, const int32_t* d_rows, int fp64,
                          void* d_rhs, void* stream);

/* The whole warm assembly in one launch: nfx_assemble_darcy_rows followed by
 * nfx_assemble_darcy_bc (same reference interfaces: the element loop of
 * LaplaceDualWeakForm plus the boundary loop of LaplaceDualWeakFormBCDirichlet in
 * the Darcy driver), with the facet term evaluated by the thread of the row it
 * belongs to. A row owned by Dirichlet facet f (facet DoF a) carries second
 * incidence -2 - (f * fd + a) in d_inc (such a row has one element incidence, and
 * it is the facet's neighbour); d_rows (F, fd) are its row codes. Results equal
 * the two-launch sequence. n_facets = 0: same as nfx_assemble_darcy_rows. */
int nfx_assemble_darcy_rows_bc(int dim, int k, int n_rows, const int32_t* d_inc,
                               const uint8_t* d_pos, const int32_t* d_rowptr,
                               const uint32_t* d_signs, const int32_t* d_cells,
                               const void* d_coords, const void* d_tables, const void* d_kinv,
                               int kinv_stride_e, int kinv_stride_q, const void* d_f,
                               int f_stride_e, int f_stride_q, int n_facets,
                               const int32_t* d_facet, const int32_t* d_combo,
                               const void* d_trace, const void* d_fw, int nqf, const void* d_pD,
                               const int32_t* d_rows, const void* d_geo,
                               const int32_t* d_tile_map, int n_tile_map, int fp64, void* d_vals,
                               void* d_rhs, void* stream);

/* ---- assembly: dual-mixed elasticity (BDM1 - P0 - P0, tetrahedra) --------
 * mode 0 replaces LEDualWeakForm.evaluate_form_vectorized
 * (src/weak_forms/le_dual_weak_form.py:354-510), mode 1
 * LERieszMapWeakForm.evaluate_form_vectorized (src/weak_forms/le_riesz_map_weak_form.py:258-360),
 * each with the SBAIJ scatter of examples/cauchy_elasticity/mfe_elasticity_example_1.py:128-192
 * expressed by the slot codes (dropped entries = -1, explicit zero diagonals = bit 28).
 *  d_tables  w[8] | phi[8][12][3] | div[8][12]  (BDM1 reference basis)
 *  d_lam, d_mu  Lame parameters at quadrature points (stride_e, stride_q)
 *  d_f       body force + rotation source, 6 components: f[e*fs_e + q*fs_q + c*fs_c]
 *  d_slot    slot codes, element-major: slot_pairs = 0: (E, 42*42), entry (I, J) of
 *            element e at e*1764 + 42 I + J; slot_pairs = 1: (E, 903), one code per
 *            upper-triangle entry I <= J (row-major) receiving the symmetric value
 *            (an upper SBAIJ operator keeps at most one of (I, J), (J, I));
 *            d_rhs_row (6, E) rows of r(u_c), r(t_c)
 * One wavefront per element. */
int nfx_assemble_elasticity(int mode, int n_elem, const int32_t* d_cells,
                            const void* d_coords, const void* d_tables,
                            const void* d_lam, const void* d_mu, int coef_stride_e,
                            int coef_stride_q, const void* d_f, int f_stride_e,
                            int f_stride_q, int f_stride_c, const int32_t* d_slot, int slot_pairs,
                            const int32_t* d_rhs_row, int fp64, void* d_vals,
                            void* d_stage, void* d_rhs, void* stream);

/* The same operators (A: mode 0 with the residual, P: mode 1), facet-owner form, for
 * tetrahedra (dim 3) and triangles (dim 2): replaces the element contributions of
 * LEDualWeakForm.evaluate_form_vectorized and LERieszMapWeakForm.evaluate_form_vectorized
 * (src/weak_forms/le_dual_weak_form.py:354-510, src/weak_forms/le_riesz_map_weak_form.py:258-360)
 * together with their SBAIJ scatter (examples/cauchy_elasticity/mfe_elasticity_example_1.py:128-192).
 * One wavefront owns the CSR rows of one facet (a face: 3 BDM1 DoFs x 3 stress rows, rows
 * 9 p .. 9 p + 8; an edge: 2 x 2, rows 4 p .. 4 p + 3), sums the rows of its (at most two)
 * incident elements in element order and stores the contiguous CSR slice; cell rows
 * (u: n_s + NC e + c, t: n_s + n_u + NT e + c; NC = dim, NT = 3 or 1) hold only their diagonal
 * (0 in A, the mass in P). n_s = (9 | 4) n_faces, n_u = NC n_elem.
 *   d_face_inc   (n_faces, 2) incident elements in element order (-1 none)
 *   d_face_lf    (n_faces) local facet index of each incidence (bits 0-1, 2-3)
 *   d_el_face    (n_elem, dim + 1) DoF block of each local facet
 *   d_el_orient  (n_elem) uint64: bits 2b, 2b+1 facet slot of reference basis b, 32+b its sign
 *   d_geom       (n_elem, dim dim + 1): J / |det J| row-major, then |det J| (J[c][d] = X_{d+1,c} - X_{0,c})
 *   d_tables     w[NQ] | phi[NQ][NB][dim] | divref[NQ][NB] | Uref[NB] | Phi[NB][dim] |
 *                DDref[NB][NB]: sum_q w divref_b, sum_q w phihat_be, sum_q w divref_a divref_b
 *   d_rowptr     CSR row pointer of the operator (stress rows first)
 *   d_codes      uint16 per CSR entry of the stress rows: bits 0-3 row in the facet block
 *                (j NC + c), 4-9 / 10-15 element-local reference column in the first / second
 *                incident element (63: no contribution)
 * Coefficients as nfx_assemble_elasticity (f has NC + NT components); d_rhs gets every row
 * (0 on stress rows) in mode 0. */
int nfx_assemble_elasticity_faces(int dim, int mode, int n_faces, int n_elem,
                                  const int32_t* d_face_inc, const int32_t* d_face_lf,
                                  const int32_t* d_el_face, const uint64_t* d_el_orient,
                                  const void* d_geom, const void* d_tables,
                                  const void* d_lam, const void* d_mu, int coef_stride_e,
                                  int coef_stride_q, const void* d_f, int f_stride_e, int f_stride_q,
                                  int f_stride_c, const int32_t* d_rowptr, const uint16_t* d_codes,
                                  int fp64, void* d_vals, void* d_rhs, void* stream);

/* Block preconditioner of the elasticity system from its full FP64 CSR (n = n_s + 2 n_u:
 * stress rows, then u, then t): the inverse of each 9 x 9 face block of the stress rows
 * (rows 9b .. 9b+8, columns of the same block) and of each 6 x 6 cell block of the Schur
 * complement B D^-1 B^T on (u, t) (B the stress columns of the cell's u rows n_s + 3c + k
 * and t rows n_s + n_u + 3c + k), FP64 Gauss-Jordan, symmetrised. The field-split
 * preconditioner of mfe_elasticity_example_1.py:238-274 solves with the Riesz map
 * exactly; this is its block-diagonal stand-in. Output: FP32 values of the CSR with 9
 * entries per stress row (row 9b + i: columns 9b .. 9b+8) then 6 per (u, t) row (the
 * cell's u then t DoFs); d_dinv (n_s / 9 x 81 FP64) workspace. A non-positive pivot sets
 * bits of *d_status (1 face block, 2 more than 4 faces per cell, 4 cell block). */
int nfx_elasticity_mass_schur_pc(int n, int n_s, int n_u, const int32_t* d_rowptr, const int32_t* d_col,
                                 const double* d_vals64, double* d_dinv, float* d_pc_val, int* d_status,
                                 void* stream);

/* Stress Dirichlet term of LEDualWeakFormBCDirichlet.evaluate_form
 * (src/weak_forms/le_dual_weak_form.py:514-578):
 * r[(a, c)] = -sum_q w |detJ_f| u_D,c(x_q) psi_a(x_q) . n for the dim BDM1 DoFs of each
 * boundary facet; dim 3: d_uD (F, 4, 3), d_rows (F, 9); dim 2: d_uD (F, 2, 2), d_rows (F, 4)
 * row codes (entry a * dim + c). Other arguments as nfx_assemble_darcy_bc. */
int nfx_assemble_elasticity_bc(int dim, int n_facets, const int32_t* d_facet,
                               const int32_t* d_neigh, const int32_t* d_combo,
                               const int32_t* d_cells, const void* d_coords,
                               const void* d_trace, const void* d_fw, int nqf,
                               const void* d_uD, const int32_t* d_rows, int fp64,
                               void* d_rhs, void* stream);

/* dst[i] = src[d_idx[i]] (float or double): the full symmetric matrix from the
 * SBAIJ upper triangle for the Krylov solver. */
int nfx_gather(int fp64, int64_t n, const int32_t* d_idx, const void* d_src, void* d_dst,
               void* stream);

/* ---- scatter reduction ---------------------------------------------------
 * Deterministic segmented sum of staged contributions into CSR slots - the
 * multi-contributor part of the reference's serial setValue(addv=True) scatter
 * (src/assembly/ScatterFormData.py:5-11): for segment m,
 *   vals[d_target[m]] = sum of stage[d_segptr[m] .. d_segptr[m+1]) in element order. */
int nfx_stage_reduce(int n_seg, const int32_t* d_segptr, const int32_t* d_target,
                     int fp64, const void* d_stage, void* d_vals, void* stream);

/* Fill a float / double array with a value (a kernel, safe under graph capture). */
int nfx_fill(int fp64, int64_t n, double value, void* d_x, void* stream);

/* ---- geometry -------------------------------------------------------------
 * Physical quadrature points x_q = sum_a bary[q][a] X_a of every element in FP64:
 * the points at which the reference evaluates its coefficient callables
 * (src/geometry/mapping.py:39-41). d_out (E, nq, 3). */
int nfx_quadrature_points(int dim, int n_elem, int nq, const int32_t* d_cells,
                          const double* d_coords64, const double* d_bary,
                          double* d_out, void* stream);

/* ---- sparse kernels --------------------------------------------------------
 * y = alpha A x + beta y, CSR FP32 (PETSc MatMult inside the reference's KSP,
 * examples/md_elliptic_scalar/md_elliptic_scalar.py:351-364). */
int nfx_spmv(int n_rows, const int32_t* d_rowptr, const int32_t* d_col,
             const float* d_val, float alpha, const float* d_x, float beta,
             float* d_y, void* stream);

/* FP64 residual r = b - A x with FP64 values (the refinement residual). */
int nfx_residual64(int n_rows, const int32_t* d_rowptr, const int32_t* d_col,
                   const double* d_val, const double* d_x, const double* d_b,
                   double* d_r, void* stream);

/* nfx_residual64 without any allocation: no counterpart in the reference; for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). The residual kernel's partial sums go to a static device
 * array of 2 x 2048 doubles (32 KiB per GPU, held until the process ends); the kernel grid is
 * capped at 2048 blocks, so any n_rows fits. Same outputs as nfx_residual64, bit for bit.
 * Calls that overlap on one GPU share that array and race. */
int nfx_residual64_noalloc(int n_rows, const int32_t* d_rowptr, const int32_t* d_col,
                           const double* d_val, const double* d_x, const double* d_b,
                           double* d_r, void* stream);

/* FP64 row diagonal of a CSR matrix, and the FP64 Schur-complement diagonal
 * S[k - n_u] = sum_{j < n_u} A[k, j] A[j, k] / A[j, j] for rows k >= n_u (the
 * pressure part of the block preconditioner; pattern must be structurally
 * symmetric with sorted columns). */
int nfx_csr_diag64(int n_rows, const int32_t* d_rowptr, const int32_t* d_col,
                   const double* d_val, double* d_diag, void* stream);
int nfx_schur_diag64(int n_rows, int n_u, const int32_t* d_rowptr,
                     const int32_t* d_col, const double* d_val,
                     const double* d_diag, double* d_schur, void* stream);

/* ---- Krylov --------------------------------------------------------------
 * Preconditioned MINRES (Paige-Saunders) in FP32 on S = D A, D = diag(d_sign)
 * (-1 on rows that turn the saddle point symmetric; NULL = identity), with an SPD
 * preconditioner given as a sparse matrix P^-1 (CSR d_pc_*, FP32: a diagonal, or
 * the block-diagonal inverse of a field split), inside FP64 iterative refinement
 * on the as-assembled FP64 matrix. Replaces KSP minres +
 * PCFIELDSPLIT (examples/cauchy_elasticity/mfe_elasticity_example_1.py:238-274)
 * and KSP preonly + LU/MUMPS (examples/md_elliptic_scalar/md_elliptic_scalar.py:351-364)
 * by an iterative solve to ||b - A x||_2 / ||b||_2 <= rtol evaluated in FP64.
 * Synchronous. h_info[0] = refinement steps, h_info[1] = total MINRES iterations;
 * h_resid[0] = final relative FP64 residual; h_hist (optional, hist_cap values) =
 * relative residual before each refinement step. Returns NFX_E_NOT_CONVERGED if
 * rtol is not reached within max_outer refinement steps.
 * A solver object owns its Krylov workspace (11 FP32 vectors + 1 FP64 vector).
 */
typedef struct nfx_solver nfx_solver;
int nfx_solver_create(nfx_solver** out, int n_rows, int nnz, void* stream);
int nfx_solver_destroy(nfx_solver* s);
/* Give the solver a 3 x 3 block CSR copy of the FP32 operator it is then called with
 * (n_rows a multiple of 3; block row R holds rows 3R .. 3R+2, d_browptr (n_rows/3 + 1),
 * d_bcol (n_blocks) block columns, d_bval 9 n_blocks values with entry (a, b) of block j
 * at (3a + b) n_blocks + j): the inner MINRES multiplies with it (one index and three
 * consecutive vector loads per block); the FP64 refinement keeps the CSR values. n_blocks
 * = 0 or null pointers: back to the CSR operator. The arrays must outlive the solves. */
int nfx_solver_set_block3(nfx_solver* s, int n_blocks, const int32_t* d_browptr, const int32_t* d_bcol,
                          const float* d_bval);
int nfx_solve_minres_ir(nfx_solver* s, const int32_t* d_rowptr, const int32_t* d_col,
                        const float* d_val32, const double* d_val64,
                        const float* d_sign, const int32_t* d_pc_rowptr,
                        const int32_t* d_pc_col, const float* d_pc_val,
                        const double* d_b, double* d_x,
                        double rtol, double inner_rtol, int max_inner, int max_outer,
                        int* h_info, double* h_resid, double* h_hist, int hist_cap);

/* Two-grid block preconditioner of the dual-mixed elasticity operator (stress rows first,
 * 9 per face; then u rows n_s + 3c + k and t rows n_s + n_u + 3c + k of cell c), the
 * GPU replacement of the reference's fieldsplit preconditioner built on the Riesz map
 * (exact Cholesky of its stress block, ICC on the rest:
 * examples/cauchy_elasticity/mfe_elasticity_example_1.py:238-269). P^-1 = diag(D^-1,
 * psign TG(S)): D^-1 the inverted 9 x 9 face blocks of the compliance block (FP32, 81 per
 * face), S = B D^-1 B^T on the cells (6 x 6 blocks: the cell's own, then one per local face
 * neighbour), TG(S) one symmetric two-grid cycle with the block-Jacobi smoother W = omega
 * S_cc^-1 and piecewise-constant aggregation with the dense coarse inverse ainv ((6 n_agg)^2
 * FP32, row-major). Cell data are in aggregate order: position m holds cell agg_mem[m], the
 * positions of aggregate A are agg_ptr[A] .. agg_ptr[A+1]-1, agg_of[m] its aggregate,
 * nbr[4m+k] the position of the neighbour across local face k (-1 none); sblk holds entry e
 * of block b (0: own, 1 + k: neighbour k) of position m at (36 b + e) n_cells + m, winv
 * entry e at e n_cells + m (the layout nfx_elasticity_tg_blocks writes). work: 12 n_cells +
 * 12 n_agg floats, 16-byte aligned. psign = +1 gives an SPD preconditioner (MINRES). */
typedef struct nfx_elasticity_tg {
  int n_s, n_u, n_cells, n_agg;
  const float* dinv;
  const int32_t* nbr;
  const float* sblk;
  const float* winv;
  const int32_t* agg_ptr;
  const int32_t* agg_mem;
  const int32_t* agg_of;
  const float* ainv;
  float psign;
  float* work;
} nfx_elasticity_tg;
int nfx_elasticity_tg_apply(const nfx_elasticity_tg* pc, const float* d_r, float* d_z, void* stream);
/* Set-up of the blocks (FP64), cells in mesh order: for cell c and local face k (face
 * d_efac[4c+k], neighbour cell d_nbr[4c+k] or -1, the neighbour's local index of that face
 * d_nbrk[4c+k]) with B_ck the 6 x 9 block of the (u, t) x stress coupling (d_bblk, 54 per
 * (c, k), coefficient-independent) and d_dinv the FP64 face-block inverses: d_S (n_cells x
 * 5 x 36, cell order) = S blocks; d_S32 and d_W32 = omega S_cc^-1 in the entry-major
 * aggregate-order layout of nfx_elasticity_tg (cell c at position d_pos[c]); d_part
 * (n_cells x 4 x 36) scratch. Bit 0 of *d_status is set when a cell block is not SPD. */
int nfx_elasticity_tg_blocks(int n_cells, const int32_t* d_efac, const int32_t* d_nbr,
                             const int32_t* d_nbrk, const double* d_bblk, const double* d_dinv,
                             const int32_t* d_pos, double* d_part, double* d_S, float* d_S32,
                             float* d_W32, double omega, int* d_status, void* stream);
/* d_Ac += P^T S P (FP64, (6 n_agg)^2 row-major; zero it first), deterministic; cells in
 * mesh order: d_agg_mem[agg_ptr[A] ..] the cells of aggregate A, d_agg_of / d_nbr per cell
 * (d_S as nfx_elasticity_tg_blocks writes it). */
int nfx_elasticity_tg_coarse(int n_agg, const int32_t* d_agg_ptr, const int32_t* d_agg_mem,
                             const int32_t* d_agg_of, const int32_t* d_nbr, const double* d_S,
                             double* d_Ac, void* stream);
/* nfx_solve_minres_ir with the elasticity two-grid preconditioner (applied as a function,
 * psign = +1) in place of a sparse P^-1; the solver's 3 x 3 block copy of the operator is
 * used when set. */
int nfx_solve_minres_ir_tg(nfx_solver* s, const int32_t* d_rowptr, const int32_t* d_col,
                           const float* d_val32, const double* d_val64, const nfx_elasticity_tg* pc,
                           const double* d_b, double* d_x, double rtol, double inner_rtol,
                           int max_inner, int max_outer, int* h_info, double* h_resid, double* h_hist,
                           int hist_cap);

/* Batched MINRES: many independent systems (the batched configuration's instances, each
 * replacing one KSP solve of examples/md_elliptic_scalar/md_elliptic_scalar.py:351-364), one
 * set of launches per iteration. Each instance has its own CSR (local column indices), FP32
 * and FP64 values, optional row signs D (the symmetric form D A) and diagonal preconditioner
 * (FP32 inverse diagonal, NULL = identity), its own MINRES recurrence, stopping test and FP64
 * refinement to ||b - A x|| / ||b|| <= rtol; vectors are concatenated in instance order
 * (row_off is filled in by nfx_minres_batch_create). Partial sums are reduced per instance in
 * a fixed order: results do not depend on the other instances of the batch. */
typedef struct nfx_minres_batch_item {
  int n_rows;
  int lanes; /* filled in by nfx_minres_batch_create (lanes per row of the SpMV) */
  int64_t row_off;
  const int32_t* rowptr;
  const int32_t* col;
  const float* val32;
  const double* val64;
  const float* sign;
  const float* pinv;
} nfx_minres_batch_item;
typedef struct nfx_minres_batch nfx_minres_batch;
int nfx_minres_batch_create(nfx_minres_batch** out, int n_inst, const nfx_minres_batch_item* h_items,
                            void* stream);
int nfx_minres_batch_destroy(nfx_minres_batch* s);
/* Total rows of the batch (the length of d_b and d_x); a count, not a status. */
int64_t nfx_minres_batch_rows(const nfx_minres_batch* s);
/* Synchronous. Each refinement step of an instance stops its inner MINRES at the relative
 * reduction still missing, max(inner_rtol, rtol / (2 ||r|| / ||b||)) (at most 0.5). Per
 * instance: h_iters total MINRES iterations, h_outer refinement steps, h_resid final relative
 * FP64 residual. NFX_E_NOT_CONVERGED if any instance stays above rtol. */
int nfx_solve_minres_batch_ir(nfx_minres_batch* s, const double* d_b, double* d_x, double rtol,
                              double inner_rtol, int max_inner, int max_outer, int* h_iters, int* h_outer,
                              double* h_resid);

/* Two-grid Schur preconditioner of the pressure rows of a group of Darcy instances of one
 * degree in the batched MINRES (RT1-P0: nb = 1 pressure unknown per cell; RT2-P1: nb = 4):
 * the learned / two-grid replacement of the Schur diagonal of the block preconditioner
 * (no reference counterpart; the solve it preconditions replaces
 * examples/md_elliptic_scalar/md_elliptic_scalar.py:351-364 per instance). Cells of one
 * instance are contiguous; cell c's unknowns sit at rows prow[c nb + i] of the stacked batch
 * vector, cell_inst[c] is its instance (its index in the batch, for the mode gate); nbr[4c+k]
 * the neighbour across local face k (-1 none); sd (n_cells, nb, nb) and so (n_cells, 4, nb,
 * nb) the blocks of S = B diag(M)^-1 B^T, rs (n_cells, 5, nb) their row sums (own block,
 * then neighbour k; nfx_batch_tg_rowsums); wd (n_cells, nb, nb) the smoother blocks, wo
 * (n_cells, 4) its couplings (nb = 1 only, NULL = none); aggregates (agg_ptr / agg_mem /
 * agg_of) never span two instances; ainv holds every instance's dense coarse inverse, row
 * a of a's instance at ainv + arow[a] with alen[a] entries for the aggregates a0[a] ..
 * a0[a] + alen[a] - 1. work: 4 n_cells nb + 2 n_agg floats. The cycle
 * z = x2 + W (r - S x2), x2 = x1 + P Ac^-1 P^T (r - S x1), x1 = W r writes z on the
 * group's pressure rows; with flux blocks (n_fb > 0) z = D_g^-1 r on their flux rows too. */
typedef struct nfx_batch_tg {
  int nb, n_cells, n_agg;
  const int32_t* cell_inst;
  const int32_t* prow;
  const int32_t* nbr;
  const float* sd;
  const float* so;
  const float* rs;
  const float* wd;
  const float* wo;
  const int32_t* agg_ptr;
  const int32_t* agg_mem;
  const int32_t* agg_of;
  const int64_t* arow;
  const int32_t* a0;
  const int32_t* alen;
  const float* ainv;
  float* work;
  int n_fb;                /* flux blocks (3 x 3) replacing the diagonal on flux rows, or 0 */
  const int32_t* fb_inst;  /* (n_fb) instance of each block */
  const int32_t* fb_rows;  /* (n_fb, 3) rows in the stacked vector */
  const float* fb_inv;     /* (n_fb, 3, 3) inverse blocks */
} nfx_batch_tg;
/* S blocks from each instance's FP64 CSR values (d_vals[cell_inst[c]]): d_posb (n_cells,
 * nbf, nb) positions of B(pressure unknown i, flux DoF a) of the cell's nbf flux DoFs, d_posd
 * (n_cells, nbf) positions of their flux diagonals, d_fdofs (4, fd) the local flux DoFs of
 * each local face, d_nloc (n_cells, 4, fd) their local index in the neighbour. FP64 and
 * FP32 outputs in the layout of nfx_batch_tg (for nb = 1 the FP64 ones are the sdiag /
 * soff of nfx_gnn_features). */
int nfx_batch_tg_schur(int nb, int n_cells, int nbf, int fd, const int32_t* d_cell_inst, const double* const* d_vals,
                       const int32_t* d_posb, const int32_t* d_posd, const int32_t* d_nbr, const int32_t* d_fdofs,
                       const int32_t* d_nloc, double* d_sd64, double* d_so64, float* d_sd, float* d_so,
                       void* stream);
/* RT2-P1 variant of nfx_batch_tg_schur with the flux mass in 3 x 3 blocks (a face's three
 * DoFs, a cell's three interior DoFs, local DoFs 12..14) instead of its diagonal: S = B D^-1
 * B^T with d_posmf (n_cells, 4, 3, 3) / d_posmi (n_cells, 3, 3) the CSR positions of the blocks
 * in the cell's local DoF order; bit 1 of *d_status when a block is not SPD. */
int nfx_batch_tg_schur_blk(int n_cells, const int32_t* d_cell_inst, const double* const* d_vals,
                           const int32_t* d_posb, const int32_t* d_posmf, const int32_t* d_posmi,
                           const int32_t* d_nbr, const int32_t* d_fdofs, const int32_t* d_nloc, double* d_sd64,
                           double* d_so64, float* d_sd, float* d_so, int32_t* d_status, void* stream);
/* Inverses of the 3 x 3 flux blocks (FP32 out) from the CSR positions d_fb_pos (n_fb, 3, 3)
 * of each block's instance values; bit 2 of *d_status when a block is not SPD. */
int nfx_batch_tg_fluxinv(int n_fb, const int32_t* d_fb_inst, const double* const* d_vals, const int32_t* d_fb_pos,
                         float* d_fb_inv, int32_t* d_status, void* stream);
/* Block-Jacobi smoother wd = omega S_cc^-1 (FP64 inverse); bit 0 of *d_status when a block
 * is not SPD (its smoother block is then zero). */
int nfx_batch_tg_smoother(int nb, int n_cells, const double* d_sd64, double omega, float* d_wd, int32_t* d_status,
                          void* stream);
/* Per-instance coarse operators P^T S P (FP64, packed as ainv; d_ac must be zero). */
int nfx_batch_tg_coarse(int nb, int n_agg, const int32_t* d_agg_ptr, const int32_t* d_agg_mem,
                        const int32_t* d_agg_of, const int32_t* d_nbr, const double* d_sd64, const double* d_so64,
                        const int64_t* d_arow, const int32_t* d_a0, double* d_ac, void* stream);
/* Row sums rs of the S blocks (FP64 in, FP32 out). */
int nfx_batch_tg_rowsums(int nb, int n_cells, const double* d_sd64, const double* d_so64, float* d_rs,
                         void* stream);
/* One application z = TG(r) on the rows of one group, or of two groups of different nb in
 * shared launches (d_mode: NULL, or only instances whose mode is 0 get z written). */
int nfx_batch_tg_apply(int n_groups, const nfx_batch_tg* pc, const int32_t* d_mode, const float* d_r, float* d_z,
                       void* stream);
/* Attach (n_groups 1 or 2, of different nb) or detach (0) two-grid groups to a batched MINRES: each application
 * of its preconditioner then runs the diagonal one and, on the groups' rows, the cycle. The
 * descriptors are copied; the arrays they point to must outlive the solves. */
int nfx_minres_batch_set_pc(nfx_minres_batch* s, int n_groups, const nfx_batch_tg* h_groups);

/* ---- sort-based numbering (cold path) ---------------------------------------
 * The key step of the reference's mesh builder and of the PETSc pattern: entity ids in
 * first-appearance order (src/mesh/mesh.py:239-339: edges, faces and volumes numbered as
 * the cells are walked; src/mesh/mesh_topology.py:28-35: tag-filtered entity order) and
 * the union of element (row, col) entries (the pattern setValue builds,
 * src/assembly/ScatterFormData.py:5-11). Keys are non-negative int64 below 2^key_bits;
 * K < 2^31. d_work: nfx_sort_workspace(K) bytes. */
int nfx_sort_workspace(int64_t K, size_t* h_bytes);
/* Stable ascending sort: d_sorted_keys, and d_perm[i] = input position of the i-th key
 * (equal keys keep their input order). */
int nfx_stable_sort64(int64_t K, int key_bits, const int64_t* d_keys, int64_t* d_sorted_keys, int32_t* d_perm,
                      void* d_work, size_t work_bytes, void* stream);
/* d_ids[i] = rank of keys[i] among the distinct keys ordered by first appearance,
 * d_first[r] = input position of the first key of rank r (capacity K), *d_count = the
 * number of distinct keys (device int). */
int nfx_first_appearance(int64_t K, int key_bits, const int64_t* d_keys, int32_t* d_ids, int32_t* d_first,
                         int32_t* d_count, void* d_work, size_t work_bytes, void* stream);

/* ---- L2 projection ---------------------------------------------------------
 * Element mass matrix and right-hand side of one field of an L2 projection, as
 * L2ProjectorWeakForm.evaluate_form (src/weak_forms/l2_projector_weak_form.py:8-93)
 * scattered by l2_projector (src/postprocess/projectors.py:8-91). kind 0: vector
 * field (RT / BDM rows, contravariant Piola, the first dim components as the
 * reference), 1: scalar (Lagrange). The field occupies element-local DoFs off +
 * a * n_comp + c of nl; d_slot (nl*nl, E) and d_rhs_slot (nl, E) are slot codes
 * (CSR / stage index, bit 30 = orientation sign; code < 0 skips); multi-contributor
 * slots go to d_stage / d_rhs_stage, summed afterwards by nfx_stage_reduce. d_f: the
 * function at the element quadrature points, (E, nq, n_comp, dim) for kind 0 or
 * (E, nq, n_comp). d_tables64: w[nq] | phi[nq][nb][dim] (kind 0) or w[nq] | psi[nq][nb].
 * Matrix contributions within 1e-8 of 1e-16 are dropped, as the reference's scatter
 * filter (numpy.isclose) drops them. FP64 arithmetic; values stored as double when
 * fp64, else float. The right-hand side is b = -rg of the reference's driver. */
int nfx_assemble_projector(int kind, int dim, int nb, int nq, int n_comp, int nl, int off, int n_elem,
                           const int32_t* d_cells, const double* d_coords64,
                           const double* d_tables64, const double* d_f, const int32_t* d_slot,
                           const int32_t* d_rhs_slot, int fp64, void* d_vals, void* d_stage,
                           void* d_rhs, void* d_rhs_stage, void* stream);

/* ---- error norms (FP64) ---------------------------------------------------
 * Squared error of one field summed over elements, as l2_error / div_error
 * (src/postprocess/l2_error_post_processor.py:7-63,185-237). kind 0: vector
 * (RT/BDM rows) L2, 1: scalar (Lagrange) L2, 2: vector div. d_rows (E, nb, n_comp)
 * row codes (bit 30 = orientation sign); d_exact the exact field at the element
 * quadrature points: (E, nq, n_comp, dim) for kind 0, (E, nq, n_comp) otherwise.
 * d_tables64: w | phi | div (vector) or w | psi (scalar). Synchronous. */
int nfx_field_error(int kind, int dim, int nb, int nq, int n_comp, int n_elem,
                    const int32_t* d_cells, const double* d_coords64,
                    const double* d_tables64, const int32_t* d_rows,
                    const double* d_alpha, const double* d_exact, double* h_out,
                    void* stream);

/* Field values at n_points reference points of every element, FP64, as
 * node_average_quatity / cell_centered_quatity evaluate them for the VTK output
 * (src/postprocess/solution_post_processor.py:11-103): kind 0 (vector, RT / BDM rows)
 * d_out (E, n_points, n_comp, dim) = the first dim components of the Piola-mapped field;
 * kind 1 (scalar) d_out (E, n_points, n_comp). d_tables64: phi[n_points][nb][dim] or
 * psi[n_points][nb] (reference basis at the points); d_rows as nfx_field_error. */
int nfx_field_values(int kind, int dim, int nb, int n_points, int n_comp, int n_elem,
                     const int32_t* d_cells, const double* d_coords64, const double* d_tables64,
                     const int32_t* d_rows, const double* d_alpha, double* d_out, void* stream);

/* ---- learned correction of the Darcy block preconditioner (RT1-P0) --------
 * No reference counterpart: the reference factorises the Darcy system with
 * LU/MUMPS (examples/md_elliptic_scalar/md_elliptic_scalar.py:351-364). These
 * entry points build the pressure Schur approximation S = B diag(M)^-1 B^T of the
 * assembled matrix, run the message-passing network that maps S to a symmetric
 * sparse smoother W (pattern of S), train it, and apply the two-grid cycle that
 * replaces diag(S)^-1 in the block preconditioner (and, with nfx_flux_sweep_coef, two
 * damped Jacobi sweeps on the flux mass block that replace diag(M)^-1). Layouts:
 *   ELL: row p has 4 slots (the cells across its faces in ascending face order),
 *        d_nbr (E,4) int32 (-1 = boundary face), values (E,4);
 *   sample blocks: (ns, E) row-major (one sample per row);
 *   theta: nfx_gnn_num_params() parameters, FP32 master copy + bf16 copy (same
 *        layout: W0 (8,64) b0 (64) | per layer Wh (64,64) Wm (64,64) b (64) | ud (64)
 *        cd ue (64) ce);
 *   activations: h (5,E,64) bf16, m (4,E,64) bf16, z (5,E,64) FP32.
 * All launches are asynchronous on `stream`. */
int nfx_gnn_num_params(void);
/* S in ELL form (FP64) from the assembled FP64 values: pos_b / pos_qb / pos_m are
 * (E,4) value indices of B_pf, B_qf and M_ff (-1 = none). */
int nfx_schur_ell(int E, const int32_t* d_pos_b, const int32_t* d_pos_qb,
                  const int32_t* d_pos_m, const double* d_vals64, double* d_sdiag,
                  double* d_soff, void* stream);
/* Couplings a (E,4), node features x (E,8), FP32 copies of S (diag, off). */
int nfx_gnn_features(int E, const int32_t* d_nbr, const double* d_sdiag,
                     const double* d_soff, float* d_a, float* d_x, float* d_sd32,
                     float* d_so32, void* stream);
/* Forward pass: smoother entries Wd (E), Wo (E,4); activations kept for backward. */
int nfx_gnn_forward(int E, const int32_t* d_nbr, const float* d_a, const float* d_x,
                    const float* d_theta32, const uint16_t* d_theta16, uint16_t* d_h,
                    uint16_t* d_m, float* d_z, const double* d_sdiag, float* d_Wd,
                    float* d_Wo, void* stream);
/* Floats needed by the d_partial argument of nfx_gnn_backward. */
int nfx_gnn_backward_workspace(int E);
/* Backward pass from gamma_d = dL/dW_pp, gamma_o = dL/dW_pq + dL/dW_qp to the
 * parameter gradient (FP32, nfx_gnn_num_params() values). d_dh, d_dm: (E,64) FP32. */
int nfx_gnn_backward(int E, const int32_t* d_nbr, const float* d_a, const float* d_x,
                     const uint16_t* d_theta16, const uint16_t* d_h, const uint16_t* d_m,
                     const float* d_z, const double* d_sdiag, const float* d_gamma_d,
                     const float* d_gamma_o, float* d_dh, float* d_dm, float* d_partial,
                     float* d_grad, void* stream);
/* Y = beta Y0 + alpha (diag d + ELL off) X on a sample block (Y0 may be NULL). */
int nfx_ell_spmm(int E, int ns, const int32_t* d_nbr, const float* d_d, const float* d_off,
                 float alpha, const float* d_X, float beta, const float* d_Y0, float* d_Y,
                 void* stream);
/* Loss gradient w.r.t. the smoother entries:
 * gamma_d[p] = sum_s w_s Su_sp Se_sp, gamma_o[p,j] = sum_s w_s (Su_sp Se_sq + Su_sq Se_sp). */
int nfx_smoother_grad(int E, int ns, const int32_t* d_nbr, const float* d_Su,
                      const float* d_Se, const float* d_w, float* d_gamma_d,
                      float* d_gamma_o, void* stream);
/* C[s][a] = sum of X[s][p] over the members of aggregate a (CSR d_agg_ptr / d_agg_mem). */
int nfx_agg_sum(int n_agg, int ns, int E, const int32_t* d_agg_ptr, const int32_t* d_agg_mem,
                const float* d_X, float* d_C, void* stream);
/* Dense coarse operator P^T S P (FP64, row-major n_agg x n_agg); d_Ac must be zero. */
int nfx_coarse_dense(int n_agg, const int32_t* d_agg_ptr, const int32_t* d_agg_mem,
                     const int32_t* d_agg_of, const int32_t* d_nbr, const double* d_sdiag,
                     const double* d_soff, double* d_Ac, void* stream);
/* In-place inverse of an SPD block (n <= 128, FP32, row stride lda): the diagonal
 * blocks of the recursive 2 x 2 block inversion of the coarse operator. */
int nfx_spd_inverse_small(int n, float* d_A, int lda, void* stream);

/* Two-grid Schur block preconditioner: flux rows z_u = dMinv r_u; pressure rows
 * z_p = psign (x2 + W (r_p - S x2)), x2 = x1 + P Ainv P^T (r_p - S x1), x1 = W r_p,
 * with W = (Wd, Wo) and S = (sd, so) in ELL form (nbr), the aggregates as CSR
 * (agg_ptr, agg_mem) plus agg_of (cell -> aggregate), and the coarse inverse either
 * FP32 (Ainv) or row-scaled FP16 (Ainv16, Ascale; used when Ainv16 is not NULL).
 * Rows after the cell block, n_tail of them (the fracture pressures of the
 * mixed-dimensional system; n_tail = 0 for none), are scaled by tail_dinv, or, when
 * slot_tail is not NULL, eliminated by block LU with the cells: cell p's ELL slot j
 * couples to tail row slot_tail[4p+j] (-1 none) with coefficient slot_coef[4p+j] =
 * S_pf / S_ff, tail row f to cells tail_cells[2f+k] (-1 none) with tail_coef[2f+k] = S_fp,
 * and tail_dinv = 1 / S_ff; then r_p <- r_p - sum slot_coef r_f before the cycle (S in
 * ELL form then holds the cells' Schur complement S - S_pf S_ff^-1 S_fq) and
 * z_f = tail_dinv (r_f - sum tail_coef z_p) after it.
 * Flux rows: when flux_coef is not NULL, z_u = C r_u with C a sparse n_u x n_u operator
 * in slot-major ELL form, flux_width slots per row: z_i = sum_s flux_coef[s n_u + i]
 * r[flux_col[s n_u + i]] (padding slots: coefficient 0, any column < n_u); for example the
 * two damped Jacobi sweeps on the flux mass block of nfx_flux_sweep_coef. NULL:
 * z_u = dMinv r_u.
 * work: 3 n_cells + 2 n_agg floats (4 n_cells + 2 n_agg with slot_tail). */
typedef struct nfx_twogrid_pc {
  int n_u, n_cells, n_agg;
  const int32_t* nbr;
  const float* Wd;
  const float* Wo;
  const float* sd;
  const float* so;
  const int32_t* agg_ptr;
  const int32_t* agg_mem;
  const int32_t* agg_of;
  const float* Ainv;
  const uint16_t* Ainv16;
  const float* Ascale;
  const float* dMinv;
  float psign;
  float* work;
  int n_tail;
  const float* tail_dinv;
  const int32_t* slot_tail;
  const float* slot_coef;
  const int32_t* tail_cells;
  const float* tail_coef;
  int flux_width;
  const int32_t* flux_col;
  const float* flux_coef;
} nfx_twogrid_pc;
int nfx_twogrid_apply(const nfx_twogrid_pc* pc, const float* d_r, float* d_z, void* stream);
/* Flux-row operator of nfx_twogrid_pc for two damped Jacobi sweeps (weight omega) on the
 * flux mass block M from a zero start: C = 2 omega D^-1 - omega^2 D^-1 M D^-1, D = diag(M)
 * (positive), from the FP64 CSR values. d_pos / d_col (width x n_u, slot-major): CSR
 * position and column of the s-th entry of flux row i with column < n_u (d_pos -1 for a
 * padding slot); d_diag_pos[i]: CSR position of M_ii. Writes d_coef (width x n_u).
 * d_bound (one float, may be NULL when omega > 0): receives b = max_i sum_j |M_ij| / M_ii,
 * an upper bound of the spectral radius of D^-1 M; omega <= 0 selects omega = 2 / (0.5 + b)
 * (0.8 on meshes with b = 2, and omega b < 2 always, so C stays positive definite). */
int nfx_flux_sweep_coef(int n_u, int width, const int32_t* d_pos, const int32_t* d_col,
                        const int32_t* d_diag_pos, const double* d_vals64, double omega, float* d_coef,
                        float* d_bound, void* stream);
/* C = beta C + alpha op(A) op(B) (FP32, row-major, leading dimensions lda/ldb/ldc;
 * op(A) = A (M x K) or A^T when transA (A stored K x M); op(B) = B (K x N) or B^T when
 * transB); when d_C2 is not null the result is also stored transposed into it
 * (N x M, leading dimension ldc2). The products of the coarse-operator inversion
 * (K <= 256, MFMA, 64 x 64 or 32 x 32 output tiles). beta = 0 does not read C. */
int nfx_sgemm(int M, int N, int K, float alpha, const float* d_A, int lda, int transA,
              const float* d_B, int ldb, int transB, float beta, float* d_C, int ldc,
              float* d_C2, int ldc2, void* stream);
/* Row-scaled FP16 copy of a dense n x n FP32 matrix (Ascale[r] = max |A[r][:]|). */
int nfx_half_rows(int n, const float* d_A, uint16_t* d_A16, float* d_scale, void* stream);

/* ---- Mixed-dimensional transient Darcy -----------------------------------
 * Interface coupling of CouplingWeakForm.evaluate_form
 * (examples/md_elliptic_scalar/CouplingWeakForm.py:11-151), RT1 matrix facet DoFs and P0
 * fracture pressure, for m fracture triangles: triangle i lies on a matrix face with
 * vertex ids d_facet_nodes[3i..3i+2]; side s = 0 (positive) / 1 (negative) has element
 * d_side_elem[2i+s], combo index d_side_combo[2i+s] into d_trace (the facet trace
 * tables of nfx_assemble_darcy_bc) and row code d_side_code[2i+s] (its facet DoF, bit
 * 30 = negated). With s_q the normal trace out of that element and
 * 1/alpha = d_inv_alpha[i] = delta / (2 kappa_n):
 *   vals[pos[6i+3s]] += (1/alpha) sum_q w |det J_F| s_q^2,
 *   vals[pos[6i+3s+1]] and vals[pos[6i+3s+2]] += sum_q w |det J_F| s_q (signed),
 * the (u_s, u_s), (u_s, p) and (p, u_s) entries; all positions distinct. exchange_sign
 * multiplies the (p, u_s) entries: +1 the reference form (its fracture mass row adds the
 * matrix outflow u_s . n_s with the sign of the fracture divergence), -1 the conservative
 * exchange (the outflow is a source of the fracture). */
int nfx_md_coupling(int m, const int32_t* d_side_elem, const int32_t* d_side_combo,
                    const int32_t* d_side_code, const int32_t* d_facet_nodes, const int32_t* d_cells,
                    const void* d_coords, const void* d_trace, const void* d_fw,
                    const void* d_inv_alpha, const int32_t* d_pos, int exchange_sign, int fp64,
                    void* d_vals, void* stream);
/* dst[map[i]] = src[i] (float or double by fp64): places a block system's CSR values
 * or residual into the coupled system's (the reference scatters every block into one
 * PETSc matrix, md_elliptic_scalar.py:250-343). */
int nfx_md_scatter(int n, const void* d_src, const int32_t* d_map, int fp64, void* d_dst,
                   void* stream);
/* vals[pos[i]] = v[i] (mode 0) or vals[pos[i]] += v[i] (mode 1), distinct positions:
 * storage terms on the pressure diagonal, elimination of the fracture-tip flux DoFs. */
int nfx_md_update(int n, const int32_t* d_pos, const double* d_v, int mode, int fp64,
                  void* d_vals, void* stream);
/* Backward-Euler right-hand side (SundusWeakForm.py:148-150, mixed_transport_model.py:
 * 254-350): b[i] = b0[i] for i < n_u, b0[i] + mass[i - n_u] x[i] for the pressure rows. */
int nfx_md_rhs(int n, int n_u, const double* d_b0, const double* d_mass, const double* d_x,
               double* d_b, void* stream);
/* Diagonal block preconditioner of the coupled system (no reference counterpart: the
 * reference factorises with MUMPS): flux rows 1/|a_kk|, pressure rows
 * 1/(|a_kk| + sum_j |a_kj a_jk| / |a_jj|) over flux columns j. */
int nfx_md_block_diag(int n, int n_u, const int32_t* d_rowptr, const int32_t* d_col,
                      const double* d_vals, float* d_pinv, void* stream);

/* ---- FGMRES + FP64 iterative refinement -----------------------------------
 * Replaces KSP preonly + LU/MUMPS of the Darcy driver
 * (examples/md_elliptic_scalar/md_elliptic_scalar.py:351-364) when the learned
 * (flexible) preconditioner is used: FGMRES(restart) in FP32 with right
 * preconditioning by `pc` (two-grid Schur block preconditioner), else by the
 * diagonal d_pinv (NULL = none), inside FP64 refinement on the as-assembled FP64
 * matrix until ||b - A x||_2 / ||b||_2 <= rtol. reorth = 1 repeats classical
 * Gram-Schmidt once; the host reads the residual estimate every check_every
 * iterations. Synchronous. h_info[0] = refinement steps, h_info[1] = total FGMRES
 * iterations. Returns NFX_E_NOT_CONVERGED if rtol is not reached. */
typedef struct nfx_fgmres nfx_fgmres;
int nfx_fgmres_create(nfx_fgmres** out, int n_rows, int restart, void* stream);
int nfx_fgmres_destroy(nfx_fgmres* s);
int nfx_solve_fgmres_ir(nfx_fgmres* s, const int32_t* d_rowptr, const int32_t* d_col,
                        const float* d_val32, const double* d_val64, const nfx_twogrid_pc* pc,
                        const float* d_pinv, const double* d_b, double* d_x, double rtol,
                        double inner_rtol, int max_inner, int max_outer, int reorth,
                        int check_every, int* h_info, double* h_resid, double* h_hist,
                        int hist_cap);

#ifdef __cplusplus
}
#endif

#endif /* NFX_H */
