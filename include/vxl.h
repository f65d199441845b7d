/* libvxl: MI355X (gfx950) periodic HEX8 effective-stiffness solve path.
 *
 * C ABI.  Plain pointers and sizes only; device pointers are HIP device
 * addresses owned by the caller unless stated otherwise; `stream` arguments
 * are hipStream_t values passed as void*.  Every function returns VXL_OK (0)
 * or a negative VXL_ERR_* code; vxl_last_error() then describes the failure
 * (thread-local).  A handle is bound to the stream it was created with and
 * is not re-entrant: concurrent solves use separate handles (the reference's
 * process-level configuration, src/voxel_lmfem/solver.py:13-15).
 *
 * Layouts.  Grid (nz, ny, nx), node id (z*ny + y)*nx + x, periodic.
 * Displacement-like fields are component planes: field[r][c][node] with r the
 * right-hand side (load case) and c in {x, y, z}.  Material coefficients are
 * the 21 upper-triangle entries (i <= j, row major) of the engineering-Voigt
 * 6x6 stiffness in order 11,22,33,23,13,12 (the reference's PAIRS order,
 * src/voxel_lmfem/_core/materials/material_field.py:14).  Macro strains are
 * engineering strains in the same order.
 */
#ifndef VXL_H
#define VXL_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  VXL_OK = 0,
  VXL_ERR_ARG = -1,      /* invalid argument: the reference raises ValueError */
  VXL_ERR_HIP = -2,      /* HIP runtime failure: the reference raises RuntimeError */
  VXL_ERR_CONVERGENCE = -3 /* true residual above tolerance (solver.py:78-79) */
};

enum { VXL_INDEX_U8 = 1, VXL_INDEX_U16 = 2, VXL_INDEX_I32 = 4 };
enum { VXL_F32 = 0, VXL_F64 = 1 };
/* Preconditioner precision only: FP32 arithmetic with the spectrum stored as
 * BF16 pairs between passes (power-of-two axes >= 8). */
enum { VXL_F32_BF16 = 2 };
/* Solve modes.  FP64: FP64 operator, FP64 FFT preconditioner, FP64 CG.
 * MIXED: FP32 operator, FP32 FFT preconditioner and FP32 CG with FP64
 * reliable updates (true residual f - K u, dots and stress in FP64). */
enum { VXL_SOLVE_FP64 = 0, VXL_SOLVE_MIXED = 1 };
/* Preconditioners of the solver: the FFT Green operator (default) or the
 * scalar Jacobi control. */
enum { VXL_PRECOND_FFT = 0, VXL_PRECOND_JACOBI = 1, VXL_PRECOND_FFT_BF16 = 2 };
enum { VXL_ELEMENT_FI = 0, VXL_ELEMENT_C3D8R_HG = 1 };
/* Krylov method of a solver: one CG recurrence per load case sharing every
 * launch (default), or one block CG over all load cases. */
enum { VXL_METHOD_CG = 0, VXL_METHOD_BLOCK_CG = 1 };

typedef struct vxl_cell vxl_cell;
typedef struct vxl_precond vxl_precond;
typedef struct vxl_solver vxl_solver;

/* Thread-local text of the last failure ("" if none). */
const char* vxl_last_error(void);
/* Library ABI version (major*10000 + minor*100 + patch). */
int vxl_version(void);

/* Device allocation cache.  Handles allocate their workspaces through a
 * per-(device, size) cache: a destroyed handle's blocks are kept and reused by
 * the next handle that asks for the same sizes, so repeated set-up of one grid
 * does not return to the driver (the reference gets the same from the PyTorch
 * caching allocator; torch.cuda.empty_cache is the counterpart of
 * vxl_release_cached_memory).  Running out of device memory releases the
 * cache of that device and retries once. */
int vxl_cached_bytes(uint64_t* bytes);
int vxl_release_cached_memory(void);

/* Material front end on the device: index_dev[e] = lut_host[labels[e] - label_lo]
 * over a resident label array (label_bytes 4 = int32, 8 = int64; C order
 * z, y, x), where lut_host (span entries) holds each label's coefficient row
 * or -1 for labels without a material, or -2 for a local-frame phase whose
 * voxel e gets its own row local_base + e (see vxl_local_rows; int32 index).
 * Returns present_host[span] (1 where
 * the label occurs), *undefined_host (1 if a voxel's label has no row; its
 * index is then 0) and samples_host, the rows at the Zeman sampling lattice
 * (every max(1, n/16)-th voxel per axis, C order,
 * ceil(nz/sz) * ceil(ny/sy) * ceil(nx/sx) entries).  Synchronises `stream`.
 * Replaces the label half of build_hybrid_material (hybrid_material.py:38-48,
 * rows in phase_table order, material_field.py:57-68) and the sampling of
 * _compute_reference_medium (hybrid_adapter.py:101-125). */
int vxl_label_rows(int nz, int ny, int nx, int label_bytes, const void* labels_dev, long long label_lo, int span,
                   const int* lut_host, int index_kind, void* index_dev, long long local_base, int* present_host,
                   int* undefined_host, int* samples_host, void* stream);

/* Local-frame phases on the device (the per-voxel half of
 * build_hybrid_material, hybrid_material.py:49-61, with the frame rules of
 * _orientation_values, material_field.py:71-102, and the Bond rotation of
 * _rotate_packed, material_field.py:105-114).  A lookup entry of -2 in
 * vxl_label_rows gives voxel e the row local_base + e (int32 index); this call
 * fills those rows: for every voxel whose label has slot_host[label - lo] =
 * p >= 0, rows_dev[(local_base + e) * 21 ...] = Bond(R_e) C_p Bond(R_e)^T with
 * C_p = c21_host[p * 21 ...].  mode 1: R_e from the axis pair a_dev, b_dev
 * (columns a, c x a, c with c = a x b normalised); mode 2: a_dev holds R_e
 * row major.  stride is 0 (one frame for every voxel) or 3 (mode 1) / 9
 * (mode 2) doubles per voxel.  bad_host[5] flags, as the host checks raise:
 * 0 nonfinite, 1 zero axis, 2 collinear axes, 3 not orthonormal (1e-10),
 * 4 det != +1 (1e-10).  Only the voxel count nz ny nx matters here, so a
 * list of selected voxels may be passed as a 1 x 1 x m "grid" with their
 * labels and frames.  Synchronises `stream`. */
int vxl_local_rows(int nz, int ny, int nx, int label_bytes, const void* labels_dev, long long label_lo, int span,
                   const int* slot_host, int phases, const double* c21_host, int mode, const double* a_dev,
                   const double* b_dev, long long stride, double* rows_dev, long long local_base, int* bad_host,
                   void* stream);

/* Cell: grid, voxel size and material field.
 * Replaces the operator construction of the reference's hybrid FI storage:
 *   HybridFIOperator.__init__  src/voxel_lmfem/_core/materials/hybrid_adapter.py:45-68
 *   build_hybrid_material      src/voxel_lmfem/_core/materials/hybrid_material.py:14-70
 * index_dev: per-voxel row index (uint8/uint16/int32 per index_kind), device,
 *   caller-owned, must outlive the cell.
 * c21_host: rows x 21 FP64 coefficients (host); copied to device in FP64 and
 *   FP32.
 * scale_dev: optional per-voxel FP32 factor multiplying the row (NULL = 1),
 *   device, caller-owned.  The reference has no continuous-modulus field; this
 *   carries density (SIMP) cells.
 * h = voxel size (hx, hy, hz) = domain / (nx, ny, nz)
 *   (hybrid_adapter.py:147). */
int vxl_cell_create(vxl_cell** out, int nz, int ny, int nx, double hx, double hy, double hz,
                    int index_kind, const void* index_dev, int rows, const double* c21_host,
                    const float* scale_dev, void* stream);
/* vxl_cell_create with the rows x 21 FP64 coefficient table already on the
 * device (c21_dev, caller-owned, read once). */
int vxl_cell_create_device_rows(vxl_cell** out, int nz, int ny, int nx, double hx, double hy, double hz,
                                int index_kind, const void* index_dev, int rows, const double* c21_dev,
                                const float* scale_dev, void* stream);
int vxl_cell_destroy(vxl_cell* cell);

/* out = K u for nrhs fields (precision VXL_F32/VXL_F64), component-plane
 * layout.  Replaces the pybind `apply`:
 *   operator_kernels.cu:88-116 (split background + contrast FP64 K.u) and
 *   hybrid_kernels.cu:117-127 (pool-indexed FP64 K.u). */
int vxl_apply(vxl_cell* cell, int precision, int nrhs, const void* u_dev, void* out_dev);

/* Stencil schedule of the cell per precision: 1 = per-layer 8-node modal
 * transform, 2 = persistent face columns, 3 = 2 with the plane after next
 * prefetched into registers (default 3 for both precisions; the
 * CG form of the stencil used inside vxl_solver_run has schedules 2 and 3
 * only and runs 2 when 1 is set).  All schedules compute the same sums.
 * No reference counterpart: the reference's variant choice is recorded in
 * _core/hybrid_operator/recommended.json:1-6 by hybrid_operator/study.py:64-82;
 * vxl_tune_stencil is this build's measured equivalent: it times every
 * schedule on the given buffers (reps launches each, HIP events on the cell's
 * stream), keeps the fastest, and returns it in *best and the per-launch
 * milliseconds of schedules 1..3 in ms3 (may be NULL). */
int vxl_set_stencil_variant(vxl_cell* cell, int precision, int variant);
int vxl_get_stencil_variant(vxl_cell* cell, int precision);
int vxl_tune_stencil(vxl_cell* cell, int precision, int nrhs, const void* u_dev, void* out_dev, int reps,
                     int* best, double* ms3);
/* Full stencil configuration per precision: tile height (workgroup tile of
 * 32 x tile_y output nodes; VXL_F32: 14 (default), 8 or 6; VXL_F64: 6
 * (default), 8 or 4), schedule (1-3, as above) and the fewest z planes per
 * workgroup chunk (2..64, default 8; each chunk recomputes one element
 * layer, and shallower chunks give small cells more workgroups).  Every
 * configuration computes the same element sums.  No reference counterpart
 * (see vxl_tune_stencil). */
int vxl_set_stencil_config(vxl_cell* cell, int precision, int tile_y, int variant, int zmin);
/* Element of the cell's operator: VXL_ELEMENT_FI (default; fully integrated
 * HEX8, 2x2x2 Gauss) or VXL_ELEMENT_C3D8R_HG, the reference's reduced-
 * integration control (femlib/c3d8r.py:62-78: centroid rule V B0^T C B0 plus
 * the hourglass penalty s(C) H H^T, s = coef * max((I1 + 7 I2)/60, 1e-12) *
 * V / hbar^2; femlib/batched/c3d8r_hourglass_operator.py).  RHS, mean stress
 * and centre fields are the same for both elements (both use the centroid
 * strain of a constant macro strain).  The HG element runs the default tile
 * and schedule 3; preconditioners built afterwards use its symbol (FFT) or
 * diagonal (Jacobi).  Set it before creating solvers or preconditioners.
 * hourglass_coefficient must be finite and positive (the reference default
 * is 0.1); it is ignored for VXL_ELEMENT_FI.  VXL_ERR_ARG for the HG element
 * while the matrix-core stencil or per-voxel frames are on (FI only). */
int vxl_cell_set_element(vxl_cell* cell, int element, double hourglass_coefficient);
int vxl_get_stencil_config(vxl_cell* cell, int precision, int* tile_y, int* variant, int* zmin);
/* Matrix-core FP32 stencil (enable 1) or the modal VALU stencil (enable 0,
 * the default).  The matrix-core form is the label-grouped dense element
 * product f_e = K_L (u_e - u_e0) of the reference's grouped kernel
 * (_core/runtime/fi/grouped_kernels.cu:12-74): the dense 24x24 FI element
 * matrix of every coefficient row is built once here (host FP64, 2x2x2
 * Gauss rule), scaled by a power of two and split into FP16 hi/lo operands;
 * each element's displacement differences are split the same way, and the
 * three leading products run on v_mfma_f32_16x16x32_f16 with FP32
 * accumulation (about 2^-21 relative per element column), one product pass
 * per distinct row among the 64 elements of a wave.  It replaces every FP32
 * apply of the cell, plain and CG form (vxl_apply, vxl_solver_run in mixed
 * mode); FP64 is unchanged.  FI element and at most 4096 coefficient rows
 * (label-only cells); VXL_ERR_ARG otherwise.  Tile 32 x 7 nodes, schedule
 * and tile settings of the modal stencil do not apply to it. */
int vxl_cell_set_mfma(vxl_cell* cell, int enable);
/* Per-voxel material frames for the FP32 stencil: element stiffness rebuilt
 * in the stencil from each voxel's phase and orientation, in place of its
 * coefficient row.  Replaces the per-voxel rows that the reference's hybrid
 * pool gives every local-frame voxel (hybrid_material.py:49-61, rotated by
 * _rotate_packed, material_field.py:105-114) in every FP32 apply of the cell
 * (vxl_apply, the mixed solve's inner iterations); FP64 applies, the RHS, the
 * certificate and the stress keep the cell's FP64 rows.
 * phase_rows: nphase x 21 FP64 (host), each phase's C21 in its own frame;
 * phase_local[p] != 0 marks local-frame phases (rotated per voxel; a phase
 * that is exactly cubic in its frame takes the cubic closed form, one that
 * is orthotropic in its frame (exact zeros outside the normal block and the
 * shear diagonal) the orthotropic closed form, any other the Bond
 * rotation), 0 global-frame phases (used as given).  phase_index_dev: uint8
 * phase slot of every voxel (device); quat_dev: 4 floats (w, x, y, z) per
 * voxel (device), the unit quaternion of the voxel's rotation R (columns =
 * local axes in cell coordinates), read for local-frame voxels.  Both are
 * copied (17 B per voxel).  phase_index_dev == NULL turns the frames off.
 * FI element, no scale field, not with the matrix-core stencil; 1 to 256
 * phases; VXL_ERR_ARG if an index is >= nphase or a quaternion is not finite
 * with |q|^2 within 1e-5 of 1.  No counterpart in the reference, which stores
 * one rotated row per local-frame voxel. */
int vxl_cell_set_frames(vxl_cell* cell, int nphase, const double* phase_rows, const int* phase_local,
                        const void* phase_index_dev, const float* quat_dev);
/* Stacked cells: declare that the cell's index array holds `groups` index
 * blocks of n entries each (cells of this grid and voxel size, one after the
 * other, indexing the one row table passed at creation), so one launch runs
 * all of them: right-hand side r of an nrhs-RHS call belongs to cell
 * r / (nrhs / groups), and nrhs must be a multiple of groups.  The
 * reference runs one cell per call (solver.py:9-85) and batches only
 * same-shape cells of one material (_core/runtime/repro/source/femlib/
 * batched/driver.py:13-60); here the cells may differ in material and each
 * keeps its own reference medium (vxl_solver_set_media).  Stencil, RHS,
 * mean stress, FFT preconditioner and solver (per-case CG, mixed or FP64)
 * take stacked cells; centre fields, the energy gradient, the Jacobi
 * diagonal, block CG, per-voxel frames, the matrix-core form, a scale field
 * and the HG element do not (VXL_ERR_ARG).  1 to 64 cells; set before
 * creating solvers. */
int vxl_cell_set_groups(vxl_cell* cell, int groups);
/* Autotuner: times all 27 (tile, schedule, chunk depth in {8, 4, 2})
 * configurations on the given buffers (reps launches each, HIP events on the
 * cell's stream), keeps the fastest on the cell and returns it.  ms_table
 * (27 entries, may be NULL) receives the per-launch milliseconds at
 * [(tile_index * 3 + schedule - 1) * 3 + depth_index], tiles in the order
 * listed above.  The persistent cache of tuned configurations (JSON keyed by
 * grid, precision, load count and material index width) lives in the Python
 * layer (voxel_lmfem.autotune). */
int vxl_autotune_stencil(vxl_cell* cell, int precision, int nrhs, const void* u_dev, void* out_dev, int reps,
                         int* tile_y, int* variant, int* zmin, double* ms_table);

/* f = -V sum_e B0^T C_e E for nload macro strains (host, nload x 6), FP64.
 * Replaces `rhs`  hybrid_kernels.cu:128-137 (rhs_kernel :49-68). */
int vxl_rhs(vxl_cell* cell, int nload, const double* macro_host, double* f_dev);

/* out_host (nload x 6) = <C (E + B0 u)> over all elements, FP64.
 * Replaces `stress`  hybrid_kernels.cu:138-147 (stress_kernel :69-106). */
int vxl_avg_stress(vxl_cell* cell, int nload, const double* u_dev, const double* macro_host,
                   double* out_host);

/* Element-centre fields of nload FP64 solutions (macro_host: nload x 6):
 * strain = E + B0 u_e (engineering Voigt), stress = C_e strain, both written
 * as planes [r][i][e] (nload x 6 x N), and energy = strain . stress / 2 as
 * [r][e] (nload x N); any output may be NULL.  Replaces recover_center_fields
 * (femlib/batched/c3d8r_hourglass_operator.py:752-758, called by
 * benchmark.workload for its field statistics, benchmark.py:99-107). */
int vxl_center_fields(vxl_cell* cell, int nload, const double* u_dev, const double* macro_host,
                      double* strain_dev, double* stress_dev, double* energy_dev);

/* Adjoint of C_eff.  For nload converged FP64 fluctuation fields u_dev
 * (nload x 3 x N planes, e.g. the u_dev_out of vxl_solver_run) of macro
 * strains E (macro_host, nload x 6) and a symmetric weight matrix W (w_host,
 * nload x nload, row major):
 *   out_dev[e] = sum_ij W_ij int_e (E_i + B u_i)^T D_e (E_j + B u_j) dV
 * with the operator's 2x2x2 Gauss rule, where D_e is the cell's FP64
 * coefficient row of e (without the per-voxel scale) or, if d21_dev is not
 * NULL, the per-element tensor d21_dev[e] (N x 21, PAIRS order).  Because the
 * problem is self-adjoint (no second solve), with C_ij = E_i . CH E_j:
 *   d(sum_ij W_ij C_ij)/d s_e = out_dev[e] / |Omega|   for C_e = s_e C_row(e),
 *   d(sum_ij W_ij C_ij)/d t_e = out_dev[e] / |Omega|   for d21 = dC_e/dt_e.
 * No counterpart in the reference, which has no gradient; checked against
 * central finite differences of the C_eff solve. */
int vxl_energy_gradient(vxl_cell* cell, int nload, const double* u_dev, const double* macro_host,
                        const double* w_host, const double* d21_dev, double* out_dev);

/* Subtract the per-component mean of nrhs FP64 fields in place.
 * Replaces HybridFIOperator.project_zero_mean  hybrid_adapter.py:78-80. */
int vxl_project_zero_mean(vxl_cell* cell, int nrhs, double* u_dev);
/* vxl_project_zero_mean without any allocation, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  The
 * partial sums and means use a static per-GPU device array of 768 KiB (held
 * until the process ends), so nrhs is at most 16 (VXL_ERR_ARG beyond).  The
 * work is enqueued on the cell's stream and the call returns without waiting
 * for it.  Every call on one GPU shares that array: calls that overlap on one
 * GPU (other streams, other threads) race and must be ordered by the caller.
 * Same kernels and results as vxl_project_zero_mean, bit for bit. */
int vxl_project_zero_mean_noalloc(vxl_cell* cell, int nrhs, double* u_dev);

/* FFT Green preconditioner of the isotropic reference medium (lam, mu):
 * z = IFFT(K0(k)^-1 FFT(r)), k = 0 mode removed, hand-written LDS pencil FFTs.
 * Replaces enable_fi_reference/fi_coeff (formula_adapter.py:40-68,
 * formula_kernels.cu:104-137, 199-208), contest prepare/spectral
 * (contest_kernels.cu:46-94) and NativeFFT (native_fft.cu:11-58).
 * precision: VXL_F32 or VXL_F64 (arithmetic and storage of the spectra), or
 * VXL_F32_BF16. */
int vxl_precond_create(vxl_precond** out, vxl_cell* cell, int precision, double lam, double mu,
                       int nrhs_max);
/* r_dev, z_dev: nrhs fields in the precond's precision. */
int vxl_precond_apply(vxl_precond* pc, int nrhs, const void* r_dev, void* z_dev);
int vxl_precond_destroy(vxl_precond* pc);

/* Six-load (or nrhs-load) C_eff solver: RHS, block PCG, certification,
 * stress recovery.  Replaces benchmark.workload/solve
 * (src/voxel_lmfem/_core/runtime/benchmark.py:77-112) driving the rewritten
 * solve_pcg_batched (femlib/batched/pcg.py:20-252, chain_adapter.py:80-142).
 * Workspaces are allocated here, not in vxl_solver_run. */
int vxl_solver_create(vxl_solver** out, vxl_cell* cell, int mode, int nrhs, double lam, double mu);
/* macro_host: nrhs x 6 engineering strains.
 * ch_host: 6 x nrhs, column j = <sigma> of load j (CH[:, j], benchmark.py:98).
 * iterations/restarts/residuals: nrhs each (host); residuals are the
 *   certified FP64 true relative residuals ||f - K u||/||f||.
 * u_dev_out: optional (NULL) device buffer nrhs x 3 x N FP64 receiving the
 *   zero-mean fluctuation fields.
 * Returns VXL_ERR_CONVERGENCE if a load does not certify within max_iter. */
int vxl_solver_run(vxl_solver* s, const double* macro_host, double tol, int max_iter,
                   double* ch_host, int* iterations, int* restarts, double* residuals,
                   double* u_dev_out);
/* out_dev (3 x N FP64 planes) = 1 / diag K, the operator's diagonal assembled
 * from the 8 elements around each node (2x2x2 Gauss rule); the reference
 * caches the same diagonal for its Jacobi control (op._diag_cache,
 * chapter_aligned/study.py:84-87). */
int vxl_cell_inverse_diagonal(vxl_cell* cell, double* out_dev);
/* Selects the solver's preconditioner.  VXL_PRECOND_FFT_BF16 (mixed mode):
 * the FFT Green preconditioner with its spectrum stored in BF16 between
 * passes (VXL_F32_BF16).  VXL_PRECOND_JACOBI is the
 * reference's control (repro/bench/chapter_aligned/study.py:84-90):
 * z = project_zero_mean(r / diag K) with the assembled diagonal of K; the
 * diagonal is assembled here on the device (once per solver). */
int vxl_solver_set_preconditioner(vxl_solver* s, int kind);
/* Krylov method of the solver's runs.  VXL_METHOD_CG (default): one
 * preconditioned CG recurrence per load case, all cases sharing every launch
 * (the reference's batched PCG, femlib/batched/pcg.py:20-252, rewritten).
 * VXL_METHOD_BLOCK_CG: O'Leary's block PCG over the load cases (at most 8),
 * one Krylov space for all of them: Q = K P, alpha = (P^T Q)^-1 P^T R,
 * X += P alpha, R -= Q alpha, Z = M R, beta = -(P^T Q)^-1 Q^T Z,
 * P = Z + P beta, with the R x R algebra on the host in FP64; directions
 * whose Cholesky pivot of P^T Q falls below 1e-12 of its diagonal are
 * dropped from the step (rank deficiency), certified systems leave the
 * block, and the FP64 certificate and reliable updates are those of the
 * default method.  FFT preconditioner only.  The reference's batched PCG
 * solves each case separately; the block method is this build's. */
int vxl_solver_set_method(vxl_solver* s, int method);
/* Stacked cells: the isotropic reference medium (lam[g], mu[g]) of each of
 * the cell's groups (host arrays of vxl_cell_set_groups' count), used by the
 * FFT preconditioner for that cell's right-hand sides; a reliable update
 * re-certifies the systems of the cell that triggered it. */
int vxl_solver_set_media(vxl_solver* s, const double* lam, const double* mu);
/* Number of four-iteration blocks the last vxl_solver_run replayed from its
 * captured HIP graph (graph replay is opt-in with VXL_GRAPH=1; 0 without it,
 * with the fused forward pass, or when capture is unavailable).  No reference counterpart (the reference launches
 * through PyTorch eagerly). */
int vxl_solver_graph_blocks(vxl_solver* s);
/* Runs count independent solvers (each bound to its own cell and stream)
 * concurrently from `workers` host threads, largest cells first, so small
 * cells overlap on the GPU instead of waiting on each other's host checks.
 * macro_host packs each solver's R x 6 strains one after the other; ch_host
 * packs each solver's 6 x R CH block, and iterations/restarts/residuals R
 * entries per solver, in the same order.  Returns the first failure (the
 * message names the cell).  The reference runs cells one process at a time
 * (solver.py:13-15); this is the batched-cell driver of the uneven-cell
 * workload. */
int vxl_solver_run_batch(vxl_solver* const* solvers, int count, const double* macro_host, double tol, int max_iter,
                         double* ch_host, int* iterations, int* restarts, double* residuals, int workers);
/* Per-phase device time of the last run in milliseconds:
 * [0] setup+rhs, [1] iterations, [2] certification+stress. */
int vxl_solver_timings(vxl_solver* s, double* ms3);
/* Residual history of system k from the last run, as in the reference's
 * per-call `trace` (benchmark.py:62-75): writes up to `capacity` (iteration,
 * relative residual) pairs, a negative iteration marking an FP64 true
 * residual; returns the number of entries (>= 0) or an error code. */
int vxl_solver_trace(vxl_solver* s, int k, int* iterations, double* residuals, int capacity);
int vxl_solver_destroy(vxl_solver* s);

#ifdef __cplusplus
}
#endif

#endif /* VXL_H */
