// CG vector-kernel launchers (cg.hip) and the FFT preconditioner (fft.hip).
#pragma once

#include "vxl_internal.h"

namespace vxl {

// Device scalar slots, scal[slot * nrhs + k].
enum ScalarSlot { kRho = 0, kPQ, kAlpha, kBeta, kRR, kFF, kTmp, kNumSlots };
// finalize_kernel operations.
enum FinalizeOp { kOpStore = 0, kOpRho, kOpFF, kOpRR, kOpAlpha, kOpBeta };

constexpr int kVecBlocks = 2048;  // max workgroups of a vector kernel (partials per RHS)

int vec_blocks(int64_t len);

template <typename A, typename B>
int launch_dot(const A* a, const B* b, int64_t len, int nrhs, double* partial, const int* active,
               hipStream_t s);
// kOpRR with `stop` latches the first iteration below tol: stop[k] = iteration,
// active[k] = 0 (on the device).
int launch_finalize(const double* partial, int nblk, int op, double* scal, int nrhs, int* active,
                    int* restart, hipStream_t s, int iteration = 0, double tol = 0.0,
                    int* stop = nullptr, double* hist = nullptr);
// Stored search directions p_t (nrhs fields each) whose steps alpha_t are
// still to be added to x64.
constexpr int kMaxDirections = 16;
template <typename T>
struct DirectionSet {
  const T* p[kMaxDirections];
  int count;
};
template <typename T>
int launch_update_r(T* r, const T* q, int64_t len, int nrhs, const double* alpha, double* partial,
                    const int* active, hipStream_t s);
// out[i * 6 + j] = scale * f_i . u_j for six stacked FP64 fields (partial:
// 36 * kGramBlocks doubles).
constexpr int kGramBlocks = 1024;
int launch_gram6(const double* f, const double* u, int64_t len, double* partial, double scale, double* out,
                 hipStream_t s);
// f -= its per-(RHS, component) mean, r = f (in r's precision) and the
// ||f||^2 partials, which start at partial + 3 nrhs (nblk_x + 1) with
// *nblk_out blocks per RHS.
template <typename TR>
int launch_center_copy(double* f, TR* r, int64_t n, int nrhs, double* partial, int* nblk_out, hipStream_t s);
// out[row] = scale * sum of partial[row * nblk + 0 .. nblk) (fixed order).
int launch_row_totals(const double* partial, int rows, int nblk, double scale, double* out, hipStream_t s);
// partial[c * blocks + block]: per-block sums of the 21 stiffness
// coefficients (times the per-voxel scale) over all elements.
int launch_stiffness_sums(const vxl_cell* cell, double* partial, int blocks, hipStream_t s);
// x64 += sum_t alpha[t * nrhs + k] p_t for every system k (assign: x64 =
// the sum, its old contents ignored); terms with alpha = 0 are skipped.
template <typename T>
int launch_flush(double* x64, const DirectionSet<T>& d, const double* alpha, int nrhs, int64_t len, bool assign,
                 hipStream_t s);
template <typename T>
int launch_update_xr(T* x, T* r, const T* p, const T* q, int64_t len, int nrhs, const double* alpha,
                     double* partial, const int* active, hipStream_t s);
int launch_residual(double* r, const double* f, const double* q, float* r32, int64_t len, int nrhs,
                    double* partial, const int* active, hipStream_t s);
int launch_accumulate(double* x64, float* x32, int64_t len, int nrhs, const int* active, hipStream_t s);
template <typename A, typename B>
int launch_convert(B* dst, const A* src, int64_t len, int nrhs, const int* active, hipStream_t s);
template <typename T>
int launch_fill_zero(T* dst, int64_t len, int nrhs, const int* active, hipStream_t s);
// partial: >= 3 * nrhs * (vec_blocks(n) + 1) doubles.
int launch_project_zero_mean(double* u, int64_t n, int nrhs, double* partial, const int* active,
                             hipStream_t s);

// Block CG over the load cases (block.hip), at most block_max_rhs() systems.
// out (R x R, device) = A^T B over len entries per system (partial: R * R *
// block_gram_blocks(len) doubles); X_j += sum_i P_i alpha_ij (X may be null)
// and R_j -= sum_i Q_i alpha_ij for active systems with ||R_j||^2 partials
// rr (R * block_gram_blocks(len)); P_j = Z_j + sum_i P_i beta_ij for keep[j],
// 0 otherwise.  alpha, beta: R x R row major (device).
int block_max_rhs();
int block_gram_blocks(int64_t len);
template <typename T>
int launch_block_gram(const T* A, const T* B, int64_t len, int R, double* partial, double* out, hipStream_t s);
template <typename T>
int launch_block_update(T* X, T* Rv, const T* P, const T* Q, const double* alpha, const int* active, int64_t len,
                        int R, double* rr_partial, hipStream_t s);
template <typename T>
int launch_block_direction(T* P, const T* Z, const double* beta, const int* keep, int64_t len, int R,
                           hipStream_t s);

// Scalar Jacobi control preconditioner (jacobi.hip).  dinv = 1 / diag K per
// DOF (component planes); one CG step forms z = dinv r - mean(dinv r), the
// r.z partials, beta (kOpBeta) and p = z + beta p.  partial: >= 3 * nrhs *
// jacobi_blocks(n) doubles; rz_partial: nrhs * jacobi_blocks(n); mean: 3 nrhs.
int jacobi_blocks(int64_t n);
int launch_jacobi_diag(const vxl_cell* cell, double* dinv64, float* dinv32, hipStream_t s);
template <typename T>
int jacobi_cg_step(const T* r, T* p, const T* dinv, int64_t n, int nrhs, double* partial, double* mean,
                   double* rz_partial, double* scal, int* active, int* restart, hipStream_t s,
                   const T* p_old = nullptr);

}  // namespace vxl

// FFT Green preconditioner handle (fft.hip).
struct vxl_precond {
  vxl_cell* cell;
  int precision;
  int nrhs_max;
  double lam, mu;
  void* impl;
  // Stacked cells: device (lam, mu) per cell, right-hand side r taking cell
  // r / (nrhs / groups); nullptr: lam, mu for every right-hand side.
  double* media;
  int groups;
};

namespace vxl {
int precond_create(vxl_precond* pc);
int precond_apply(vxl_precond* pc, int nrhs, const void* r, void* z, hipStream_t s);
// CG preconditioner step: z = M r with r . z reduced by Parseval in the z
// pass (rz_partial[rhs * precond_rz_blocks + b], FP64), mid() run between the
// z pass and the inverse passes (it forms beta from the partials), and the
// last pass writing the search direction p = z + beta[rhs] p in place
// (beta = 0: p = z; the old p is then not read).  Systems with active == 0
// are skipped.
template <typename T>
int precond_apply_cg(vxl_precond* pc, int nrhs, const T* r, T* p, double* rz_partial, const double* beta,
                     const std::function<int()>& mid, const int* active, hipStream_t s);
int precond_rz_blocks(const vxl_precond* pc);
// The two halves of precond_apply_cg, for a caller that acts between them.
// Forward x pass of r; with q != nullptr it first applies r -= alpha q (and
// x += alpha p when x != nullptr) to every entry it loads and writes FP64
// ||r||^2 partials (rr_partial[rhs * precond_fused_update_blocks + b]); that
// form needs the four-step x pass (precond_fused_update_blocks > 0).
// sr (spectral residual, precond_residual_spectrum first): the residual is
// kept as its x spectrum R^ inside the preconditioner.  Without q the pass
// forms R^ from r; with q it updates R^ -= alpha q^ and writes the ||r||^2
// partials (Parseval), r is not read and x, p must be null.
template <typename T>
int precond_cg_forward(vxl_precond* pc, int nrhs, T* r, T* x, const T* p, const T* q, const double* alpha,
                       double* rr_partial, const int* active, hipStream_t s, bool sr = false);
// y, z (+ Green, r.z), mid(), y^-1 and the x inverse pass writing
// p = z + beta p_old (p_old == nullptr: p itself, in place).  sr: the y pass
// reads R^.
template <typename T>
int precond_cg_rest(vxl_precond* pc, int nrhs, T* p, double* rz_partial, const double* beta,
                    const std::function<int()>& mid, const int* active, hipStream_t s, const T* p_old = nullptr,
                    bool sr = false);
// Whether the spectral-residual form is available (FP32 spectrum storage,
// four-step x and y passes; VXL_SPECTRAL_RESIDUAL=0 turns it off), and the
// allocation of its R^ buffer (nrhs_max spectra; once, outside any capture).
bool precond_spectral_residual_ok(const vxl_precond* pc);
int precond_residual_spectrum(vxl_precond* pc);
// Workgroups per RHS of the fused forward pass, 0 if it is not available.
int precond_fused_update_blocks(const vxl_precond* pc);
// Whether the CG r update runs inside the x pass (VXL_FUSED_UPDATE=0 turns
// it off); a solver also needs its partial_fx buffer and no graph replay.
bool fused_update_enabled();
// z = M r for the systems with active[k] != 0 only.
int precond_apply_masked(vxl_precond* pc, int nrhs, const void* r, void* z, const int* active,
                         hipStream_t s);
void precond_release(vxl_precond* pc);
}  // namespace vxl

// Solver handle (solver.hip).
struct vxl_solver {
  vxl_cell* cell;
  int mode;
  int R;
  vxl_precond pc;
  bool pc_ready;
  int64_t len;  // 3N
  double *f, *x64, *r64, *q64;
  void *x, *r, *p, *q;  // inner precision (x, r, q alias the FP64 buffers in FP64 mode)
  double* scal;
  double* partial;
  double* partial_st;  // stencil p.q partials
  double* partial_rz;  // FFT z-pass r.z partials (Parseval)
  double* partial_fx;  // ||r||^2 partials of the fused forward x pass (four-step x only)
  int precond_kind;    // VXL_PRECOND_FFT or VXL_PRECOND_JACOBI
  int element;         // element of the cell when the preconditioner was built
  double* dinv64;      // Jacobi: 1 / diag K (3N FP64)
  float* dinv32;       // Jacobi, mixed mode: FP32 copy
  double* jmean;       // Jacobi: per-component means of dinv r (3 R)
  double* jrz;         // Jacobi: r.z partials
  int nblk;
  int* active_dev;
  int* restart_dev;
  int* stop_dev;
  double* macro_dev;
  double* rhs_sig;  // per-row element stresses of the RHS (nullptr: direct form)
  double* stress_partial;
  double* ch_dev;
  hipEvent_t ev[4];
  double ms[3];
  hipEvent_t flag_ev;
  int* pinned_flags;
  double* pinned_scal;
  // Instantiated graph of one block between host checks: the preconditioner
  // of the checked iteration, then kCheckStride iterations up to their
  // ||r||^2 (nullptr until first needed; rebuilt when the preconditioner
  // changes).
  hipGraphExec_t block_graph;
  double block_graph_tol;
  int block_graph_key;  // stencil schedule (and tile) the graph was captured with
  // Direction-ring state the block was captured from and leaves behind: the
  // slot holding the current direction at the start (cur0) and end (cur1),
  // and the slots whose alpha_t p_t the block leaves pending, in order.
  bool block_graph_ring;
  bool block_graph_fused;  // the block starts after a fused r update (spectrum of r formed)
  bool block_graph_sr;     // the block keeps the residual as its x spectrum
  bool sr_failed;          // no room for the residual spectrum: real-space residual from then on
  int block_graph_cur0, block_graph_cur1;
  int block_graph_pend[8];
  int block_graph_npend;
  bool graph_failed;    // capture unavailable: blocks run kernel by kernel
  hipStream_t capture_stream;  // private stream the block is recorded on (replayed on the cell's)
  int graph_blocks;     // blocks replayed from the graph in the last run
  // Search-direction ring: the CG steps alpha_t p_t are added to x64 in one
  // FP64 pass (at a certification, when the ring is full, and at the end)
  // instead of every iteration; slot 0 is p.
  int nslot;
  void* pring[vxl::kMaxDirections];
  double* hist;  // alpha_t per slot use and system, kMaxDirections x R
  double* gram;  // mean-stress partials and totals of the six-unit-strain path
  // Block CG (method VXL_METHOD_BLOCK_CG): Gram partials, the small R x R
  // matrices [G | alpha | beta] and ||r||^2 totals, direction flags.
  int method;
  double* blk_partial;
  double* blk_small;
  int* blk_keep;
  double* partial_cert;  // ||f - K x||^2 partials of the certification stencil
  // Stacked cells: (lam, mu) of each cell's reference medium on the device
  // (the preconditioner reads it through pc.media).
  double* media;
  // Per-system residual history of the last run: (iteration, relative
  // residual); a negative iteration marks an FP64 true-residual evaluation.
  std::vector<std::vector<std::pair<int, double>>> trace;
};

