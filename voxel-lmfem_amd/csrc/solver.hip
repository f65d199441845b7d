// Block multi-RHS C_eff solver driver (host side), gfx950.
//
// Follows the reference workload (src/voxel_lmfem/_core/runtime/benchmark.py:
// 77-112): zero-mean RHS per unit macro strain, preconditioned CG with the
// safeguards of femlib/batched/pcg.py:20-252 (alpha frozen when p.Kp is not
// positive and finite, host checks every `check_stride` iterations), an FP64
// true-residual certificate ||f - K u||/||f|| <= tol on the zero-mean
// projected solution, and CH[:, j] = <C (E_j + eps(u_j))>.
//
// All load cases advance together (one launch per kernel for all RHS); each
// has its own alpha/beta and a device `active` flag so converged systems stop
// costing bandwidth.
//
// FP64 mode: operator, preconditioner and vectors in FP64.  When a system's
// recursive residual reaches tol its true residual is recomputed; if the
// certificate fails the system restarts from x0 = u (reference restart loop).
// MIXED mode: FP32 operator, FP32 preconditioner and FP32 CG on a correction
// x32 with reliable updates: whenever the FP32 recursive residual has dropped
// by kReliable since the last update (or reaches tol), x64 += x32, the FP64
// residual r64 = f - K64 x64 is recomputed and copied into r32.  The search
// direction is kept, so the Krylov space survives each update.
#include "vxl_cg.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>


namespace vxl {
namespace {

constexpr double kReliable = 1e-3;  // residual drop between FP64 updates (mixed)
constexpr int kCheckStride = 4;

// Graph replay of the blocks between host checks (read at every run):
// VXL_GRAPH=1 replays each block from one captured graph; VXL_GRAPH=2 runs
// the same blocks kernel by kernel under the replay's rules (direction ring
// flushed at every host check), bit-identical to 1 (a diagnostic); anything
// else launches kernel by kernel with the ring flushed only when full or at
// a certification.
int graph_mode() {
  const char* e = std::getenv("VXL_GRAPH");
  if (e && e[0] == '1') return 1;
  if (e && e[0] == '2') return 2;
  return 0;
}
bool graphs_enabled() { return graph_mode() == 1; }

// Slot of the next search direction: the lowest slot that holds neither the
// current direction nor a direction whose alpha_t p_t is still pending.
int next_slot(int cur, const std::vector<int>& pend, int nslot) {
  for (int i = 0; i < nslot; ++i)
    if (i != cur && std::find(pend.begin(), pend.end(), i) == pend.end()) return i;
  return -1;
}

// A replayed block starts with no pending directions (the ring is flushed at
// every host check) and rotates the slots with next_slot: kCheckStride
// preconditioner steps each open a slot and kCheckStride CG updates leave
// their directions pending.  Returns the slot the block ends on.
int ring_block_end(int cur, int nslot, std::vector<int>* pend_out) {
  std::vector<int> pend;
  for (int j = 0; j < kCheckStride; ++j) {
    cur = next_slot(cur, pend, nslot);
    if (cur < 0) return -1;
    pend.push_back(cur);
  }
  if (pend_out) *pend_out = pend;
  return cur;
}

// The slot a block ends on when it starts there (iterating ring_block_end
// from slot 0 reaches it within two blocks for nslot >= kCheckStride), so
// one captured graph serves every block after the first.
int ring_block_start(int nslot) {
  int c = 0;
  for (int k = 0; k < 2 * kMaxDirections; ++k) {
    const int e = ring_block_end(c, nslot, nullptr);
    if (e < 0) return -1;
    if (e == c) return c;
    c = e;
  }
  return -1;
}

// The direction ring is on unless VXL_DIRECTION_RING=0 or fewer than two
// slots could be allocated; with graph replay or its diagnostic twin it needs
// kCheckStride slots (one block's worth).
bool ring_on(const vxl_solver* s, int gmode) {
  const char* e = std::getenv("VXL_DIRECTION_RING");
  if (e && e[0] == '0') return false;
  return s->nslot >= (gmode != 0 ? kCheckStride : 2);
}

template <typename T>
struct Inner {
  static constexpr int precision = std::is_same<T, double>::value ? VXL_F64 : VXL_F32;
};

int free_all(vxl_solver* s) {
  if (s->cell) (void)hipStreamSynchronize(s->cell->stream);
  if (s->block_graph) (void)hipGraphExecDestroy(s->block_graph);
  s->block_graph = nullptr;
  if (s->capture_stream) (void)hipStreamDestroy(s->capture_stream);
  s->capture_stream = nullptr;
  double* bufs[] = {s->f, s->x64, s->r64, s->q64, s->scal, s->partial, s->macro_dev,
                    s->stress_partial, s->ch_dev, s->partial_st, s->partial_rz, s->rhs_sig, s->partial_fx,
                    s->dinv64, s->jmean, s->jrz, s->media};
  for (double* b : bufs)
    if (b) dev_free(b);
  if (s->mode == VXL_SOLVE_MIXED) {
    void* inner[] = {s->x, s->r, s->p, s->q};
    for (void* b : inner)
      if (b) dev_free(b);
  } else if (s->p) {
    dev_free(s->p);
  }
  if (s->dinv32) dev_free(s->dinv32);
  for (int i = 1; i < s->nslot; ++i)
    if (s->pring[i]) dev_free(s->pring[i]);
  s->nslot = 0;
  if (s->hist) dev_free(s->hist);
  if (s->gram) dev_free(s->gram);
  if (s->partial_cert) dev_free(s->partial_cert);
  if (s->blk_partial) dev_free(s->blk_partial);
  if (s->blk_small) dev_free(s->blk_small);
  if (s->blk_keep) dev_free(s->blk_keep);
  s->blk_partial = s->blk_small = nullptr;
  s->blk_keep = nullptr;
  if (s->active_dev) dev_free(s->active_dev);
  if (s->restart_dev) dev_free(s->restart_dev);
  if (s->stop_dev) dev_free(s->stop_dev);
  if (s->pc_ready) precond_release(&s->pc);
  for (auto& e : s->ev)
    if (e) (void)hipEventDestroy(e);
  if (s->flag_ev) (void)hipEventDestroy(s->flag_ev);
  if (s->pinned_flags) (void)hipHostFree(s->pinned_flags);
  if (s->pinned_scal) (void)hipHostFree(s->pinned_scal);
  return VXL_OK;
}

// The kernels of one PCG iteration, bound to a solver's buffers.
// One iteration: [stencil: q = K p, p.q] [alpha] [x += alpha p, r -= alpha q,
// ||r||^2] [latch] (host check / FP64 certification) [FFT passes of M r; the
// z pass reduces r.z] [beta] [last pass: p = M r + beta p].  With the fused
// four-step forward pass the x/r update and ||r||^2 ride on the forward x
// pass of the next preconditioner apply (its input is the updated r); a
// certification that replaces r repeats that pass on the new r.
template <typename T>
struct Steps {
  vxl_solver* s;
  double tol;
  hipStream_t st;
  int R;
  int64_t len, n;
  T *x, *r, *p, *q;
  int nblk, nblk_st, nblk_rz, nblk_fx;
  bool jacobi;
  const T* dinv;
  std::function<int()> form_beta;
  // Direction ring (ring == true): the stencil reads pcur, the preconditioner
  // step writes pnext = M r + beta pcur, alpha goes to hist_row and x is not
  // touched.  Otherwise pcur = pnext = p (in place) and x += alpha p here.
  bool ring = false;
  T* pcur;
  T* pnext;
  double* hist_row = nullptr;
  // Spectral residual (with the ring and the fused forward pass): r lives as
  // its x spectrum inside the preconditioner; the forward pass reads q and
  // updates that spectrum in place, so r is neither read nor written there.
  bool sr = false;

  Steps(vxl_solver* sv, double t, hipStream_t stream)
      : s(sv), tol(t), st(stream), R(sv->R), len(sv->len), n(sv->cell->grid.n),
        x(static_cast<T*>(sv->x)), r(static_cast<T*>(sv->r)), p(static_cast<T*>(sv->p)),
        q(static_cast<T*>(sv->q)), nblk(sv->nblk), nblk_st(apply_cg_blocks<T>(sv->cell, sv->R)),
        nblk_rz(precond_rz_blocks(&sv->pc)), jacobi(sv->precond_kind == VXL_PRECOND_JACOBI),
        dinv(nullptr), pcur(p), pnext(p) {
    if (jacobi)
      dinv = std::is_same<T, double>::value ? reinterpret_cast<const T*>(s->dinv64)
                                            : reinterpret_cast<const T*>(s->dinv32);
    nblk_fx = (jacobi || !s->partial_fx || !fused_update_enabled()) ? 0 : precond_fused_update_blocks(&s->pc);
    vxl_solver* so = s;
    const int nrz = nblk_rz;
    hipStream_t on = st;
    form_beta = [so, nrz, on]() -> int {
      return launch_finalize(so->partial_rz, nrz, kOpBeta, so->scal, so->R, so->active_dev, so->restart_dev, on);
    };
  }
  int jacobi_step() {
    return jacobi_cg_step<T>(r, pnext, dinv, n, R, s->partial, s->jmean, s->jrz, s->scal, s->active_dev,
                             s->restart_dev, st, pcur);
  }
  // [stencil, alpha, update, ||r||^2 with the latch at iteration `iter`]
  int cg_update(int iter, bool& forward_done) {
    int* act = s->active_dev;
    double* scal = s->scal;
    VXL_TRY(launch_apply_cg<T>(s->cell, R, pcur, q, act, s->partial_st, nullptr, st));
    VXL_TRY(launch_finalize(s->partial_st, nblk_st, kOpAlpha, scal, R, act, nullptr, st, 0, 0.0, nullptr,
                            ring ? hist_row : nullptr));
    forward_done = false;
    if (ring && nblk_fx > 0) {  // r -= alpha q inside the preconditioner's x pass; x from the ring
      VXL_TRY(precond_cg_forward<T>(&s->pc, R, r, static_cast<T*>(nullptr), nullptr, q, scal + kAlpha * R,
                                    s->partial_fx, act, st, sr));
      VXL_TRY(launch_finalize(s->partial_fx, nblk_fx, kOpRR, scal, R, act, nullptr, st, iter, tol, s->stop_dev));
      forward_done = true;
    } else if (ring) {
      VXL_TRY(launch_update_r<T>(r, q, len, R, scal + kAlpha * R, s->partial, act, st));
      VXL_TRY(launch_finalize(s->partial, nblk, kOpRR, scal, R, act, nullptr, st, iter, tol, s->stop_dev));
    } else if (nblk_fx > 0) {
      VXL_TRY(precond_cg_forward<T>(&s->pc, R, r, x, p, q, scal + kAlpha * R, s->partial_fx, act, st));
      VXL_TRY(launch_finalize(s->partial_fx, nblk_fx, kOpRR, scal, R, act, nullptr, st, iter, tol, s->stop_dev));
      forward_done = true;
    } else {
      VXL_TRY(launch_update_xr<T>(x, r, p, q, len, R, scal + kAlpha * R, s->partial, act, st));
      VXL_TRY(launch_finalize(s->partial, nblk, kOpRR, scal, R, act, nullptr, st, iter, tol, s->stop_dev));
    }
    return VXL_OK;
  }
  // [FFT passes of M r, beta, p = M r + beta p] (or the Jacobi step); the
  // beta kernel consumes the restart flags on the device.
  int precond_step(bool forward_done) {
    if (jacobi) return jacobi_step();
    int* act = s->active_dev;
    if (!forward_done)
      VXL_TRY(precond_cg_forward<T>(&s->pc, R, r, static_cast<T*>(nullptr), nullptr, nullptr, nullptr, nullptr, act,
                                    st, sr));
    return precond_cg_rest<T>(&s->pc, R, pnext, s->partial_rz, s->scal + kBeta * R, form_beta, act, st,
                              pcur == pnext ? nullptr : pcur, sr);
  }
};

// Whether a solve with these steps keeps the residual as its x spectrum: mixed
// mode (FP32 spectra), direction ring, fused forward pass, four-step x and y
// passes, and room for the R^ buffer (allocated here, once; if it does not
// fit, the solver keeps the real-space residual from then on, so eager runs
// and captured blocks always agree).  At 256^3 x 6 this takes the forward
// pass from 4.8 to 3.6 GB.
template <typename T>
bool spectral_on(vxl_solver* s, const Steps<T>& steps) {
  if (!(std::is_same<T, float>::value && steps.ring && steps.nblk_fx > 0 && !steps.jacobi && !s->sr_failed &&
        precond_spectral_residual_ok(&s->pc)))
    return false;
  if (precond_residual_spectrum(&s->pc) != VXL_OK) {
    (void)hipGetLastError();
    set_error("");
    s->sr_failed = true;
    return false;
  }
  return true;
}

// Blocks of kCheckStride iterations between host checks replay one captured
// graph: the preconditioner of the checked iteration, then kCheckStride
// iterations up to their ||r||^2.  No host work happens inside a block
// (flags, scalars and the latch live on the device), and the latch records
// the iteration relative to the block's start.  The graph bakes in the
// kernels and their arguments, so it is keyed on the latch tolerance and the
// stencil schedule (and, in mixed mode, the matrix-core switch).  With the
// direction ring the block is captured from the ring state every block after
// the first starts in (no pending directions, current direction in slot
// ring_block_start), and it leaves the pending slots ring_block_end lists.
// With the fused r update the block starts after the previous iteration's
// fused pass (the spectrum of r is formed); a block that starts after a
// certification replaced r forms it eagerly first.
int graph_key(const vxl_solver* s) {
  const int k = s->mode == VXL_SOLVE_MIXED ? 0 : 1;
  const int mfma = k == 0 && s->cell->mfma ? 1 : 0;
  return s->cell->variant[k] + 4 * (s->cell->tile_y[k] + 64 * (s->cell->zmin[k] + 128 * mfma));
}

// The block is recorded on a private non-blocking stream (the cell's stream
// may be the legacy default stream, which cannot be captured) and replayed
// on the cell's stream.
template <typename T>
int prepare_block_graph(vxl_solver* s, double tol) {
  const bool ring = ring_on(s, 1);
  Steps<T> probe(s, tol, s->cell->stream);
  probe.ring = ring;
  const bool sr = spectral_on<T>(s, probe);
  if (s->block_graph && s->block_graph_tol == tol && s->block_graph_key == graph_key(s) &&
      s->block_graph_ring == ring && s->block_graph_sr == sr)
    return VXL_OK;
  if (s->block_graph) {
    VXL_HIP(hipGraphExecDestroy(s->block_graph));
    s->block_graph = nullptr;
  }
  auto give_up = [&]() {  // capture is an optimisation: launch kernel by kernel instead
    (void)hipGetLastError();
    s->graph_failed = true;
    return VXL_OK;
  };
  if (!s->capture_stream && hipStreamCreateWithFlags(&s->capture_stream, hipStreamNonBlocking) != hipSuccess) {
    s->capture_stream = nullptr;
    return give_up();
  }
  hipStream_t st = s->capture_stream;
  Steps<T> steps(s, tol, st);
  steps.ring = ring;
  steps.sr = sr;
  const bool fused = steps.nblk_fx > 0;
  int cur = 0;
  std::vector<int> pend;
  if (ring) {
    cur = ring_block_start(s->nslot);
    if (cur < 0) return give_up();
  }
  const int cur0 = cur;
  auto slot = [&](int i) { return static_cast<T*>(s->pring[i]); };
  auto precond = [&](bool fd) -> int {
    if (ring) {
      const int next = next_slot(cur, pend, s->nslot);
      if (next < 0) return fail("search-direction ring exhausted");
      steps.pcur = slot(cur);
      steps.pnext = slot(next);
      cur = next;
    }
    return steps.precond_step(fd);
  };
  auto update = [&](int iter, bool& fd) -> int {
    if (ring) {
      steps.pcur = slot(cur);
      steps.hist_row = s->hist + int64_t(pend.size()) * s->R;
    }
    const int rc = steps.cg_update(iter, fd);
    if (ring) pend.push_back(cur);
    return rc;
  };
  hipGraph_t g = nullptr;
  if (hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed) != hipSuccess) return give_up();
  int rc = precond(fused);
  for (int j = 1; rc == VXL_OK && j <= kCheckStride; ++j) {
    bool fd = false;
    rc = update(j, fd);
    if (rc == VXL_OK && j < kCheckStride) rc = precond(fd);
  }
  const hipError_t end = hipStreamEndCapture(st, &g);
  hipGraphExec_t exec = nullptr;
  hipError_t inst = hipErrorUnknown;
  if (rc == VXL_OK && end == hipSuccess && g) inst = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
  if (g) (void)hipGraphDestroy(g);
  if (rc != VXL_OK || end != hipSuccess || inst != hipSuccess) {
    set_error("");
    return give_up();
  }
  s->block_graph = exec;
  s->block_graph_tol = tol;
  s->block_graph_key = graph_key(s);
  s->block_graph_ring = ring;
  s->block_graph_fused = fused;
  s->block_graph_sr = sr;
  s->block_graph_cur0 = cur0;
  s->block_graph_cur1 = cur;
  s->block_graph_npend = int(pend.size());
  for (size_t i = 0; i < pend.size(); ++i) s->block_graph_pend[i] = pend[i];
  return VXL_OK;
}

// End of a run (both methods): mean stress of every system (the zero-mean
// projection only for returned fields; B0 and the zero-mean RHS both
// annihilate constants), outputs, per-phase times.  Systems that never started
// (zero RHS) keep x64 = 0.
template <typename Upload>
int finish_run(vxl_solver* s, const double* macro_host, double* ch_host, int* iterations, int* restarts,
               double* residuals, double* u_out, const std::vector<int>& its, const std::vector<int>& nrest,
               const std::vector<double>& res_true, bool converged, Upload&& upload_flags,
               std::vector<int>& active) {
  vxl_cell* cell = s->cell;
  hipStream_t st = cell->stream;
  const int R = s->R;
  const int64_t len = s->len;
  const int64_t n = cell->grid.n;
  double* partial = s->partial;
  std::fill(active.begin(), active.end(), 1);
  VXL_TRY(upload_flags());
  if (u_out) VXL_TRY(launch_project_zero_mean(s->x64, n, R, partial, nullptr, st));
  // Six unit strains: <sigma_i>(e_k) = <C>_ik - f_i . u_k / |Omega| (f_i = -sum_e
  // V B0^T C_e e_i, the RHS of load i), two streaming passes instead of the
  // element gather.
  bool unit = R == 6;
  for (int k = 0; unit && k < 6; ++k)
    for (int i = 0; i < 6; ++i) unit = unit && macro_host[k * 6 + i] == (i == k ? 1.0 : 0.0);
  std::vector<double> ch(6 * R);
  double* gp = s->gram;                                 // 36 + 21 rows of kGramBlocks partials
  double* gsum = s->gram + 57 * size_t(kGramBlocks);    // 36 + 21 totals
  if (unit) {
    const double omega = double(n) * cell->volume;
    VXL_TRY(launch_gram6(s->f, s->x64, len, gp, -1.0 / omega, gsum, st));
    VXL_TRY(launch_stiffness_sums(cell, gp + 36 * size_t(kGramBlocks), kGramBlocks, st));
    VXL_TRY(launch_row_totals(gp + 36 * size_t(kGramBlocks), 21, kGramBlocks, 1.0 / double(n), gsum + 36, st));
    std::vector<double> g(57);
    VXL_HIP(hipMemcpyAsync(g.data(), gsum, sizeof(double) * 57, hipMemcpyDeviceToHost, st));
    VXL_HIP(hipStreamSynchronize(st));
    for (int k = 0; k < 6; ++k)
      for (int i = 0; i < 6; ++i) ch[k * 6 + i] = g[36 + pidx(i, k)] + g[i * 6 + k];
  } else {
    VXL_TRY(launch_avg_stress(cell, R, s->x64, s->macro_dev, s->stress_partial, s->ch_dev, st));
    VXL_HIP(hipMemcpyAsync(ch.data(), s->ch_dev, sizeof(double) * 6 * R, hipMemcpyDeviceToHost, st));
  }
  if (u_out)
    VXL_HIP(hipMemcpyAsync(u_out, s->x64, sizeof(double) * len * R, hipMemcpyDeviceToDevice, st));
  VXL_HIP(hipEventRecord(s->ev[3], st));
  VXL_HIP(hipStreamSynchronize(st));
  for (int k = 0; k < R; ++k)
    for (int i = 0; i < 6; ++i) ch_host[i * R + k] = ch[k * 6 + i];
  for (int k = 0; k < R; ++k) {
    if (iterations) iterations[k] = its[k];
    if (restarts) restarts[k] = std::max(0, nrest[k] - 1);
    if (residuals) residuals[k] = res_true[k];
  }
  float t0, t1, t2;
  VXL_HIP(hipEventElapsedTime(&t0, s->ev[0], s->ev[1]));
  VXL_HIP(hipEventElapsedTime(&t1, s->ev[1], s->ev[2]));
  VXL_HIP(hipEventElapsedTime(&t2, s->ev[2], s->ev[3]));
  s->ms[0] = t0;
  s->ms[1] = t1;
  s->ms[2] = t2;
  if (!converged) {
    set_error("CG did not certify ||f - K u||/||f|| <= tol within max_iter");
    return VXL_ERR_CONVERGENCE;
  }
  return VXL_OK;
}


template <typename T>
int run_impl(vxl_solver* s, const double* macro_host, double tol, int max_iter, double* ch_host,
             int* iterations, int* restarts, double* residuals, double* u_out) {
  vxl_cell* cell = s->cell;
  hipStream_t st = cell->stream;
  const int R = s->R;
  const int64_t len = s->len;
  const int64_t n = cell->grid.n;
  const bool mixed = s->mode == VXL_SOLVE_MIXED;
  T* x = static_cast<T*>(s->x);
  T* r = static_cast<T*>(s->r);
  T* p = static_cast<T*>(s->p);
  T* q = static_cast<T*>(s->q);
  double* scal = s->scal;
  double* partial = s->partial;
  const int nblk = s->nblk;
  int* act = s->active_dev;

  std::vector<int> active(R, 1), restart(R, 0), its(R, 0), nrest(R, 0);
  std::vector<double> ff(R), rr(R), rr_ref(R), res_true(R, 0.0);
  for (auto& t : s->trace) t.clear();
  // Host<->device flag and scalar traffic goes through pinned staging; every
  // upload waits for the previous one so the staging area is never rewritten
  // while a copy is pending.
  int* pin_flags = s->pinned_flags;
  int* pin_stop = s->pinned_flags + 2 * R;
  double* pin_scal = s->pinned_scal;
  auto upload_flags = [&]() -> int {
    VXL_HIP(hipEventSynchronize(s->flag_ev));
    for (int k = 0; k < R; ++k) {
      pin_flags[k] = active[k];
      pin_flags[R + k] = restart[k];
    }
    VXL_HIP(hipMemcpyAsync(act, pin_flags, sizeof(int) * R, hipMemcpyHostToDevice, st));
    VXL_HIP(hipMemcpyAsync(s->restart_dev, pin_flags + R, sizeof(int) * R, hipMemcpyHostToDevice, st));
    VXL_HIP(hipEventRecord(s->flag_ev, st));
    return VXL_OK;
  };
  auto upload_mask = [&](const std::vector<int>& mask) -> int {
    VXL_HIP(hipEventSynchronize(s->flag_ev));
    for (int k = 0; k < R; ++k) pin_flags[k] = mask[k];
    VXL_HIP(hipMemcpyAsync(act, pin_flags, sizeof(int) * R, hipMemcpyHostToDevice, st));
    VXL_HIP(hipEventRecord(s->flag_ev, st));
    return VXL_OK;
  };
  auto read_slot = [&](int slot, std::vector<double>& out) -> int {
    VXL_HIP(hipMemcpyAsync(pin_scal, scal + slot * R, sizeof(double) * R, hipMemcpyDeviceToHost, st));
    VXL_HIP(hipStreamSynchronize(st));
    for (int k = 0; k < R; ++k) out[k] = pin_scal[k];
    return VXL_OK;
  };
  auto any_active = [&]() { return std::any_of(active.begin(), active.end(), [](int a) { return a; }); };

  VXL_HIP(hipEventRecord(s->ev[0], st));
  VXL_HIP(hipMemcpyAsync(s->macro_dev, macro_host, sizeof(double) * 6 * R, hipMemcpyHostToDevice, st));
  VXL_HIP(hipMemsetAsync(scal, 0, sizeof(double) * kNumSlots * R, st));
  VXL_HIP(hipMemsetAsync(s->stop_dev, 0, sizeof(int) * R, st));
  VXL_TRY(upload_flags());
  // f = P0 rhs(E), r = f and ||f||^2 (one pass after the component means)
  VXL_TRY(launch_rhs(cell, R, s->macro_dev, s->f, s->rhs_sig, st));
  int nblk_ff = 0;
  if (mixed)
    VXL_TRY(launch_center_copy<float>(s->f, reinterpret_cast<float*>(r), n, R, partial, &nblk_ff, st));
  else
    VXL_TRY(launch_center_copy<double>(s->f, s->r64, n, R, partial, &nblk_ff, st));
  VXL_TRY(launch_finalize(partial + int64_t(3) * R * (vec_blocks(n) + 1), nblk_ff, kOpFF, scal, R, nullptr, nullptr,
                          st));
  VXL_TRY(read_slot(kFF, ff));
  for (int k = 0; k < R; ++k) {
    if (!(ff[k] > 0.0)) {  // zero RHS: exact zero solution (benchmark.py:79)
      active[k] = 0;
      res_true[k] = 0.0;
    }
    rr_ref[k] = ff[k];
  }
  VXL_TRY(upload_flags());
  // p = z = M r with rho = r.z: beta = 0 because rho_old = 0 (slots cleared).
  Steps<T> steps(s, tol, st);
  const bool jacobi = steps.jacobi;
  // Direction ring (unless VXL_DIRECTION_RING=0 asks for the per-iteration x
  // update).  Under graph replay (and its diagnostic twin) the ring is
  // flushed at every host check, so each block starts with no pending
  // directions.
  const int gmode = graph_mode();
  const bool ring = ring_on(s, gmode);
  const bool flush_each_check = ring && gmode != 0;
  steps.ring = ring;
  steps.sr = spectral_on<T>(s, steps);
  // x = 0.  With the ring, x64 is first written by the first flush (an
  // assignment) and the FP32 x is not used.
  bool x64_set = !ring;
  if (!ring) VXL_TRY(launch_fill_zero<double>(s->x64, len, R, nullptr, st));
  if (mixed && !ring) VXL_TRY(launch_fill_zero<float>(reinterpret_cast<float*>(x), len, R, act, st));
  std::vector<T*> slots(size_t(std::max(1, s->nslot)));
  for (int i = 0; i < s->nslot; ++i) slots[i] = static_cast<T*>(s->pring[i]);
  if (s->nslot < 1) slots[0] = p;
  int cur = 0;
  std::vector<int> pend;  // slots whose alpha_t p_t are not in x64 yet (hist rows in order)
  auto flush = [&]() -> int {
    if (!ring) return VXL_OK;
    if (pend.empty()) {
      if (!x64_set) VXL_TRY(launch_fill_zero<double>(s->x64, len, R, nullptr, st));
      x64_set = true;
      return VXL_OK;
    }
    DirectionSet<T> d{};
    for (size_t t = 0; t < pend.size(); ++t) d.p[t] = slots[pend[t]];
    d.count = int(pend.size());
    VXL_TRY(launch_flush<T>(s->x64, d, s->hist, R, len, !x64_set, st));
    x64_set = true;
    pend.clear();
    return VXL_OK;
  };
  steps.pcur = steps.pnext = slots[0];
  if (jacobi)
    VXL_TRY(steps.jacobi_step());
  else if (steps.sr)
    VXL_TRY(steps.precond_step(false));  // R^ from r, then the rest with p = z
  else
    VXL_TRY(precond_apply_cg<T>(&s->pc, R, r, slots[0], s->partial_rz, scal + kBeta * R, steps.form_beta, act, st));
  VXL_HIP(hipEventRecord(s->ev[1], st));

  int it = 0;
  // FP64 certification of the given systems.
  auto certify = [&](const std::vector<int>& which) -> int {
    std::vector<int> mask(R, 0);
    for (int k = 0; k < R; ++k) mask[k] = which[k] && active[k];
    VXL_TRY(upload_mask(mask));
    if (ring)
      VXL_TRY(flush());
    else if (mixed)
      VXL_TRY(launch_accumulate(s->x64, reinterpret_cast<float*>(x), len, R, act, st));
    // K annihilates the mean, so the zero-mean projection of x64 is left to
    // the end of the run; mixed mode needs the residual only in FP32.
    int rblocks = 0;
    VXL_TRY(launch_residual_apply(cell, R, s->x64, s->f, mixed ? reinterpret_cast<float*>(r) : nullptr,
                                  mixed ? nullptr : s->r64, act, s->partial_cert, &rblocks, st));
    VXL_TRY(launch_finalize(s->partial_cert, rblocks, kOpStore, scal, R, act, nullptr, st));
    std::vector<double> t(R);
    VXL_TRY(read_slot(kTmp, t));
    for (int k = 0; k < R; ++k) {
      if (!mask[k]) continue;
      const double rel = std::sqrt(t[k] / ff[k]);
      if (!std::isfinite(rel)) return fail("Nonfinite true residual");
      nrest[k] += 1;
      s->trace[k].push_back({-it, rel});  // negative iteration: FP64 true residual
      rr_ref[k] = t[k];
      if (rel <= tol) {
        active[k] = 0;
        res_true[k] = rel;
      } else if (!mixed) {
        restart[k] = 1;  // restart from x0 = u: r = f - K u already in r64
      }
    }
    return VXL_OK;
  };

  const bool graphs = gmode == 1 && !s->graph_failed;
  auto cg_update = [&](int iter, bool& fd) -> int {
    if (ring) {
      steps.pcur = slots[cur];
      steps.hist_row = s->hist + int64_t(pend.size()) * R;
      VXL_TRY(steps.cg_update(iter, fd));
      pend.push_back(cur);
      return VXL_OK;
    }
    return steps.cg_update(iter, fd);
  };
  auto precond_step = [&](bool fd) -> int {
    if (!ring) return steps.precond_step(fd);
    if (int(pend.size()) >= s->nslot) VXL_TRY(flush());  // no free slot for the next direction
    const int next = next_slot(cur, pend, s->nslot);
    if (next < 0) return fail("search-direction ring exhausted");
    steps.pcur = slots[cur];
    steps.pnext = slots[next];
    VXL_TRY(steps.precond_step(fd));
    cur = next;
    return VXL_OK;
  };
  auto block_graph = [&]() { return prepare_block_graph<T>(s, tol); };

  bool forward_done = false, pending = false;  // pending: M r of iteration `it` not formed yet
  s->graph_blocks = 0;
  while (any_active() && it < max_iter) {
    int base = -1;  // >= 0: this block ran as a graph starting after iteration `base`
    bool replay = pending && graphs && it % kCheckStride == 0 && it + kCheckStride <= max_iter;
    if (replay) {
      VXL_TRY(block_graph());
      // The ring must be in the state the block was captured from.
      replay = s->block_graph != nullptr && (!ring || (pend.empty() && cur == s->block_graph_cur0));
    }
    if (replay) {
      if (s->block_graph_fused && !forward_done)  // r was replaced: form its spectrum first
        VXL_TRY(precond_cg_forward<T>(&s->pc, R, r, static_cast<T*>(nullptr), nullptr, nullptr, nullptr, nullptr,
                                      act, st, steps.sr));
      VXL_HIP(hipGraphLaunch(s->block_graph, st));
      ++s->graph_blocks;
      base = it;
      it += kCheckStride;
      std::fill(restart.begin(), restart.end(), 0);
      forward_done = s->block_graph_fused;
      if (ring) {
        pend.assign(s->block_graph_pend, s->block_graph_pend + s->block_graph_npend);
        cur = s->block_graph_cur1;
      }
    } else {
      if (pending) {
        VXL_TRY(precond_step(forward_done));
        if (std::any_of(restart.begin(), restart.end(), [](int v) { return v; }))
          std::fill(restart.begin(), restart.end(), 0);  // consumed by the beta kernel
      }
      ++it;
      VXL_TRY(cg_update(it, forward_done));
    }
    pending = true;
    if (it == 1 || it % kCheckStride == 0 || it == max_iter) {
      VXL_TRY(read_slot(kRR, rr));
      VXL_HIP(hipMemcpyAsync(pin_stop, s->stop_dev, sizeof(int) * R, hipMemcpyDeviceToHost, st));
      VXL_HIP(hipStreamSynchronize(st));
      std::vector<int> which(R, 0), latched(R, 0);
      bool anyw = false;
      for (int k = 0; k < R; ++k) {
        if (!active[k]) continue;
        if (!std::isfinite(rr[k])) return fail("Nonfinite residual in CG recursion");
        const int at = pin_stop[k] ? (base >= 0 ? base + pin_stop[k] : pin_stop[k]) : it;
        s->trace[k].push_back({at, std::sqrt(rr[k] / ff[k])});
        its[k] = at;
        latched[k] = pin_stop[k] != 0;
        const bool drop = mixed && rr[k] < kReliable * kReliable * rr_ref[k];
        if (latched[k] || drop) {
          which[k] = 1;
          anyw = true;
        }
      }
      if (anyw) {
        if (mixed) {  // reliable update of every active system of a cell keeps them in step
          const int per = R / std::max(1, cell->groups);
          for (int g0 = 0; g0 < R; g0 += per) {
            bool any = false;
            for (int k = g0; k < g0 + per; ++k) any = any || which[k];
            if (any)
              for (int k = g0; k < g0 + per; ++k) which[k] = active[k];
          }
        }
        VXL_TRY(certify(which));
        forward_done = false;  // r was replaced
        if (steps.sr) {
          // Only the certified systems have a current r (the others live in
          // R^ alone): form R^ from r for them now, under the certification
          // mask the device flags still hold, and keep the rest.
          VXL_TRY(precond_cg_forward<T>(&s->pc, R, r, static_cast<T*>(nullptr), nullptr, nullptr, nullptr,
                                        nullptr, act, st, true));
          forward_done = true;
        }
        // A latched system skipped the iterations since its latch, so its
        // search direction is stale: restart it (beta = 0) if it continues.
        for (int k = 0; k < R; ++k)
          if (active[k] && latched[k]) restart[k] = 1;
        VXL_HIP(hipMemsetAsync(s->stop_dev, 0, sizeof(int) * R, st));  // eager, never captured
      }
      if (flush_each_check) VXL_TRY(flush());
      VXL_TRY(upload_flags());
      if (!any_active()) break;
    }
  }
  VXL_TRY(flush());
  VXL_HIP(hipEventRecord(s->ev[2], st));
  return finish_run(s, macro_host, ch_host, iterations, restarts, residuals, u_out, its, nrest, res_true,
                    !any_active(), upload_flags, active);
}

// ---- block CG over the load cases ---------------------------------------------
// Y[idx][j] = G[idx][idx]^-1 B[idx][j] for the columns j with cols[j] != 0 (G
// symmetric, R x R row major; other entries of Y are 0).  idx starts as the
// kept directions and loses every direction whose Cholesky pivot falls below
// 1e-12 of its diagonal (a direction dependent on the others).
void block_solve(const std::vector<double>& G, int R, std::vector<int>& idx, const std::vector<double>& B,
                 const std::vector<int>& cols, double sign, std::vector<double>& Y) {
  std::fill(Y.begin(), Y.end(), 0.0);
  for (;;) {
    const int m = int(idx.size());
    if (m == 0) return;
    std::vector<double> L(size_t(m) * m, 0.0);
    int bad = -1;
    for (int a = 0; a < m && bad < 0; ++a) {
      for (int b = 0; b <= a; ++b) {
        double v = 0.5 * (G[idx[a] * R + idx[b]] + G[idx[b] * R + idx[a]]);
        for (int k = 0; k < b; ++k) v -= L[a * m + k] * L[b * m + k];
        if (a == b) {
          if (!(v > 1e-12 * std::fabs(G[idx[a] * R + idx[a]]))) {
            bad = a;
            break;
          }
          L[a * m + a] = std::sqrt(v);
        } else {
          L[a * m + b] = v / L[b * m + b];
        }
      }
    }
    if (bad >= 0) {
      idx.erase(idx.begin() + bad);
      continue;
    }
    for (int j = 0; j < R; ++j) {
      if (!cols[j]) continue;
      std::vector<double> y(m);
      for (int a = 0; a < m; ++a) {  // L y = b
        double v = B[idx[a] * R + j];
        for (int k = 0; k < a; ++k) v -= L[a * m + k] * y[k];
        y[a] = v / L[a * m + a];
      }
      for (int a = m - 1; a >= 0; --a) {  // L^T x = y
        double v = y[a];
        for (int k = a + 1; k < m; ++k) v -= L[k * m + a] * y[k];
        y[a] = v / L[a * m + a];
      }
      for (int a = 0; a < m; ++a) Y[idx[a] * R + j] = sign * y[a];
    }
    return;
  }
}

// O'Leary's block PCG over the R load cases, conjugate-direction form: with
// Q = K P, alpha = (P^T Q)^-1 P^T R, X += P alpha, R -= Q alpha, Z = M R,
// beta = -(P^T Q)^-1 Q^T Z and P = Z + P beta, all R x R matrices formed on the
// host from FP64 Gram reductions.  A system whose residual certifies leaves
// the block (its direction is zeroed); a direction that becomes dependent on
// the others is dropped from the step.  Mixed mode runs the block recurrence
// in FP32 on a correction x32 with the reliable updates of the batched
// method (x64 += x32 and r = f - K x64 in FP64 when a residual has dropped
// 10^3 times or reaches tol; the direction block is kept); FP64 mode updates
// x64 directly and certifies the same way.
template <typename T>
int run_block_impl(vxl_solver* s, const double* macro_host, double tol, int max_iter, double* ch_host,
                   int* iterations, int* restarts, double* residuals, double* u_out) {
  vxl_cell* cell = s->cell;
  hipStream_t st = cell->stream;
  const int R = s->R;
  const int64_t len = s->len;
  const int64_t n = cell->grid.n;
  const bool mixed = s->mode == VXL_SOLVE_MIXED;
  if (R > block_max_rhs()) return fail("block CG handles at most 8 load cases per solver");
  if (s->nslot < 2) return fail("block CG needs one more field buffer than could be allocated");
  if (s->precond_kind == VXL_PRECOND_JACOBI) return fail("block CG runs with the FFT preconditioner");
  T* X = mixed ? static_cast<T*>(s->x) : reinterpret_cast<T*>(s->x64);
  T* r = static_cast<T*>(s->r);
  T* P = static_cast<T*>(s->p);
  T* Q = static_cast<T*>(s->q);
  T* Z = static_cast<T*>(s->pring[1]);
  double* scal = s->scal;
  double* partial = s->partial;
  int* act = s->active_dev;
  double* G = s->blk_small;               // R x R Gram result
  double* dalpha = s->blk_small + R * R;  // alpha or beta (device)
  double* drr = s->blk_small + 2 * R * R; // ||r||^2 totals
  const int gblk = block_gram_blocks(len);

  std::vector<int> active(R, 1), restart_flags(R, 0), its(R, 0), nrest(R, 0);
  std::vector<double> ff(R), rr(R), rr_ref(R), res_true(R, 0.0);
  for (auto& t : s->trace) t.clear();
  int* pin_flags = s->pinned_flags;
  double* pin_scal = s->pinned_scal;
  auto upload_flags = [&]() -> int {
    VXL_HIP(hipEventSynchronize(s->flag_ev));
    for (int k = 0; k < R; ++k) {
      pin_flags[k] = active[k];
      pin_flags[R + k] = 0;
    }
    VXL_HIP(hipMemcpyAsync(act, pin_flags, sizeof(int) * R, hipMemcpyHostToDevice, st));
    VXL_HIP(hipMemcpyAsync(s->restart_dev, pin_flags + R, sizeof(int) * R, hipMemcpyHostToDevice, st));
    VXL_HIP(hipEventRecord(s->flag_ev, st));
    return VXL_OK;
  };
  auto read_doubles = [&](const double* src, int count, std::vector<double>& out) -> int {
    out.resize(count);
    VXL_HIP(hipMemcpyAsync(out.data(), src, sizeof(double) * count, hipMemcpyDeviceToHost, st));
    VXL_HIP(hipStreamSynchronize(st));
    return VXL_OK;
  };
  auto upload_small = [&](const std::vector<double>& v) -> int {
    VXL_HIP(hipMemcpyAsync(dalpha, v.data(), sizeof(double) * v.size(), hipMemcpyHostToDevice, st));
    VXL_HIP(hipStreamSynchronize(st));  // v is host memory of this frame
    return VXL_OK;
  };
  auto upload_keep = [&](const std::vector<int>& keep) -> int {
    VXL_HIP(hipMemcpyAsync(s->blk_keep, keep.data(), sizeof(int) * R, hipMemcpyHostToDevice, st));
    VXL_HIP(hipStreamSynchronize(st));
    return VXL_OK;
  };
  auto any_active = [&]() { return std::any_of(active.begin(), active.end(), [](int a) { return a; }); };
  auto gram = [&](const T* A, const T* B, std::vector<double>& out) -> int {
    VXL_TRY(launch_block_gram<T>(A, B, len, R, s->blk_partial, G, st));
    return read_doubles(G, R * R, out);
  };

  VXL_HIP(hipEventRecord(s->ev[0], st));
  VXL_HIP(hipMemcpyAsync(s->macro_dev, macro_host, sizeof(double) * 6 * R, hipMemcpyHostToDevice, st));
  VXL_HIP(hipMemsetAsync(scal, 0, sizeof(double) * kNumSlots * R, st));
  VXL_TRY(upload_flags());
  VXL_TRY(launch_rhs(cell, R, s->macro_dev, s->f, s->rhs_sig, st));
  int nblk_ff = 0;
  VXL_TRY(launch_center_copy<T>(s->f, r, n, R, partial, &nblk_ff, st));
  VXL_TRY(launch_finalize(partial + int64_t(3) * R * (vec_blocks(n) + 1), nblk_ff, kOpFF, scal, R, nullptr, nullptr,
                          st));
  VXL_TRY(read_doubles(scal + kFF * R, R, ff));
  for (int k = 0; k < R; ++k) {
    if (!(ff[k] > 0.0)) active[k] = 0;  // zero RHS: exact zero solution (benchmark.py:79)
    rr_ref[k] = ff[k];
  }
  VXL_TRY(upload_flags());
  VXL_TRY(launch_fill_zero<double>(s->x64, len, R, nullptr, st));
  if (mixed) VXL_TRY(launch_fill_zero<float>(reinterpret_cast<float*>(s->x), len, R, nullptr, st));
  VXL_HIP(hipEventRecord(s->ev[1], st));

  std::vector<double> Gpq, Gpr, Gqz, coef(size_t(R) * R), rr_read;
  std::vector<int> keep(R, 0);
  bool fresh = true;  // P = M r to be formed (the first block)
  int it = 0;
  auto certify = [&]() -> int {
    std::vector<int> mask = active;
    if (mixed) VXL_TRY(launch_accumulate(s->x64, reinterpret_cast<float*>(s->x), len, R, act, st));
    int rblocks = 0;
    VXL_TRY(launch_residual_apply(cell, R, s->x64, s->f, mixed ? reinterpret_cast<float*>(s->r) : nullptr,
                                  mixed ? nullptr : s->r64, act, s->partial_cert, &rblocks, st));
    VXL_TRY(launch_finalize(s->partial_cert, rblocks, kOpStore, scal, R, act, nullptr, st));
    std::vector<double> t;
    VXL_TRY(read_doubles(scal + kTmp * R, R, t));
    for (int k = 0; k < R; ++k) {
      if (!mask[k]) continue;
      const double rel = std::sqrt(t[k] / ff[k]);
      if (!std::isfinite(rel)) return fail("Nonfinite true residual");
      nrest[k] += 1;
      s->trace[k].push_back({-it, rel});
      rr_ref[k] = t[k];
      if (rel <= tol) {
        active[k] = 0;
        res_true[k] = rel;
      }
    }
    return upload_flags();
  };

  while (any_active() && it < max_iter) {
    if (fresh) {  // P = Z = M r for the active systems, 0 for the others
      VXL_TRY(precond_apply_masked(&s->pc, R, r, Z, act, st));
      keep = active;
      std::fill(coef.begin(), coef.end(), 0.0);
      VXL_TRY(upload_small(coef));
      VXL_TRY(upload_keep(keep));
      VXL_TRY(launch_block_direction<T>(P, Z, dalpha, s->blk_keep, len, R, st));
      fresh = false;
    }
    VXL_TRY(launch_apply<T>(cell, R, P, Q, st));
    VXL_TRY(gram(P, Q, Gpq));
    VXL_TRY(gram(P, r, Gpr));
    std::vector<int> idx;
    for (int k = 0; k < R; ++k)
      if (keep[k]) idx.push_back(k);
    block_solve(Gpq, R, idx, Gpr, active, 1.0, coef);  // alpha
    VXL_TRY(upload_small(coef));
    VXL_TRY(launch_block_update<T>(X, r, P, Q, dalpha, act, len, R, s->blk_partial, st));
    VXL_TRY(launch_row_totals(s->blk_partial, R, gblk, 1.0, drr, st));
    ++it;
    VXL_TRY(read_doubles(drr, R, rr_read));
    bool check = false;
    for (int k = 0; k < R; ++k) {
      if (!active[k]) continue;
      if (!std::isfinite(rr_read[k])) return fail("Nonfinite residual in block CG recursion");
      its[k] = it;
      s->trace[k].push_back({it, std::sqrt(rr_read[k] / ff[k])});
      if (rr_read[k] <= tol * tol * ff[k]) check = true;                         // reached tol
      if (mixed && rr_read[k] < kReliable * kReliable * rr_ref[k]) check = true;  // reliable update
    }
    if (check) {
      VXL_TRY(certify());  // r of every active system is now its true residual
      if (!any_active()) break;
    }
    // Next direction block: beta = -(P^T Q)^-1 Q^T Z over the directions kept in
    // the step, for every active system.  After a reliable update the block
    // keeps its directions, as the per-case method keeps its search direction:
    // the new directions are made K-conjugate to the last block from the true
    // residuals.
    VXL_TRY(precond_apply_masked(&s->pc, R, r, Z, act, st));
    VXL_TRY(gram(Q, Z, Gqz));
    block_solve(Gpq, R, idx, Gqz, active, -1.0, coef);
    keep = active;
    VXL_TRY(upload_small(coef));
    VXL_TRY(upload_keep(keep));
    VXL_TRY(launch_block_direction<T>(P, Z, dalpha, s->blk_keep, len, R, st));
  }
  VXL_HIP(hipEventRecord(s->ev[2], st));
  if (!mixed) {
    // x64 is X itself; nothing to flush.
  }
  return finish_run(s, macro_host, ch_host, iterations, restarts, residuals, u_out, its, nrest, res_true,
                    !any_active(), upload_flags, active);
}

}  // namespace

int solver_create(vxl_solver* s) {
  vxl_cell* cell = s->cell;
  const int R = s->R;
  s->len = 3 * cell->grid.n;
  const size_t b64 = sizeof(double) * size_t(s->len) * R;
  VXL_HIP(dev_malloc(&s->f, b64));
  VXL_HIP(dev_malloc(&s->x64, b64));
  if (s->mode != VXL_SOLVE_MIXED) VXL_HIP(dev_malloc(&s->r64, b64));
  VXL_HIP(dev_malloc(&s->q64, b64));
  if (s->mode == VXL_SOLVE_MIXED) {
    const size_t b32 = sizeof(float) * size_t(s->len) * R;
    VXL_HIP(dev_malloc(&s->x, b32));
    VXL_HIP(dev_malloc(&s->r, b32));
    VXL_HIP(dev_malloc(&s->p, b32));
    VXL_HIP(dev_malloc(&s->q, b32));
  } else {
    s->x = s->x64;
    s->r = s->r64;
    s->q = s->q64;
    VXL_HIP(dev_malloc(&s->p, b64));
  }
  // Direction ring: slot 0 is p; the other slots are optional (a failed
  // allocation leaves a shorter ring, and fewer than two slots mean the
  // solver updates x every iteration instead).
  s->pring[0] = s->p;
  s->nslot = 1;
  {
    const size_t bp = (s->mode == VXL_SOLVE_MIXED ? sizeof(float) : sizeof(double)) * size_t(s->len) * R;
    for (int i = 1; i < kMaxDirections; ++i) {
      if (dev_malloc(&s->pring[i], bp) != hipSuccess) {
        (void)hipGetLastError();
        s->pring[i] = nullptr;
        break;
      }
      s->nslot = i + 1;
    }
  }
  VXL_HIP(dev_malloc(&s->hist, sizeof(double) * kMaxDirections * R));
  VXL_HIP(dev_malloc(&s->gram, sizeof(double) * (57 * size_t(kGramBlocks) + 64)));
  VXL_HIP(dev_malloc(&s->partial_cert, sizeof(double) * size_t(residual_apply_blocks(cell, R)) * R));
  s->nblk = vec_blocks(s->len);
  const int nblk_rows = vec_blocks(cell->grid.n);
  // row sums and means of the projection, then one ||f||^2 partial per block
  const size_t np = size_t(3 * R * (nblk_rows + 1) + s->nblk * R);
  VXL_HIP(dev_malloc(&s->partial, sizeof(double) * np));
  {
    // sized for any tile / chunk depth the cell can be switched to
    const int st_blocks = apply_cg_blocks_max(cell, R, s->mode == VXL_SOLVE_MIXED ? VXL_F32 : VXL_F64);
    VXL_HIP(dev_malloc(&s->partial_st, sizeof(double) * size_t(st_blocks) * R));
  }
  VXL_HIP(dev_malloc(&s->scal, sizeof(double) * kNumSlots * R));
  VXL_HIP(dev_malloc(&s->active_dev, sizeof(int) * R));
  VXL_HIP(dev_malloc(&s->restart_dev, sizeof(int) * R));
  VXL_HIP(dev_malloc(&s->macro_dev, sizeof(double) * 6 * R));
  if (const int64_t ns = rhs_scratch_size(cell, R)) VXL_HIP(dev_malloc(&s->rhs_sig, sizeof(double) * size_t(ns)));
  VXL_HIP(dev_malloc(&s->stress_partial, sizeof(double) * size_t(stress_partial_size(cell, R))));
  VXL_HIP(dev_malloc(&s->ch_dev, sizeof(double) * 6 * R));
  for (auto& e : s->ev) VXL_HIP(hipEventCreate(&e));
  VXL_HIP(hipEventCreate(&s->flag_ev));
  VXL_HIP(hipHostMalloc(reinterpret_cast<void**>(&s->pinned_flags), sizeof(int) * 3 * R));
  VXL_HIP(dev_malloc(&s->stop_dev, sizeof(int) * R));
  VXL_HIP(hipHostMalloc(reinterpret_cast<void**>(&s->pinned_scal), sizeof(double) * R));
  s->trace.assign(R, {});
  s->pc.cell = cell;
  s->pc.precision = s->mode == VXL_SOLVE_MIXED ? VXL_F32 : VXL_F64;
  s->pc.nrhs_max = R;
  s->pc_ready = true;
  VXL_TRY(precond_create(&s->pc));
  VXL_HIP(dev_malloc(&s->partial_rz, sizeof(double) * size_t(precond_rz_blocks(&s->pc)) * R));
  if (const int fx = precond_fused_update_blocks(&s->pc))
    VXL_HIP(dev_malloc(&s->partial_fx, sizeof(double) * size_t(fx) * R));
  return VXL_OK;
}

int solver_run(vxl_solver* s, const double* macro_host, double tol, int max_iter, double* ch_host,
               int* iterations, int* restarts, double* residuals, double* u_out) {
  if (s->mode == VXL_SOLVE_MIXED)
    return s->method == VXL_METHOD_BLOCK_CG
               ? run_block_impl<float>(s, macro_host, tol, max_iter, ch_host, iterations, restarts, residuals, u_out)
               : run_impl<float>(s, macro_host, tol, max_iter, ch_host, iterations, restarts, residuals, u_out);
  return s->method == VXL_METHOD_BLOCK_CG
             ? run_block_impl<double>(s, macro_host, tol, max_iter, ch_host, iterations, restarts, residuals, u_out)
             : run_impl<double>(s, macro_host, tol, max_iter, ch_host, iterations, restarts, residuals, u_out);
}

int solver_release(vxl_solver* s) { return free_all(s); }

int solver_set_method(vxl_solver* s, int method) {
  if (method == VXL_METHOD_BLOCK_CG && !(s->blk_partial && s->blk_small && s->blk_keep)) {
    if (s->R > block_max_rhs()) return fail("block CG handles at most 8 load cases per solver");
    if (s->nslot < 2) return fail("block CG needs one more field buffer than could be allocated");
    const int64_t gb = block_gram_blocks(s->len);
    if (!s->blk_partial) VXL_HIP(dev_malloc(&s->blk_partial, sizeof(double) * size_t(s->R) * s->R * gb));
    if (!s->blk_small) VXL_HIP(dev_malloc(&s->blk_small, sizeof(double) * size_t(2 * s->R * s->R + s->R)));
    if (!s->blk_keep) VXL_HIP(dev_malloc(&s->blk_keep, sizeof(int) * size_t(s->R)));
  }
  s->method = method;
  return VXL_OK;
}

int solver_set_media(vxl_solver* s, const double* lam, const double* mu) {
  const int G = s->cell->groups;
  for (int g = 0; g < G; ++g)
    if (!(lam[g] >= 0.0) || !(mu[g] > 0.0) || !std::isfinite(lam[g]) || !std::isfinite(mu[g]))
      return fail("reference medium requires lam >= 0, mu > 0");
  if (s->block_graph) {  // captured without the per-cell media
    VXL_HIP(hipStreamSynchronize(s->cell->stream));
    VXL_HIP(hipGraphExecDestroy(s->block_graph));
    s->block_graph = nullptr;
  }
  std::vector<double> m(2 * size_t(G));
  for (int g = 0; g < G; ++g) {
    m[2 * g] = lam[g];
    m[2 * g + 1] = mu[g];
  }
  if (!s->media) VXL_HIP(dev_malloc(&s->media, sizeof(double) * m.size()));
  VXL_HIP(hipMemcpyAsync(s->media, m.data(), sizeof(double) * m.size(), hipMemcpyHostToDevice, s->cell->stream));
  VXL_HIP(hipStreamSynchronize(s->cell->stream));
  s->pc.media = s->media;
  s->pc.groups = G;
  return VXL_OK;
}

int solver_prepare_graph(vxl_solver* s, double tol) {
  if (!graphs_enabled() || s->graph_failed || s->method != VXL_METHOD_CG) return VXL_OK;
  return s->mode == VXL_SOLVE_MIXED ? prepare_block_graph<float>(s, tol) : prepare_block_graph<double>(s, tol);
}

int solver_set_preconditioner(vxl_solver* s, int kind) {
  if (s->block_graph) {  // captured with the old preconditioner's kernels
    VXL_HIP(hipStreamSynchronize(s->cell->stream));
    VXL_HIP(hipGraphExecDestroy(s->block_graph));
    s->block_graph = nullptr;
  }
  // FFT storage precision: rebuild the preconditioner when it changes.
  const int want = kind == VXL_PRECOND_FFT_BF16 ? VXL_F32_BF16
                                                : (s->mode == VXL_SOLVE_MIXED ? VXL_F32 : VXL_F64);
  if (kind != VXL_PRECOND_JACOBI && s->pc.precision != want) {
    precond_release(&s->pc);
    s->pc.precision = want;
    const int rc = precond_create(&s->pc);
    if (rc != VXL_OK) {  // restore the default so the solver stays usable
      const std::string msg = vxl_last_error();
      precond_release(&s->pc);
      s->pc.precision = s->mode == VXL_SOLVE_MIXED ? VXL_F32 : VXL_F64;
      VXL_TRY(precond_create(&s->pc));
      set_error(msg);
      return rc;
    }
  }
  if (kind == VXL_PRECOND_JACOBI && !s->dinv64) {
    const int64_t len = 3 * s->cell->grid.n;
    VXL_HIP(dev_malloc(&s->dinv64, sizeof(double) * size_t(len)));
    if (s->mode == VXL_SOLVE_MIXED) VXL_HIP(dev_malloc(&s->dinv32, sizeof(float) * size_t(len)));
    VXL_HIP(dev_malloc(&s->jmean, sizeof(double) * 3 * s->R));
    VXL_HIP(dev_malloc(&s->jrz, sizeof(double) * size_t(jacobi_blocks(s->cell->grid.n)) * s->R));
    VXL_TRY(launch_jacobi_diag(s->cell, s->dinv64, s->dinv32, s->cell->stream));
    VXL_HIP(hipStreamSynchronize(s->cell->stream));
  }
  s->precond_kind = kind;
  return VXL_OK;
}

}  // namespace vxl
