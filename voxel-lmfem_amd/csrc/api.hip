// extern "C" entry points of libvxl (declared in include/vxl.h).
#include "vxl_cg.h"

#include <cmath>
#include <cstring>
#include <atomic>
#include <algorithm>
#include <mutex>
#include <thread>

namespace vxl {

static thread_local std::string g_error;

void set_error(const std::string& message) { g_error = message; }

int fail(const std::string& message) {
  g_error = message;
  if (message.find("hip") != std::string::npos || message.find("HIP") != std::string::npos)
    return VXL_ERR_HIP;
  return VXL_ERR_ARG;
}

int solver_create(vxl_solver* s);
int solver_run(vxl_solver* s, const double* macro_host, double tol, int max_iter, double* ch_host,
               int* iterations, int* restarts, double* residuals, double* u_out);
int solver_release(vxl_solver* s);
int solver_set_preconditioner(vxl_solver* s, int kind);
int solver_prepare_graph(vxl_solver* s, double tol);
int solver_set_method(vxl_solver* s, int method);
int solver_set_media(vxl_solver* s, const double* lam, const double* mu);

namespace {

template <typename Idx>
__global__ void index_max_kernel(const Idx* index, int64_t n, unsigned int* out) {
  unsigned int m = 0;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const long long v = static_cast<long long>(index[i]);
    const unsigned int u = v < 0 ? 0xffffffffu : static_cast<unsigned int>(v);
    m = u > m ? u : m;
  }
  atomicMax(out, m);
}

// Device buffer returned to the allocation cache on scope exit, after the
// stream that used it has drained (temporaries of the non-solver entry points).
struct DevBuf {
  void* p = nullptr;
  hipStream_t stream = nullptr;
  explicit DevBuf(hipStream_t s) : stream(s) {}
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() {
    if (p) {
      (void)hipStreamSynchronize(stream);
      dev_free(p);
    }
  }
};

int check_index(vxl_cell* c) {
  DevBuf buf(c->stream);
  VXL_HIP(dev_malloc(&buf.p, sizeof(unsigned int)));
  unsigned int* d = static_cast<unsigned int*>(buf.p);
  unsigned int h = 0;
  VXL_HIP(hipMemcpyAsync(d, &h, sizeof(h), hipMemcpyHostToDevice, c->stream));
  const int64_t n = c->grid.n;
  const int blocks = int(std::min<int64_t>((n + 255) / 256, 1024));
  switch (c->mat.index_kind) {
    case VXL_INDEX_U8:
      hipLaunchKernelGGL(index_max_kernel<uint8_t>, dim3(blocks), dim3(256), 0, c->stream,
                         static_cast<const uint8_t*>(c->mat.index), n, d);
      break;
    case VXL_INDEX_U16:
      hipLaunchKernelGGL(index_max_kernel<uint16_t>, dim3(blocks), dim3(256), 0, c->stream,
                         static_cast<const uint16_t*>(c->mat.index), n, d);
      break;
    case VXL_INDEX_I32:
      hipLaunchKernelGGL(index_max_kernel<int32_t>, dim3(blocks), dim3(256), 0, c->stream,
                         static_cast<const int32_t*>(c->mat.index), n, d);
      break;
  }
  VXL_HIP(hipGetLastError());
  VXL_HIP(hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  VXL_HIP(hipStreamSynchronize(c->stream));
  if (h >= static_cast<unsigned int>(c->rows)) return fail("material row index out of range");
  return VXL_OK;
}

// Shared part of the two cell constructors: argument checks, the handle, and
// the padded FP64/FP32 tables filled by `fill` (host upload or device pad).
template <typename Fill>
int cell_create_common(vxl_cell** out, int nz, int ny, int nx, double hx, double hy, double hz, int index_kind,
                       const void* index_dev, int rows, const float* scale_dev, void* stream, Fill fill) {
  g_error.clear();
  if (!out) return fail("out handle is NULL");
  *out = nullptr;
  if (nz < 2 || ny < 2 || nx < 2) return fail("grid dimensions must be >= 2");
  if (int64_t(nz) * ny * nx > (int64_t(1) << 40) || int64_t(ny) * nx > INT32_MAX) return fail("grid too large");
  if (!(hx > 0 && hy > 0 && hz > 0) || !std::isfinite(hx * hy * hz))
    return fail("voxel sizes must be finite and positive");
  if (index_kind != VXL_INDEX_U8 && index_kind != VXL_INDEX_U16 && index_kind != VXL_INDEX_I32)
    return fail("index_kind must be VXL_INDEX_U8, VXL_INDEX_U16 or VXL_INDEX_I32");
  if (!index_dev || rows < 1) return fail("index, coefficient table and rows are required");
  auto* c = new vxl_cell();
  c->grid.nx = nx;
  c->grid.ny = ny;
  c->grid.nz = nz;
  c->grid.n = int64_t(nx) * ny * nz;
  c->h[0] = hx;
  c->h[1] = hy;
  c->h[2] = hz;
  c->volume = hx * hy * hz;
  c->rows = rows;
  c->stream = static_cast<hipStream_t>(stream);
  c->mat.index_kind = index_kind;
  c->mat.index = index_dev;
  c->mat.scale = scale_dev;
  c->c64_owned = nullptr;
  c->c32_owned = nullptr;
  c->variant[0] = default_stencil_variant(VXL_F32);
  c->variant[1] = default_stencil_variant(VXL_F64);
  c->tile_y[0] = default_stencil_tile(VXL_F32);
  c->tile_y[1] = default_stencil_tile(VXL_F64);
  c->zmin[0] = c->zmin[1] = kDefaultChunkPlanes;
  c->element = VXL_ELEMENT_FI;
  c->hg_coef = 0.1;
  c->mfma = 0;
  c->mfma_table = nullptr;
  c->mfma_exp = nullptr;
  c->frame_index = nullptr;
  c->frame_table = nullptr;
  c->frame_quat = nullptr;
  c->frame_forms = 0;
  c->groups = 1;
  (void)hipGetDevice(&c->device);
  auto cleanup = [&]() {
    (void)hipStreamSynchronize(c->stream);
    if (c->c64_owned) dev_free(c->c64_owned);
    if (c->c32_owned) dev_free(c->c32_owned);
    delete c;
  };
  if (dev_malloc(&c->c64_owned, size_t(rows) * kRowWidth * sizeof(double)) != hipSuccess ||
      dev_malloc(&c->c32_owned, size_t(rows) * kRowWidth * sizeof(float)) != hipSuccess) {
    cleanup();
    g_error = "hipMalloc failed for the coefficient table";
    return VXL_ERR_HIP;
  }
  int rc = fill(c);
  if (rc == VXL_OK) {
    c->mat.c64 = c->c64_owned;
    c->mat.c32 = c->c32_owned;
    rc = check_index(c);
  }
  if (rc != VXL_OK) {
    std::string msg = g_error;
    cleanup();
    g_error = msg;
    return rc;
  }
  *out = c;
  return VXL_OK;
}

// 21-coefficient FP64 rows -> padded FP64 and FP32 tables; bad[0] = 1 on a
// nonfinite coefficient.
__global__ void __launch_bounds__(256)
pad_rows_kernel(const double* c21, int64_t rows, double* t64, float* t32, int* bad) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < rows * kRowWidth;
       i += int64_t(gridDim.x) * blockDim.x) {
    const int64_t r = i / kRowWidth;
    const int k = int(i - r * kRowWidth);
    const double v = k < 21 ? c21[r * 21 + k] : 0.0;
    if (!isfinite(v)) bad[0] = 1;
    t64[i] = v;
    t32[i] = float(v);
  }
}

}  // namespace
}  // namespace vxl

// ---- allocation-free zero-mean projection ------------------------------------
// The partial sums and means of vxl_project_zero_mean_noalloc live in this
// file-scope device array instead of an allocated workspace.  Each GPU the
// library runs on holds its own copy (768 KiB) until the process ends; every
// call on one GPU shares it, so calls that overlap on one GPU race.  The
// partials and means are written before they are read in every call, so
// nothing has to be cleared between calls.
namespace vxl {
constexpr int kNoAllocMaxRhs = 16;
constexpr size_t kNoAllocScratchDoubles = size_t(3) * kNoAllocMaxRhs * (kVecBlocks + 1);
}  // namespace vxl

static __device__ __align__(256) double scratch[vxl::kNoAllocScratchDoubles];

namespace vxl {
namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on `device`, looked up once per device with that
// device current (the host-side name is only a stand-in, never a device
// address).
int static_scratch(int device, double** out) {
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (device < 0 || device >= kMaxDevices) return fail("static scratch: device ordinal out of range");
  if (!g_scratch_ptr[device]) {
    int prev = -1;
    VXL_HIP(hipGetDevice(&prev));
    if (prev != device) VXL_HIP(hipSetDevice(device));
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (prev != device) (void)hipSetDevice(prev);
    if (e != hipSuccess || !ptr)
      return fail(std::string("HIP error in hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = static_cast<double*>(ptr);
  }
  *out = g_scratch_ptr[device];
  return VXL_OK;
}

}  // namespace
}  // namespace vxl

using namespace vxl;

extern "C" {

const char* vxl_last_error(void) { return g_error.c_str(); }

int vxl_version(void) { return 100; }

int vxl_cached_bytes(uint64_t* bytes) {
  if (!bytes) return fail("vxl_cached_bytes: invalid arguments");
  *bytes = cached_bytes();
  return VXL_OK;
}

int vxl_release_cached_memory(void) {
  release_cached();
  return VXL_OK;
}

int vxl_label_rows(int nz, int ny, int nx, int label_bytes, const void* labels_dev, long long label_lo, int span,
                   const int* lut_host, int index_kind, void* index_dev, long long local_base, int* present_host,
                   int* undefined_host, int* samples_host, void* stream) {
  g_error.clear();
  if (nz < 2 || ny < 2 || nx < 2 || (label_bytes != 4 && label_bytes != 8) || !labels_dev || span < 1 ||
      span > (1 << 24) || !lut_host || !index_dev || !present_host || !undefined_host || !samples_host ||
      (index_kind != VXL_INDEX_U8 && index_kind != VXL_INDEX_U16 && index_kind != VXL_INDEX_I32))
    return fail("vxl_label_rows: invalid arguments");
  int max_row = -1;
  bool local = false;
  for (int i = 0; i < span; ++i) {
    if (lut_host[i] < -2) return fail("vxl_label_rows: lookup entries must be rows, -1 or -2");
    max_row = std::max(max_row, lut_host[i]);
    local = local || lut_host[i] == -2;
  }
  const int64_t n = int64_t(nz) * ny * nx;
  if (local) {
    if (index_kind != VXL_INDEX_I32 || local_base <= max_row || local_base + n > INT32_MAX)
      return fail("vxl_label_rows: local-frame rows need an int32 index and local_base above the phase rows");
  }
  const int limit = index_kind == VXL_INDEX_U8 ? 255 : index_kind == VXL_INDEX_U16 ? 65535 : INT32_MAX;
  if (max_row > limit) return fail("vxl_label_rows: row index does not fit the index kind");
  return label_rows(nz, ny, nx, label_bytes, labels_dev, label_lo, span, lut_host, index_kind, index_dev,
                    present_host, undefined_host, samples_host, local_base, static_cast<hipStream_t>(stream));
}

int vxl_local_rows(int nz, int ny, int nx, int label_bytes, const void* labels_dev, long long label_lo, int span,
                   const int* slot_host, int phases, const double* c21_host, int mode, const double* a_dev,
                   const double* b_dev, long long stride, double* rows_dev, long long local_base, int* bad_host,
                   void* stream) {
  g_error.clear();
  if (nz < 1 || ny < 1 || nx < 1 || (label_bytes != 4 && label_bytes != 8) || !labels_dev || span < 1 ||
      span > (1 << 24) || !slot_host || phases < 1 || !c21_host || (mode != 1 && mode != 2) || !a_dev ||
      (mode == 1 && !b_dev) || (stride != 0 && stride != (mode == 1 ? 3 : 9)) || !rows_dev || local_base < 0 ||
      !bad_host)
    return fail("vxl_local_rows: invalid arguments");
  for (int i = 0; i < span; ++i)
    if (slot_host[i] < -1 || slot_host[i] >= phases) return fail("vxl_local_rows: phase slot out of range");
  return local_rows(int64_t(nz) * ny * nx, label_bytes, labels_dev, label_lo, span, slot_host, phases, c21_host,
                    mode, a_dev, b_dev, stride, rows_dev, local_base, bad_host, static_cast<hipStream_t>(stream));
}

int vxl_cell_create(vxl_cell** out, int nz, int ny, int nx, double hx, double hy, double hz,
                    int index_kind, const void* index_dev, int rows, const double* c21_host,
                    const float* scale_dev, void* stream) {
  return cell_create_common(out, nz, ny, nx, hx, hy, hz, index_kind, index_dev, rows, scale_dev, stream,
                            [&](vxl_cell* c) -> int {
                              if (!c21_host) return fail("index, coefficient table and rows are required");
                              for (int64_t i = 0; i < int64_t(rows) * 21; ++i)
                                if (!std::isfinite(c21_host[i]))
                                  return fail("coefficient table contains nonfinite values");
                              std::vector<double> t64(size_t(rows) * kRowWidth, 0.0);
                              std::vector<float> t32(size_t(rows) * kRowWidth, 0.0f);
                              for (int r = 0; r < rows; ++r)
                                for (int k = 0; k < 21; ++k) {
                                  t64[size_t(r) * kRowWidth + k] = c21_host[size_t(r) * 21 + k];
                                  t32[size_t(r) * kRowWidth + k] = float(c21_host[size_t(r) * 21 + k]);
                                }
                              VXL_HIP(hipMemcpy(c->c64_owned, t64.data(), t64.size() * sizeof(double),
                                                hipMemcpyHostToDevice));
                              VXL_HIP(hipMemcpy(c->c32_owned, t32.data(), t32.size() * sizeof(float),
                                                hipMemcpyHostToDevice));
                              return VXL_OK;
                            });
}

int vxl_cell_create_device_rows(vxl_cell** out, int nz, int ny, int nx, double hx, double hy, double hz,
                                int index_kind, const void* index_dev, int rows, const double* c21_dev,
                                const float* scale_dev, void* stream) {
  return cell_create_common(out, nz, ny, nx, hx, hy, hz, index_kind, index_dev, rows, scale_dev, stream,
                            [&](vxl_cell* c) -> int {
                              if (!c21_dev) return fail("index, coefficient table and rows are required");
                              DevBuf flag(c->stream);
                              VXL_HIP(dev_malloc(&flag.p, sizeof(int)));
                              VXL_HIP(hipMemsetAsync(flag.p, 0, sizeof(int), c->stream));
                              const int64_t total = int64_t(rows) * kRowWidth;
                              const unsigned blocks = unsigned(std::min<int64_t>((total + 255) / 256, 65536));
                              hipLaunchKernelGGL(pad_rows_kernel, dim3(blocks), dim3(256), 0, c->stream, c21_dev,
                                                 int64_t(rows), static_cast<double*>(c->c64_owned),
                                                 static_cast<float*>(c->c32_owned), static_cast<int*>(flag.p));
                              VXL_HIP(hipGetLastError());
                              int bad = 0;
                              VXL_HIP(hipMemcpyAsync(&bad, flag.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
                              VXL_HIP(hipStreamSynchronize(c->stream));
                              if (bad) return fail("coefficient table contains nonfinite values");
                              return VXL_OK;
                            });
}

int vxl_cell_destroy(vxl_cell* cell) {
  if (!cell) return VXL_OK;
  (void)hipStreamSynchronize(cell->stream);
  mfma_release(cell);
  frames_release(cell);
  if (cell->c64_owned) dev_free(cell->c64_owned);
  if (cell->c32_owned) dev_free(cell->c32_owned);
  delete cell;
  return VXL_OK;
}

int vxl_apply(vxl_cell* cell, int precision, int nrhs, const void* u_dev, void* out_dev) {
  g_error.clear();
  if (!cell || !u_dev || !out_dev || nrhs < 1) return fail("vxl_apply: invalid arguments");
  if (nrhs % cell->groups) return fail("stacked cells: nrhs must be a multiple of the cell count");
  if (u_dev == out_dev) return fail("vxl_apply: input and output must not alias");
  if (precision == VXL_F64)
    return launch_apply<double>(cell, nrhs, static_cast<const double*>(u_dev), static_cast<double*>(out_dev),
                                cell->stream);
  if (precision == VXL_F32)
    return launch_apply<float>(cell, nrhs, static_cast<const float*>(u_dev), static_cast<float*>(out_dev),
                               cell->stream);
  return fail("precision must be VXL_F32 or VXL_F64");
}

int vxl_set_stencil_variant(vxl_cell* cell, int precision, int variant) {
  g_error.clear();
  if (!cell || (precision != VXL_F32 && precision != VXL_F64) || variant < 1 || variant > 3)
    return fail("vxl_set_stencil_variant: invalid arguments");
  cell->variant[precision == VXL_F64 ? 1 : 0] = variant;
  return VXL_OK;
}

int vxl_cell_set_groups(vxl_cell* cell, int groups) {
  g_error.clear();
  if (!cell || groups < 1 || groups > 64) return fail("vxl_cell_set_groups: 1 to 64 cells");
  if (groups > 1 && (cell->mfma || cell->frame_index || cell->mat.scale || cell->element != VXL_ELEMENT_FI))
    return fail("stacked cells take the modal FI stencil without a scale field, frames or the matrix-core form");
  cell->groups = groups;
  return VXL_OK;
}

int vxl_cell_set_mfma(vxl_cell* cell, int enable) {
  g_error.clear();
  if (!cell || (enable != 0 && enable != 1)) return fail("vxl_cell_set_mfma: invalid arguments");
  if (enable && cell->groups > 1) return fail("the matrix-core stencil takes one cell, not a stacked cell");
  if (enable && cell->frame_index) return fail("the matrix-core stencil does not take per-voxel frames");
  if (enable) {
    VXL_TRY(mfma_prepare(cell));
    cell->mfma = 1;
  } else {
    cell->mfma = 0;
  }
  return VXL_OK;
}

int vxl_cell_set_frames(vxl_cell* cell, int nphase, const double* phase_rows, const int* phase_local,
                        const void* phase_index_dev, const float* quat_dev) {
  g_error.clear();
  if (!cell) return fail("vxl_cell_set_frames: invalid arguments");
  if (!phase_index_dev) {
    (void)hipStreamSynchronize(cell->stream);
    frames_release(cell);
    return VXL_OK;
  }
  if (nphase < 1 || nphase > 256 || !phase_rows || !phase_local || !quat_dev)
    return fail("vxl_cell_set_frames: 1 to 256 phase rows, their frame flags, a phase index and quaternions");
  if (cell->element != VXL_ELEMENT_FI) return fail("per-voxel frames need the FI element");
  if (cell->groups > 1) return fail("per-voxel frames take one cell, not a stacked cell");
  if (cell->mat.scale) return fail("per-voxel frames do not combine with a scale field");
  if (cell->mfma) return fail("the matrix-core stencil does not take per-voxel frames");
  std::vector<float> table(size_t(nphase) * kRowWidth, 0.0f);
  bool any_bond = false, any_ortho = false;
  for (int p = 0; p < nphase; ++p) {
    const double* c = phase_rows + size_t(p) * 21;
    for (int k = 0; k < 21; ++k)
      if (!std::isfinite(c[k])) return fail("phase rows contain nonfinite values");
    int kind = 0;
    if (phase_local[p]) {
      // Orthotropic in its local frame: no normal-shear and no shear-shear
      // coupling (exact zeros); cubic besides: equal normal diagonal, equal
      // normal off-diagonal, equal shear diagonal (exactly).
      bool ortho = true;
      for (int i = 0; i < 6; ++i)
        for (int j = i + 1; j < 6; ++j)
          if (j >= 3 && c[pidx(i, j)] != 0.0) ortho = false;
      const double d = c[pidx(0, 0)], o = c[pidx(0, 1)], g = c[pidx(3, 3)];
      const bool cubic = ortho && c[pidx(1, 1)] == d && c[pidx(2, 2)] == d && c[pidx(0, 2)] == o &&
                         c[pidx(1, 2)] == o && c[pidx(4, 4)] == g && c[pidx(5, 5)] == g;
      kind = cubic ? 2 : ortho ? 3 : 1;
      any_bond = any_bond || kind == 1;
      any_ortho = any_ortho || kind == 3;
    }
    for (int k = 0; k < 21; ++k) table[size_t(p) * kRowWidth + k] = float(c[k]);
    table[size_t(p) * kRowWidth + 21] = float(kind);
  }
  (void)hipStreamSynchronize(cell->stream);
  frames_release(cell);
  const int64_t n = cell->grid.n;
  auto undo = [&](int rc) {
    frames_release(cell);
    return rc;
  };
  if (dev_malloc(&cell->frame_table, table.size() * sizeof(float)) != hipSuccess ||
      dev_malloc(&cell->frame_index, size_t(n)) != hipSuccess ||
      dev_malloc(&cell->frame_quat, size_t(n) * 4 * sizeof(float)) != hipSuccess)
    return undo(fail("out of device memory for the frames"));
  int* flag = nullptr;
  if (dev_malloc(&flag, 2 * sizeof(int)) != hipSuccess) return undo(fail("out of device memory for the frames"));
  int bad[2] = {0, 0};
  const int rc = [&]() -> int {
    VXL_HIP(hipMemcpyAsync(cell->frame_table, table.data(), table.size() * sizeof(float), hipMemcpyHostToDevice,
                           cell->stream));
    VXL_HIP(hipMemcpyAsync(cell->frame_index, phase_index_dev, size_t(n), hipMemcpyDeviceToDevice, cell->stream));
    VXL_HIP(hipMemcpyAsync(cell->frame_quat, quat_dev, size_t(n) * 4 * sizeof(float), hipMemcpyDeviceToDevice,
                           cell->stream));
    VXL_TRY(launch_frames_check(cell->frame_index, cell->frame_quat, n, nphase, flag, cell->stream));
    VXL_HIP(hipMemcpyAsync(bad, flag, 2 * sizeof(int), hipMemcpyDeviceToHost, cell->stream));
    VXL_HIP(hipStreamSynchronize(cell->stream));
    return VXL_OK;
  }();
  dev_free(flag);
  if (rc != VXL_OK) return undo(rc);
  if (bad[0]) return undo(fail("phase index out of range of the phase rows"));
  if (bad[1]) return undo(fail("frames must be finite unit quaternions (|q|^2 within 1e-5 of 1)"));
  cell->frame_forms = any_bond ? 1 : any_ortho ? 3 : 2;
  return VXL_OK;
}

int vxl_get_stencil_variant(vxl_cell* cell, int precision) {
  if (!cell || (precision != VXL_F32 && precision != VXL_F64)) return fail("vxl_get_stencil_variant: invalid arguments");
  return cell->variant[precision == VXL_F64 ? 1 : 0];
}

int vxl_tune_stencil(vxl_cell* cell, int precision, int nrhs, const void* u_dev, void* out_dev, int reps,
                     int* best, double* ms3) {
  g_error.clear();
  if (!cell || !u_dev || !out_dev || nrhs < 1 || reps < 1 || (precision != VXL_F32 && precision != VXL_F64))
    return fail("vxl_tune_stencil: invalid arguments");
  hipEvent_t e0, e1;
  VXL_HIP(hipEventCreate(&e0));
  VXL_HIP(hipEventCreate(&e1));
  const int slot = precision == VXL_F64 ? 1 : 0;
  double best_ms = 1e300;
  int best_v = cell->variant[slot];
  for (int v = 1; v <= 3; ++v) {
    cell->variant[slot] = v;
    int rc = vxl_apply(cell, precision, nrhs, u_dev, out_dev);  // warm-up
    if (rc != VXL_OK) return rc;
    VXL_HIP(hipEventRecord(e0, cell->stream));
    for (int i = 0; i < reps; ++i) {
      rc = vxl_apply(cell, precision, nrhs, u_dev, out_dev);
      if (rc != VXL_OK) return rc;
    }
    VXL_HIP(hipEventRecord(e1, cell->stream));
    VXL_HIP(hipEventSynchronize(e1));
    float ms = 0;
    VXL_HIP(hipEventElapsedTime(&ms, e0, e1));
    const double per = double(ms) / reps;
    if (ms3) ms3[v - 1] = per;
    if (per < best_ms) {
      best_ms = per;
      best_v = v;
    }
  }
  cell->variant[slot] = best_v;
  if (best) *best = best_v;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return VXL_OK;
}

int vxl_cell_set_element(vxl_cell* cell, int element, double hourglass_coefficient) {
  g_error.clear();
  if (!cell || (element != VXL_ELEMENT_FI && element != VXL_ELEMENT_C3D8R_HG))
    return fail("vxl_cell_set_element: invalid arguments");
  if (element == VXL_ELEMENT_C3D8R_HG && !(std::isfinite(hourglass_coefficient) && hourglass_coefficient > 0.0))
    return fail("hourglass_coefficient must be finite and positive");
  if (element == VXL_ELEMENT_C3D8R_HG && (cell->mfma || cell->frame_index || cell->groups > 1))
    return fail("the matrix-core stencil and per-voxel frames implement the FI element only; turn them off first");
  cell->element = element;
  if (element == VXL_ELEMENT_C3D8R_HG) cell->hg_coef = hourglass_coefficient;
  return VXL_OK;
}

int vxl_set_stencil_config(vxl_cell* cell, int precision, int tile_y, int variant, int zmin) {
  g_error.clear();
  if (!cell || (precision != VXL_F32 && precision != VXL_F64))
    return fail("vxl_set_stencil_config: invalid arguments");
  if (!stencil_tile_supported(precision, tile_y)) return fail("vxl_set_stencil_config: unsupported tile height");
  if (variant < 1 || variant > 3) return fail("vxl_set_stencil_config: schedule must be 1, 2 or 3");
  if (zmin < kMinChunkPlanes || zmin > 64) return fail("vxl_set_stencil_config: chunk depth must lie in [2, 64]");
  const int slot = precision == VXL_F64 ? 1 : 0;
  cell->tile_y[slot] = tile_y;
  cell->variant[slot] = variant;
  cell->zmin[slot] = zmin;
  return VXL_OK;
}

int vxl_get_stencil_config(vxl_cell* cell, int precision, int* tile_y, int* variant, int* zmin) {
  g_error.clear();
  if (!cell || (precision != VXL_F32 && precision != VXL_F64))
    return fail("vxl_get_stencil_config: invalid arguments");
  const int slot = precision == VXL_F64 ? 1 : 0;
  if (tile_y) *tile_y = cell->tile_y[slot];
  if (variant) *variant = cell->variant[slot];
  if (zmin) *zmin = cell->zmin[slot];
  return VXL_OK;
}

int vxl_autotune_stencil(vxl_cell* cell, int precision, int nrhs, const void* u_dev, void* out_dev, int reps,
                         int* tile_y, int* variant, int* zmin, double* ms_table) {
  g_error.clear();
  if (!cell || !u_dev || !out_dev || nrhs < 1 || reps < 1 || (precision != VXL_F32 && precision != VXL_F64))
    return fail("vxl_autotune_stencil: invalid arguments");
  const int slot = precision == VXL_F64 ? 1 : 0;
  const int depths[3] = {kDefaultChunkPlanes, 4, kMinChunkPlanes};
  hipEvent_t e0, e1;
  VXL_HIP(hipEventCreate(&e0));
  VXL_HIP(hipEventCreate(&e1));
  const int keep[3] = {cell->tile_y[slot], cell->variant[slot], cell->zmin[slot]};
  int best[3] = {keep[0], keep[1], keep[2]};
  double best_ms = 1e300;
  int rc = VXL_OK;
  for (int t = 0; t < 3 && rc == VXL_OK; ++t)
    for (int v = 1; v <= 3 && rc == VXL_OK; ++v)
      for (int d = 0; d < 3 && rc == VXL_OK; ++d) {
        cell->tile_y[slot] = stencil_tile_option(precision, t);
        cell->variant[slot] = v;
        cell->zmin[slot] = depths[d];
        rc = vxl_apply(cell, precision, nrhs, u_dev, out_dev);  // warm-up
        if (rc != VXL_OK) break;
        if (hipEventRecord(e0, cell->stream) != hipSuccess) rc = fail("hipEventRecord failed");
        for (int i = 0; i < reps && rc == VXL_OK; ++i) rc = vxl_apply(cell, precision, nrhs, u_dev, out_dev);
        if (rc != VXL_OK) break;
        float ms = 0;
        if (hipEventRecord(e1, cell->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
          rc = fail("HIP event timing failed in vxl_autotune_stencil");
          break;
        }
        const double per = double(ms) / reps;
        if (ms_table) ms_table[(t * 3 + (v - 1)) * 3 + d] = per;
        if (per < best_ms) {
          best_ms = per;
          best[0] = cell->tile_y[slot];
          best[1] = v;
          best[2] = depths[d];
        }
      }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  const int* use = rc == VXL_OK ? best : keep;
  cell->tile_y[slot] = use[0];
  cell->variant[slot] = use[1];
  cell->zmin[slot] = use[2];
  if (rc != VXL_OK) return rc;
  if (tile_y) *tile_y = best[0];
  if (variant) *variant = best[1];
  if (zmin) *zmin = best[2];
  return VXL_OK;
}

int vxl_rhs(vxl_cell* cell, int nload, const double* macro_host, double* f_dev) {
  g_error.clear();
  if (!cell || !macro_host || !f_dev || nload < 1) return fail("vxl_rhs: invalid arguments");
  DevBuf m(cell->stream);
  VXL_HIP(dev_malloc(&m.p, sizeof(double) * 6 * nload));
  VXL_HIP(hipMemcpyAsync(m.p, macro_host, sizeof(double) * 6 * nload, hipMemcpyHostToDevice, cell->stream));
  DevBuf sig(cell->stream);
  const int64_t ns = rhs_scratch_size(cell, nload);
  if (ns > 0) VXL_HIP(dev_malloc(&sig.p, sizeof(double) * size_t(ns)));
  VXL_TRY(launch_rhs(cell, nload, static_cast<double*>(m.p), f_dev, static_cast<double*>(sig.p), cell->stream));
  VXL_HIP(hipStreamSynchronize(cell->stream));
  return VXL_OK;
}

int vxl_avg_stress(vxl_cell* cell, int nload, const double* u_dev, const double* macro_host,
                   double* out_host) {
  g_error.clear();
  if (!cell || !macro_host || !u_dev || !out_host || nload < 1) return fail("vxl_avg_stress: invalid arguments");
  DevBuf m(cell->stream), partial(cell->stream), out(cell->stream);
  VXL_HIP(dev_malloc(&m.p, sizeof(double) * 6 * nload));
  VXL_HIP(dev_malloc(&partial.p, sizeof(double) * size_t(stress_partial_size(cell, nload))));
  VXL_HIP(dev_malloc(&out.p, sizeof(double) * 6 * nload));
  VXL_HIP(hipMemcpyAsync(m.p, macro_host, sizeof(double) * 6 * nload, hipMemcpyHostToDevice, cell->stream));
  VXL_TRY(launch_avg_stress(cell, nload, u_dev, static_cast<double*>(m.p), static_cast<double*>(partial.p),
                            static_cast<double*>(out.p), cell->stream));
  VXL_HIP(hipMemcpyAsync(out_host, out.p, sizeof(double) * 6 * nload, hipMemcpyDeviceToHost, cell->stream));
  VXL_HIP(hipStreamSynchronize(cell->stream));
  return VXL_OK;
}

int vxl_center_fields(vxl_cell* cell, int nload, const double* u_dev, const double* macro_host,
                      double* strain_dev, double* stress_dev, double* energy_dev) {
  g_error.clear();
  if (!cell || !macro_host || !u_dev || nload < 1) return fail("vxl_center_fields: invalid arguments");
  DevBuf m(cell->stream);
  VXL_HIP(dev_malloc(&m.p, sizeof(double) * 6 * nload));
  VXL_HIP(hipMemcpyAsync(m.p, macro_host, sizeof(double) * 6 * nload, hipMemcpyHostToDevice, cell->stream));
  VXL_TRY(launch_center_fields(cell, nload, u_dev, static_cast<double*>(m.p), strain_dev, stress_dev, energy_dev,
                               cell->stream));
  VXL_HIP(hipStreamSynchronize(cell->stream));
  return VXL_OK;
}

int vxl_energy_gradient(vxl_cell* cell, int nload, const double* u_dev, const double* macro_host,
                        const double* w_host, const double* d21_dev, double* out_dev) {
  g_error.clear();
  if (!cell || nload < 1 || nload > 64 || !u_dev || !macro_host || !w_host || !out_dev)
    return fail("vxl_energy_gradient: invalid arguments");
  if (cell->element != VXL_ELEMENT_FI) return fail("vxl_energy_gradient: the contraction is for the FI element");
  if (cell->groups > 1) return fail("vxl_energy_gradient takes one cell, not a stacked cell");
  for (int i = 0; i < nload; ++i)
    for (int j = 0; j < nload; ++j)
      if (!std::isfinite(w_host[i * nload + j]) || w_host[i * nload + j] != w_host[j * nload + i])
        return fail("vxl_energy_gradient: weights must be a finite symmetric nload x nload matrix");
  return energy_gradient(cell, nload, u_dev, macro_host, w_host, d21_dev, out_dev, cell->stream);
}

int vxl_project_zero_mean(vxl_cell* cell, int nrhs, double* u_dev) {
  g_error.clear();
  if (!cell || !u_dev || nrhs < 1) return fail("vxl_project_zero_mean: invalid arguments");
  DevBuf partial(cell->stream);
  const int64_t n = cell->grid.n;
  VXL_HIP(dev_malloc(&partial.p, sizeof(double) * 3 * nrhs * size_t(vec_blocks(n) + 1)));
  VXL_TRY(launch_project_zero_mean(u_dev, n, nrhs, static_cast<double*>(partial.p), nullptr, cell->stream));
  VXL_HIP(hipStreamSynchronize(cell->stream));
  return VXL_OK;
}

int vxl_project_zero_mean_noalloc(vxl_cell* cell, int nrhs, double* u_dev) {
  g_error.clear();
  if (!cell || !u_dev || nrhs < 1) return fail("vxl_project_zero_mean_noalloc: invalid arguments");
  if (nrhs > kNoAllocMaxRhs)
    return fail("vxl_project_zero_mean_noalloc: at most 16 fields per call (the static scratch holds the "
                "partial sums of 48 component rows)");
  double* ptr = nullptr;
  VXL_TRY(static_scratch(cell->device, &ptr));
  // 3 * nrhs * (vec_blocks(n) + 1) doubles, within the array for nrhs <= 16.
  VXL_TRY(launch_project_zero_mean(u_dev, cell->grid.n, nrhs, ptr, nullptr, cell->stream));
  return VXL_OK;
}

int vxl_precond_create(vxl_precond** out, vxl_cell* cell, int precision, double lam, double mu,
                       int nrhs_max) {
  g_error.clear();
  if (!out || !cell || nrhs_max < 1) return fail("vxl_precond_create: invalid arguments");
  if (precision != VXL_F32 && precision != VXL_F64 && precision != VXL_F32_BF16)
    return fail("precision must be VXL_F32, VXL_F64 or VXL_F32_BF16");
  *out = nullptr;
  auto* pc = new vxl_precond();
  pc->cell = cell;
  pc->precision = precision;
  pc->nrhs_max = nrhs_max;
  pc->lam = lam;
  pc->mu = mu;
  pc->impl = nullptr;
  int rc = precond_create(pc);
  if (rc != VXL_OK) {
    precond_release(pc);
    delete pc;
    return rc;
  }
  *out = pc;
  return VXL_OK;
}

int vxl_precond_apply(vxl_precond* pc, int nrhs, const void* r_dev, void* z_dev) {
  g_error.clear();
  if (!pc || !r_dev || !z_dev) return fail("vxl_precond_apply: invalid arguments");
  return precond_apply(pc, nrhs, r_dev, z_dev, pc->cell->stream);
}

int vxl_precond_destroy(vxl_precond* pc) {
  if (!pc) return VXL_OK;
  precond_release(pc);
  delete pc;
  return VXL_OK;
}

int vxl_solver_create(vxl_solver** out, vxl_cell* cell, int mode, int nrhs, double lam, double mu) {
  g_error.clear();
  if (!out || !cell || nrhs < 1 || nrhs > 64) return fail("vxl_solver_create: invalid arguments");
  if (mode != VXL_SOLVE_FP64 && mode != VXL_SOLVE_MIXED) return fail("mode must be VXL_SOLVE_FP64 or VXL_SOLVE_MIXED");
  if (nrhs % cell->groups) return fail("stacked cells: nrhs must be a multiple of the cell count");
  *out = nullptr;
  auto* s = new vxl_solver();  // value-initialised: pointers null, counters zero
  s->cell = cell;
  s->mode = mode;
  s->R = nrhs;
  s->pc.lam = lam;
  s->pc.mu = mu;
  s->element = cell->element;
  int rc = solver_create(s);
  if (rc != VXL_OK) {
    std::string msg = g_error;
    solver_release(s);
    delete s;
    g_error = msg;
    return rc;
  }
  *out = s;
  return VXL_OK;
}

int vxl_solver_run(vxl_solver* s, const double* macro_host, double tol, int max_iter, double* ch_host,
                   int* iterations, int* restarts, double* residuals, double* u_dev_out) {
  g_error.clear();
  if (!s || !macro_host || !ch_host) return fail("vxl_solver_run: invalid arguments");
  if (!(tol > 0 && tol < 1)) return fail("tol must lie in (0, 1)");
  if (max_iter < 1) return fail("max_iter must be >= 1");
  if (s->element != s->cell->element)
    return fail("the cell's element changed after this solver was created; create a new solver");
  return solver_run(s, macro_host, tol, max_iter, ch_host, iterations, restarts, residuals, u_dev_out);
}

int vxl_solver_run_batch(vxl_solver* const* solvers, int count, const double* macro_host, double tol, int max_iter,
                         double* ch_host, int* iterations, int* restarts, double* residuals, int workers) {
  g_error.clear();
  if (!solvers || count < 1 || !macro_host || !ch_host || workers < 1)
    return fail("vxl_solver_run_batch: invalid arguments");
  if (!(tol > 0 && tol < 1)) return fail("tol must lie in (0, 1)");
  if (max_iter < 1) return fail("max_iter must be >= 1");
  std::vector<int64_t> first(count + 1, 0);  // first load of each solver in the packed arrays
  for (int i = 0; i < count; ++i) {
    if (!solvers[i]) return fail("vxl_solver_run_batch: null solver");
    if (solvers[i]->element != solvers[i]->cell->element)
      return fail("the cell's element changed after a solver was created; create a new solver");
    first[i + 1] = first[i] + solvers[i]->R;
  }
  // Largest cells first (LPT): a worker that takes a small cell late finishes soon.
  std::vector<int> order(count);
  for (int i = 0; i < count; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return solvers[a]->cell->grid.n * solvers[a]->R > solvers[b]->cell->grid.n * solvers[b]->R;
  });
  // Graphs are captured here, before any worker runs: a capture in progress
  // makes other threads' synchronising HIP calls fail.
  for (int i = 0; i < count; ++i) {
    (void)hipSetDevice(solvers[i]->cell->device);
    VXL_TRY(solver_prepare_graph(solvers[i], tol));
  }
  std::atomic<int> next{0};
  std::vector<int> rc(count, VXL_OK);
  std::vector<std::string> err(count);
  auto worker = [&]() {
    for (;;) {
      const int k = next.fetch_add(1);
      if (k >= count) return;
      const int i = order[k];
      vxl_solver* sv = solvers[i];
      (void)hipSetDevice(sv->cell->device);
      const int64_t o = first[i];
      rc[i] = solver_run(sv, macro_host + 6 * o, tol, max_iter, ch_host + 6 * o, iterations ? iterations + o : nullptr,
                         restarts ? restarts + o : nullptr, residuals ? residuals + o : nullptr, nullptr);
      if (rc[i] != VXL_OK) err[i] = vxl_last_error();
    }
  };
  const int nt = std::min(workers, count);
  std::vector<std::thread> pool;
  for (int t = 1; t < nt; ++t) pool.emplace_back(worker);
  worker();
  for (auto& t : pool) t.join();
  for (int i = 0; i < count; ++i)
    if (rc[i] != VXL_OK) {
      g_error = "cell " + std::to_string(i) + ": " + err[i];
      return rc[i];
    }
  return VXL_OK;
}

int vxl_cell_inverse_diagonal(vxl_cell* cell, double* out_dev) {
  g_error.clear();
  if (!cell || !out_dev) return fail("vxl_cell_inverse_diagonal: invalid arguments");
  if (cell->groups > 1) return fail("vxl_cell_inverse_diagonal takes one cell, not a stacked cell");
  VXL_TRY(launch_jacobi_diag(cell, out_dev, nullptr, cell->stream));
  VXL_HIP(hipStreamSynchronize(cell->stream));
  return VXL_OK;
}

int vxl_solver_set_preconditioner(vxl_solver* s, int kind) {
  g_error.clear();
  if (!s || (kind != VXL_PRECOND_FFT && kind != VXL_PRECOND_JACOBI && kind != VXL_PRECOND_FFT_BF16))
    return fail("vxl_solver_set_preconditioner: invalid arguments");
  if (kind == VXL_PRECOND_FFT_BF16 && s->mode != VXL_SOLVE_MIXED)
    return fail("the BF16-stored preconditioner is a mixed-mode option");
  if (kind == VXL_PRECOND_JACOBI && s->cell->groups > 1)
    return fail("stacked cells run with the FFT preconditioner");
  return solver_set_preconditioner(s, kind);
}

int vxl_solver_set_method(vxl_solver* s, int method) {
  g_error.clear();
  if (!s || (method != VXL_METHOD_CG && method != VXL_METHOD_BLOCK_CG))
    return fail("vxl_solver_set_method: invalid arguments");
  if (method == VXL_METHOD_BLOCK_CG && s->cell->groups > 1)
    return fail("block CG couples the load cases of one cell; stacked cells run per-case CG");
  return solver_set_method(s, method);
}

int vxl_solver_set_media(vxl_solver* s, const double* lam, const double* mu) {
  g_error.clear();
  if (!s || !lam || !mu) return fail("vxl_solver_set_media: invalid arguments");
  if (s->cell->groups < 2) return fail("vxl_solver_set_media: the cell is not a stacked cell");
  return solver_set_media(s, lam, mu);
}

int vxl_solver_graph_blocks(vxl_solver* s) {
  g_error.clear();
  if (!s) return fail("vxl_solver_graph_blocks: invalid arguments");
  return s->graph_blocks;
}

int vxl_solver_timings(vxl_solver* s, double* ms3) {
  if (!s || !ms3) return fail("vxl_solver_timings: invalid arguments");
  for (int i = 0; i < 3; ++i) ms3[i] = s->ms[i];
  return VXL_OK;
}

int vxl_solver_trace(vxl_solver* s, int k, int* iterations, double* residuals, int capacity) {
  g_error.clear();
  if (!s || k < 0 || k >= s->R || capacity < 0) return fail("vxl_solver_trace: invalid arguments");
  const auto& t = s->trace[k];
  const int n = int(t.size());
  for (int i = 0; i < n && i < capacity; ++i) {
    if (iterations) iterations[i] = t[i].first;
    if (residuals) residuals[i] = t[i].second;
  }
  return n;
}

int vxl_solver_destroy(vxl_solver* s) {
  if (!s) return VXL_OK;
  solver_release(s);
  delete s;
  return VXL_OK;
}

}  // extern "C"
