// Internal declarations shared by the libvxl translation units (gfx950 only).
//
// Field layout in HBM (every precision):  field[r][c][z][y][x]
//   r = right-hand side (load case), c = displacement component x,y,z,
//   node id (z*ny + y)*nx + x, periodic in all three axes.
// Material layout: one row index per voxel (uint8/uint16/int32) into a
// coefficient table of 24-wide padded rows holding the 21 upper-triangle
// entries of the engineering-Voigt 6x6 stiffness (pairs i<=j row major,
// 11,22,33,23,13,12 order).  An optional FP32 per-voxel factor multiplies the
// row (continuous density fields).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/vxl.h"

namespace vxl {

constexpr int kRowWidth = 24;  // padded C21 row

// Position of C(i, j) in a 21-coefficient upper-triangle row (i <= j, row major).
__host__ __device__ constexpr int pidx(int i, int j) {
  return i <= j ? i * 6 - (i * (i - 1)) / 2 + (j - i) : j * 6 - (j * (j - 1)) / 2 + (i - j);
}

// XCD-aware block order.  Workgroups are dealt to the 8 XCDs round-robin in
// linear block order and each XCD has its own L2; this remap gives every XCD
// a contiguous run of logical blocks (x fastest), so neighbouring blocks whose
// rows share cache lines run on the same XCD at about the same time.  A
// bijection on the grid (the last total % 8 blocks keep their ids).
__device__ __forceinline__ uint3 xcd_block3() {
  const unsigned gx = gridDim.x, gy = gridDim.y;
  const unsigned total = gx * gy * gridDim.z;
  unsigned lin = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
  const unsigned full = total & ~7u;
  if (lin < full) lin = (lin & 7u) * (full >> 3) + (lin >> 3);
  return make_uint3(lin % gx, (lin / gx) % gy, lin / (gx * gy));
}

template <typename Idx>
__device__ __forceinline__ int load_row(const void* index, int64_t v) {
  return int(static_cast<const Idx*>(index)[v]);
}

void set_error(const std::string& message);
int fail(const std::string& message);  // sets the error text, returns VXL_ERR_*

#define VXL_HIP(...)                                                            \
  do {                                                                          \
    hipError_t _e = (__VA_ARGS__);                                              \
    if (_e != hipSuccess)                                                       \
      return ::vxl::fail(std::string("HIP error in ") + #__VA_ARGS__ + ": " +   \
                         hipGetErrorString(_e));                                \
  } while (0)

#define VXL_TRY(...)                                                            \
  do {                                                                          \
    int _rc = (__VA_ARGS__);                                                    \
    if (_rc != VXL_OK) return _rc;                                              \
  } while (0)

// Geometric factors of the modal FI element action.  For strain class k
// (0: constant, 1: linear, 2: bilinear) and axis a, g[k][a] = w_k^(1/2) *
// V^(1/2) / (4 h_a) with w = {1, 1/3, 1/9}: the Gauss weights and the element
// volume are split evenly between the strain and the back-projection so the
// element matrix stays exactly symmetric.
struct ModalFactors {
  double g[3][3];
};
ModalFactors modal_factors(const double h[3]);

// Sum of v over a workgroup of NT threads (NT a multiple of 64); the result
// is valid in thread 0.  red must hold NT/64 doubles of LDS.  Every thread of
// the workgroup must call it.
template <int NT>
__device__ __forceinline__ double block_reduce_sum(double v, double* red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int w = 0; w < NT / 64; ++w) s += red[w];
  }
  __syncthreads();
  return s;
}

// Cached device allocations (devmem.hip): freed blocks are kept per
// (device, size) and reused; the owner synchronises the stream that used a
// block before dev_free.
hipError_t dev_malloc(void** p, size_t bytes);
template <typename T>
hipError_t dev_malloc(T** p, size_t bytes) {
  return dev_malloc(reinterpret_cast<void**>(p), bytes);
}
void dev_free(void* p);
size_t cached_bytes();
void release_cached();

// Sum of v over a workgroup whose size is a multiple of 64 (at most 1024);
// red holds blockDim.x / 64 doubles.  Result valid in thread 0.
__device__ __forceinline__ double block_reduce_dyn(double v, double* red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0;
  if (threadIdx.x == 0)
    for (int w = 0; w < int(blockDim.x >> 6); ++w) s += red[w];
  __syncthreads();
  return s;
}

struct Grid {
  int nx, ny, nz;
  int64_t n;
};

// Material view used by the kernels (device pointers).
struct MaterialView {
  int index_kind;           // VXL_INDEX_U8 / U16 / I32
  const void* index;        // per-voxel row
  const double* c64;        // rows x 24
  const float* c32;         // rows x 24
  const float* scale;       // optional per-voxel factor (nullptr: 1)
};

}  // namespace vxl

// Opaque handle types (complete definitions; the C header only forward-declares).
struct vxl_cell {
  vxl::Grid grid;
  double h[3];
  double volume;
  int rows;
  vxl::MaterialView mat;
  double* c64_owned;
  float* c32_owned;
  hipStream_t stream;
  int device;
  int variant[2];  // stencil schedule per precision (VXL_F32, VXL_F64): 1, 2 or 3
  int tile_y[2];   // stencil tile height (output nodes along y; the tile is 32 wide)
  int zmin[2];     // fewest z planes per workgroup chunk (each chunk recomputes one layer)
  int element;     // VXL_ELEMENT_FI (2x2x2 Gauss) or VXL_ELEMENT_C3D8R_HG (one point + hourglass)
  double hg_coef;  // hourglass coefficient of the C3D8R-HG element
  // Matrix-core FP32 stencil (mfma.hip): on when mfma != 0; the dense FP16
  // hi/lo element matrices of every row in lane order and their exponents.
  int mfma;
  void* mfma_table;
  int* mfma_exp;
  // Per-voxel material frames for the FP32 stencil (vxl_cell_set_frames): on
  // when frame_index != nullptr.  Phase slot per voxel, FP32 phase rows x 24
  // (slot 21: 0 global frame, 1 local frame, 2 local-frame cubic) and the
  // unit quaternion (w, x, y, z) of every voxel.
  uint8_t* frame_index;
  float* frame_table;
  float* frame_quat;
  int frame_forms;  // rotations the cell needs: 1 any (Bond), 2 cubic only, 3 cubic and orthotropic only
  // Stacked cells (vxl_cell_set_groups): `groups` cells of this grid and
  // voxel size whose index blocks follow each other in mat.index (n entries
  // each, into one shared row table); right-hand side r of an nrhs launch
  // belongs to cell r / (nrhs / groups).
  int groups;
};

namespace vxl {

// ---- stencil.hip -------------------------------------------------------------
template <typename T>
int launch_apply(const vxl_cell* cell, int nrhs, const T* u, T* out, hipStream_t s);
// CG stencil step: q = K p and partial[r * blocks + b] = per-workgroup sums of
// p . q (FP64); systems with active[r] == 0 are skipped.  Returns the
// workgroups per RHS in *blocks.
template <typename T>
int launch_apply_cg(const vxl_cell* cell, int nrhs, const T* p, T* q, const int* active, double* partial,
                    int* blocks, hipStream_t s);
template <typename T>
int apply_cg_blocks(const vxl_cell* cell, int nrhs);
constexpr int kDefaultChunkPlanes = 8;  // default fewest z planes per stencil chunk
constexpr int kMinChunkPlanes = 2;      // smallest chunk depth a cell may be set to
int apply_cg_blocks_max(const vxl_cell* cell, int nrhs, int precision);
bool stencil_tile_supported(int precision, int tile_y);
// Certification residual r = f - K u with the FP64 operator (r32 or r64
// written; ||r||^2 partials, *blocks per system); systems with active == 0
// are skipped.
int launch_residual_apply(const vxl_cell* cell, int nrhs, const double* u, const double* f, float* r32, double* r64,
                          const int* active, double* partial, int* blocks, hipStream_t s);
int residual_apply_blocks(const vxl_cell* cell, int nrhs);
int default_stencil_tile(int precision);
int stencil_tile_option(int precision, int k);  // k = 0, 1, 2

// C3D8R-HG hourglass penalty per element: s(C) = coef * max((I1 + 7 I2) / 60,
// 1e-12) * V / hbar^2 with I1 = sum C_ii + 2 sum_{i<j<3} C_ij, I2 = sum_{i<3}
// C_ii + 2 sum_{i>=3} C_ii, hbar the mean voxel edge
// (femlib/c3d8r.py:48-59).  hg_factor = coef * V / hbar^2; the per-element
// material part is applied where C is known.
inline double hg_factor(const vxl_cell* cell) {
  const double hbar = (cell->h[0] + cell->h[1] + cell->h[2]) / 3.0;
  return cell->hg_coef * cell->volume / (hbar * hbar);
}
int launch_rhs(const vxl_cell* cell, int nload, const double* macro_dev, double* f, double* sig, hipStream_t s);
// Doubles of the `sig` scratch of launch_rhs (rows x nload x 6 element
// stresses); 0 means pass sig = nullptr (direct per-element form).
int64_t rhs_scratch_size(const vxl_cell* cell, int nload);
// partial: workspace of stress_partial_size(cell, nload) doubles; out: nload x 6
int launch_avg_stress(const vxl_cell* cell, int nload, const double* u, const double* macro_dev,
                      double* partial, double* out_dev, hipStream_t s);
int64_t stress_partial_size(const vxl_cell* cell, int nload);
int default_stencil_variant(int precision);
// out[e] = sum_ij W_ij int_e (E_i + B u_i)^T D_e (E_j + B u_j) dV with D_e the
// cell's coefficient row of e, or d21[e] when d21 != nullptr (adjoint.hip).
int energy_gradient(const vxl_cell* cell, int nload, const double* u, const double* macro_host,
                    const double* w_host, const double* d21, double* out, hipStream_t s);
// Label -> row index on the device with label presence and the Zeman lattice
// samples (material.hip).
int label_rows(int nz, int ny, int nx, int label_bytes, const void* labels_dev, long long label_lo, int span,
               const int* lut_host, int index_kind, void* index_dev, int* present_host, int* undefined_host,
               int* samples_host, int64_t local_base, hipStream_t s);
// Bond-rotated rows of local-frame voxels at rows_dev[(local_base + e) * 21] (material.hip).
int local_rows(int64_t n, int label_bytes, const void* labels_dev, long long label_lo, int span, const int* slot_host,
               int phases, const double* c21_host, int mode, const double* a_dev, const double* b_dev, int64_t stride,
               double* rows_dev, int64_t local_base, int* bad_host, hipStream_t s);
// Matrix-core FP32 stencil (mfma.hip): label-grouped dense element matrices
// on v_mfma_f32_16x16x32_f16.  mfma_prepare builds the per-row tables once;
// launch_mfma_apply is the plain (partial == nullptr) or CG form of the
// FP32 apply, as launch_apply / launch_apply_cg.
int mfma_max_rows();
int mfma_prepare(vxl_cell* cell);
void mfma_release(vxl_cell* cell);
// Per-voxel frames (stencil.hip): frees the cell's frame buffers (frames off);
// flag[0] = 1 if a phase index is >= nphase, flag[1] = 1 if a quaternion is
// not finite or |q|^2 is off 1 by more than 1e-5 (flag cleared here).
void frames_release(vxl_cell* cell);
vxl_cell group_view(const vxl_cell* cell, int g);  // one cell of a stacked cell (stencil.hip)
int launch_frames_check(const uint8_t* index, const float* quat, int64_t n, int nphase, int* flag, hipStream_t s);
int mfma_blocks(const vxl_cell* cell, int nrhs);
int launch_mfma_apply(const vxl_cell* cell, int nrhs, const float* u, float* out, const int* active,
                      double* partial, int* blocks, hipStream_t s);
// Element-centre fields (any output may be null): strain = E + B0 u_e and
// stress = C_e strain as planes [r][6][N], energy = strain . stress / 2 as [r][N].
int launch_center_fields(const vxl_cell* cell, int nload, const double* u, const double* macro_dev,
                         double* strain, double* stress, double* energy, hipStream_t s);

}  // namespace vxl
