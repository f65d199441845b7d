// Scalar Jacobi preconditioner: the control the reference compares FFT-PCG
// against (repro/bench/chapter_aligned/study.py:84-90: z = project_zero_mean(
// r / diag K) with the assembled diagonal).  diag K is assembled here from the
// 8 elements around each node with the operator's 2x2x2 Gauss rule:
//   K_(n c),(n c) = sum_e sum_g (V/8) b^T C_e b,
// b the column of B_g for node n's corner in element e and component c.  For
// the C3D8R-HG element the rule is the centroid with weight V, plus the
// diagonal of the hourglass penalty s_e H H^T: s_e * 4 patterns * (1/8).
#include "vxl_cg.h"

namespace vxl {
namespace {

constexpr int kJBlock = 256;

template <typename Idx, bool kScale>
__global__ void __launch_bounds__(128)
jacobi_diag_kernel(const void* index, const double* table, const float* scale, int nx, int ny, int nz, double hx,
                   double hy, double hz, double hgk, double* dinv64, float* dinv32) {
  const int64_t nxy = int64_t(nx) * ny, n = nxy * nz;
  const int64_t node = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (node >= n) return;
  const int x = int(node % nx), y = int((node / nx) % ny), z = int(node / nxy);
  const bool hg = hgk > 0.0;  // C3D8R-HG: centroid rule + hourglass penalty
  const double gp = hg ? 0.0 : 0.57735026918962576451;  // 1/sqrt(3)
  const int npts = hg ? 1 : 8;
  const double wv = hx * hy * hz / double(npts);
  double d[3] = {0, 0, 0};
  for (int a = 0; a < 8; ++a) {  // the node is corner a of element (x - ax, y - ay, z - az)
    const int ax = a & 1, ay = (a >> 1) & 1, az = a >> 2;
    const int ex = (x - ax + nx) % nx, ey = (y - ay + ny) % ny, ez = (z - az + nz) % nz;
    const int64_t e = int64_t(ez) * nxy + int64_t(ey) * nx + ex;
    const double* C = table + int64_t(load_row<Idx>(index, e)) * kRowWidth;
    const double s = kScale ? double(scale[e]) : 1.0;
    const double sx = ax ? 1.0 : -1.0, sy = ay ? 1.0 : -1.0, sz = az ? 1.0 : -1.0;
    if (hg) {
      const double normal = C[pidx(0, 0)] + C[pidx(1, 1)] + C[pidx(2, 2)];
      const double i1 = normal + 2 * (C[pidx(0, 1)] + C[pidx(0, 2)] + C[pidx(1, 2)]);
      const double i2 = normal + 2 * (C[pidx(3, 3)] + C[pidx(4, 4)] + C[pidx(5, 5)]);
      const double pen = hgk * fmax(s * (i1 + 7 * i2) / 60.0, 1e-12) * 0.5;  // scaled C, as the stencil
      for (int c = 0; c < 3; ++c) d[c] += pen;
    }
    for (int g = 0; g < npts; ++g) {
      const double xi = (g & 1) ? gp : -gp, eta = ((g >> 1) & 1) ? gp : -gp, zeta = (g >> 2) ? gp : -gp;
      const double fx = 1 + sx * xi, fy = 1 + sy * eta, fz = 1 + sz * zeta;
      const double gx = sx * fy * fz / (4 * hx), gy = sy * fx * fz / (4 * hy), gz = sz * fx * fy / (4 * hz);
      // b_x = (gx,0,0,0,gz,gy), b_y = (0,gy,0,gz,0,gx), b_z = (0,0,gz,gy,gx,0)
      d[0] += s * wv *
              (gx * gx * C[pidx(0, 0)] + gz * gz * C[pidx(4, 4)] + gy * gy * C[pidx(5, 5)] +
               2 * (gx * gz * C[pidx(0, 4)] + gx * gy * C[pidx(0, 5)] + gz * gy * C[pidx(4, 5)]));
      d[1] += s * wv *
              (gy * gy * C[pidx(1, 1)] + gz * gz * C[pidx(3, 3)] + gx * gx * C[pidx(5, 5)] +
               2 * (gy * gz * C[pidx(1, 3)] + gy * gx * C[pidx(1, 5)] + gz * gx * C[pidx(3, 5)]));
      d[2] += s * wv *
              (gz * gz * C[pidx(2, 2)] + gy * gy * C[pidx(3, 3)] + gx * gx * C[pidx(4, 4)] +
               2 * (gz * gy * C[pidx(2, 3)] + gz * gx * C[pidx(2, 4)] + gy * gx * C[pidx(3, 4)]));
    }
  }
  for (int c = 0; c < 3; ++c) {
    const double v = 1.0 / d[c];
    dinv64[c * n + node] = v;
    if (dinv32) dinv32[c * n + node] = float(v);
  }
}

__device__ __forceinline__ double jblock_sum(double v, double* red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0;
  if (threadIdx.x == 0)
    for (int w = 0; w < kJBlock / 64; ++w) s += red[w];
  __syncthreads();
  return s;
}

// partial[(rhs * 3 + c) * gridDim.x + block] = sum over the block's nodes of dinv r (component c).
template <typename T>
__global__ void __launch_bounds__(kJBlock)
jacobi_sums_kernel(const T* r, const T* dinv, int64_t n, double* partial, const int* active) {
  __shared__ double red[kJBlock / 64];
  const int k = blockIdx.y;
  if (active && !active[k]) return;
  const T* rk = r + int64_t(k) * 3 * n;
  for (int c = 0; c < 3; ++c) {
    double acc = 0;
    for (int64_t i = int64_t(blockIdx.x) * kJBlock + threadIdx.x; i < n; i += int64_t(gridDim.x) * kJBlock)
      acc += double(dinv[c * n + i]) * double(rk[c * n + i]);
    acc = jblock_sum(acc, red);
    if (threadIdx.x == 0) partial[(int64_t(k) * 3 + c) * gridDim.x + blockIdx.x] = acc;
  }
}

// mean[rhs * 3 + c] = sum of the partials / n, one thread per (rhs, c), fixed order.
__global__ void jacobi_means_kernel(const double* partial, int nblk, int64_t n, int nrhs, double* mean,
                                    const int* active) {
  const int t = int(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= 3 * nrhs) return;
  if (active && !active[t / 3]) return;
  double s = 0;
  for (int b = 0; b < nblk; ++b) s += partial[int64_t(t) * nblk + b];
  mean[t] = s / double(n);
}

// z = dinv r - mean (per component); partial[rhs * gridDim.x + block] = sum r . z.
template <typename T>
__global__ void __launch_bounds__(kJBlock)
jacobi_rz_kernel(const T* r, const T* dinv, const double* mean, int64_t n, double* partial, const int* active) {
  __shared__ double red[kJBlock / 64];
  const int k = blockIdx.y;
  if (active && !active[k]) return;
  const T* rk = r + int64_t(k) * 3 * n;
  double acc = 0;
  for (int c = 0; c < 3; ++c) {
    const double m = mean[k * 3 + c];
    for (int64_t i = int64_t(blockIdx.x) * kJBlock + threadIdx.x; i < n; i += int64_t(gridDim.x) * kJBlock) {
      const double rv = double(rk[c * n + i]);
      acc += rv * (double(dinv[c * n + i]) * rv - m);
    }
  }
  acc = jblock_sum(acc, red);
  if (threadIdx.x == 0) partial[int64_t(k) * gridDim.x + blockIdx.x] = acc;
}

// p = (dinv r - mean) + beta p_old, p_old from pin (pin == p: in place;
// beta = 0: p = z, the old p is not read).
template <typename T>
__global__ void __launch_bounds__(kJBlock)
jacobi_direction_kernel(T* p, const T* pin, const T* r, const T* dinv, const double* mean, const double* beta,
                        int64_t n, const int* active) {
  const int k = blockIdx.y;
  if (active && !active[k]) return;
  const T b = T(beta[k]);
  T* pk = p + int64_t(k) * 3 * n;
  const T* gk = pin + int64_t(k) * 3 * n;
  const T* rk = r + int64_t(k) * 3 * n;
  for (int c = 0; c < 3; ++c) {
    const T m = T(mean[k * 3 + c]);
    for (int64_t i = int64_t(blockIdx.x) * kJBlock + threadIdx.x; i < n; i += int64_t(gridDim.x) * kJBlock) {
      const T zv = dinv[c * n + i] * rk[c * n + i] - m;
      pk[c * n + i] = b != T(0) ? fma(b, gk[c * n + i], zv) : zv;
    }
  }
}

}  // namespace

int jacobi_blocks(int64_t n) { return int(std::max<int64_t>(1, std::min<int64_t>((n + kJBlock - 1) / kJBlock, 1024))); }

int launch_jacobi_diag(const vxl_cell* cell, double* dinv64, float* dinv32, hipStream_t s) {
  const auto& g = cell->grid;
  const unsigned blocks = unsigned((g.n + 127) / 128);
  const double hgk = cell->element == VXL_ELEMENT_C3D8R_HG ? hg_factor(cell) : 0.0;
#define VXL_JD(I, SC)                                                                                         \
  hipLaunchKernelGGL((jacobi_diag_kernel<I, SC>), dim3(blocks), dim3(128), 0, s, cell->mat.index, cell->mat.c64, \
                     cell->mat.scale, g.nx, g.ny, g.nz, cell->h[0], cell->h[1], cell->h[2], hgk, dinv64, dinv32)
  const bool sc = cell->mat.scale != nullptr;
  switch (cell->mat.index_kind) {
    case VXL_INDEX_U8:
      if (sc) VXL_JD(uint8_t, true); else VXL_JD(uint8_t, false);
      break;
    case VXL_INDEX_U16:
      if (sc) VXL_JD(uint16_t, true); else VXL_JD(uint16_t, false);
      break;
    default:
      if (sc) VXL_JD(int32_t, true); else VXL_JD(int32_t, false);
      break;
  }
#undef VXL_JD
  VXL_HIP(hipGetLastError());
  return VXL_OK;
}

template <typename T>
int jacobi_cg_step(const T* r, T* p, const T* dinv, int64_t n, int nrhs, double* partial, double* mean,
                   double* rz_partial, double* scal, int* active, int* restart, hipStream_t s, const T* p_old) {
  const int nb = jacobi_blocks(n);
  const dim3 grid = dim3(unsigned(nb), unsigned(nrhs));
  hipLaunchKernelGGL((jacobi_sums_kernel<T>), grid, dim3(kJBlock), 0, s, r, dinv, n, partial, active);
  VXL_HIP(hipGetLastError());
  hipLaunchKernelGGL(jacobi_means_kernel, dim3(unsigned((3 * nrhs + 63) / 64)), dim3(64), 0, s, partial, nb, n, nrhs,
                     mean, active);
  VXL_HIP(hipGetLastError());
  hipLaunchKernelGGL((jacobi_rz_kernel<T>), grid, dim3(kJBlock), 0, s, r, dinv, mean, n, rz_partial, active);
  VXL_HIP(hipGetLastError());
  VXL_TRY(launch_finalize(rz_partial, nb, kOpBeta, scal, nrhs, active, restart, s));
  hipLaunchKernelGGL((jacobi_direction_kernel<T>), grid, dim3(kJBlock), 0, s, p, p_old ? p_old : p, r, dinv, mean,
                     scal + kBeta * nrhs, n, active);
  VXL_HIP(hipGetLastError());
  return VXL_OK;
}
template int jacobi_cg_step<float>(const float*, float*, const float*, int64_t, int, double*, double*, double*,
                                   double*, int*, int*, hipStream_t, const float*);
template int jacobi_cg_step<double>(const double*, double*, const double*, int64_t, int, double*, double*, double*,
                                    double*, int*, int*, hipStream_t, const double*);

}  // namespace vxl
