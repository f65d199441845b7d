// Block conjugate gradients over the load cases (O'Leary's block PCG): the
// vector kernels of one block iteration.  The systems share one Krylov space:
// the step and direction updates mix all R search directions at every DOF
// through small R x R matrices (alpha, beta), formed on the host from Gram
// matrices these kernels reduce.  Fields are component planes
// field[r][c][node] (r < R), len = 3N entries per system.
#include "vxl_cg.h"

namespace vxl {
namespace {

constexpr int kBT = 256;       // threads per workgroup
constexpr int kMaxR = 8;       // largest block handled here
constexpr int kGramWG = 1024;  // workgroups of a Gram reduction

// Sum of v over the workgroup, valid in thread 0.
__device__ __forceinline__ double wg_sum(double v, double* red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0;
  if (threadIdx.x == 0)
#pragma unroll
    for (int w = 0; w < kBT / 64; ++w) s += red[w];
  __syncthreads();
  return s;
}

// partial[(i * R + j) * gridDim.x + b] = sum over this workgroup's entries of
// A_i . B_j (FP64 products and sums).  Each thread loads the R values of A
// and of B at its entries once and keeps all R x R sums, so A and B are read
// from memory once per launch.
template <typename T, int NR>
__global__ void __launch_bounds__(kBT)
gram_kernel(const T* __restrict__ A, const T* __restrict__ B, int64_t len, double* __restrict__ partial) {
  __shared__ double red[kBT / 64];
  double acc[NR][NR];
#pragma unroll
  for (int i = 0; i < NR; ++i)
#pragma unroll
    for (int j = 0; j < NR; ++j) acc[i][j] = 0.0;
  for (int64_t e = int64_t(blockIdx.x) * kBT + threadIdx.x; e < len; e += int64_t(gridDim.x) * kBT) {
    double a[NR], b[NR];
#pragma unroll
    for (int i = 0; i < NR; ++i) {
      a[i] = double(A[int64_t(i) * len + e]);
      b[i] = double(B[int64_t(i) * len + e]);
    }
#pragma unroll
    for (int i = 0; i < NR; ++i)
#pragma unroll
      for (int j = 0; j < NR; ++j) acc[i][j] = fma(a[i], b[j], acc[i][j]);
  }
#pragma unroll
  for (int i = 0; i < NR; ++i)
#pragma unroll
    for (int j = 0; j < NR; ++j) {
      const double s = wg_sum(acc[i][j], red);
      if (threadIdx.x == 0) partial[(int64_t(i) * NR + j) * gridDim.x + blockIdx.x] = s;
    }
}

// X_j += sum_i P_i alpha[i][j] (X may be null), R_j -= sum_i Q_i alpha[i][j]
// for the systems with active[j] != 0, and FP64 ||R_j||^2 partials
// (rr[j * gridDim.x + b]) of every system.
template <typename T>
__global__ void __launch_bounds__(kBT)
block_update_kernel(T* __restrict__ X, T* __restrict__ Rv, const T* __restrict__ P, const T* __restrict__ Q,
                    const double* __restrict__ alpha, const int* __restrict__ active, int64_t len, int R,
                    double* __restrict__ rr) {
  __shared__ double red[kBT / 64];
  __shared__ T al[kMaxR * kMaxR];
  __shared__ int act[kMaxR];
  if (threadIdx.x < R * R) al[threadIdx.x] = T(alpha[threadIdx.x]);
  if (threadIdx.x < R) act[threadIdx.x] = active[threadIdx.x];
  __syncthreads();
  double acc[kMaxR];
#pragma unroll
  for (int j = 0; j < kMaxR; ++j) acc[j] = 0.0;
  for (int64_t e = int64_t(blockIdx.x) * kBT + threadIdx.x; e < len; e += int64_t(gridDim.x) * kBT) {
    T p[kMaxR], q[kMaxR];
#pragma unroll
    for (int i = 0; i < kMaxR; ++i) {
      p[i] = i < R ? P[int64_t(i) * len + e] : T(0);
      q[i] = i < R ? Q[int64_t(i) * len + e] : T(0);
    }
#pragma unroll
    for (int j = 0; j < kMaxR; ++j) {
      if (j >= R) break;
      T r = Rv[int64_t(j) * len + e];
      if (act[j]) {
        T dx = T(0), dr = T(0);
#pragma unroll
        for (int i = 0; i < kMaxR; ++i) {
          if (i >= R) break;
          dx = fma(p[i], al[i * R + j], dx);
          dr = fma(q[i], al[i * R + j], dr);
        }
        r -= dr;
        Rv[int64_t(j) * len + e] = r;
        if (X) X[int64_t(j) * len + e] += dx;
      }
      acc[j] = fma(double(r), double(r), acc[j]);
    }
  }
#pragma unroll
  for (int j = 0; j < kMaxR; ++j) {
    if (j >= R) break;
    const double s = wg_sum(acc[j], red);
    if (threadIdx.x == 0) rr[int64_t(j) * gridDim.x + blockIdx.x] = s;
  }
}

// P_j = Z_j + sum_i P_i beta[i][j] for the directions with keep[j] != 0,
// P_j = 0 for the others (in place: every thread reads all P_i of its entry
// before it writes).
template <typename T>
__global__ void __launch_bounds__(kBT)
block_direction_kernel(T* __restrict__ P, const T* __restrict__ Z, const double* __restrict__ beta,
                       const int* __restrict__ keep, int64_t len, int R) {
  __shared__ T be[kMaxR * kMaxR];
  __shared__ int kp[kMaxR];
  if (threadIdx.x < R * R) be[threadIdx.x] = T(beta[threadIdx.x]);
  if (threadIdx.x < R) kp[threadIdx.x] = keep[threadIdx.x];
  __syncthreads();
  for (int64_t e = int64_t(blockIdx.x) * kBT + threadIdx.x; e < len; e += int64_t(gridDim.x) * kBT) {
    T p[kMaxR];
#pragma unroll
    for (int i = 0; i < kMaxR; ++i) p[i] = i < R ? P[int64_t(i) * len + e] : T(0);
#pragma unroll
    for (int j = 0; j < kMaxR; ++j) {
      if (j >= R) break;
      T v = T(0);
      if (kp[j]) {
        v = Z[int64_t(j) * len + e];
#pragma unroll
        for (int i = 0; i < kMaxR; ++i) {
          if (i >= R) break;
          v = fma(p[i], be[i * R + j], v);
        }
      }
      P[int64_t(j) * len + e] = v;
    }
  }
}

int grid_blocks(int64_t len) {
  const int64_t b = (len + kBT - 1) / kBT;
  return int(b > kGramWG ? kGramWG : (b < 1 ? 1 : b));
}

}  // namespace

int block_max_rhs() { return kMaxR; }
int block_gram_blocks(int64_t len) { return grid_blocks(len); }

template <typename T>
int launch_block_gram(const T* A, const T* B, int64_t len, int R, double* partial, double* out, hipStream_t s) {
  if (R < 1 || R > kMaxR) return fail("block CG: at most 8 systems");
  const int nb = grid_blocks(len);
  switch (R) {
#define VXL_GRAM(NR)                                                                                 \
  case NR:                                                                                           \
    hipLaunchKernelGGL((gram_kernel<T, NR>), dim3(nb), dim3(kBT), 0, s, A, B, len, partial);         \
    break;
    VXL_GRAM(1)
    VXL_GRAM(2)
    VXL_GRAM(3)
    VXL_GRAM(4)
    VXL_GRAM(5)
    VXL_GRAM(6)
    VXL_GRAM(7)
    VXL_GRAM(8)
#undef VXL_GRAM
  }
  VXL_HIP(hipGetLastError());
  return launch_row_totals(partial, R * R, nb, 1.0, out, s);
}

template <typename T>
int launch_block_update(T* X, T* Rv, const T* P, const T* Q, const double* alpha, const int* active, int64_t len,
                        int R, double* rr_partial, hipStream_t s) {
  if (R < 1 || R > kMaxR) return fail("block CG: at most 8 systems");
  hipLaunchKernelGGL((block_update_kernel<T>), dim3(grid_blocks(len)), dim3(kBT), 0, s, X, Rv, P, Q, alpha, active,
                     len, R, rr_partial);
  VXL_HIP(hipGetLastError());
  return VXL_OK;
}

template <typename T>
int launch_block_direction(T* P, const T* Z, const double* beta, const int* keep, int64_t len, int R,
                           hipStream_t s) {
  if (R < 1 || R > kMaxR) return fail("block CG: at most 8 systems");
  hipLaunchKernelGGL((block_direction_kernel<T>), dim3(grid_blocks(len)), dim3(kBT), 0, s, P, Z, beta, keep, len,
                     R);
  VXL_HIP(hipGetLastError());
  return VXL_OK;
}

template int launch_block_gram<float>(const float*, const float*, int64_t, int, double*, double*, hipStream_t);
template int launch_block_gram<double>(const double*, const double*, int64_t, int, double*, double*, hipStream_t);
template int launch_block_update<float>(float*, float*, const float*, const float*, const double*, const int*,
                                        int64_t, int, double*, hipStream_t);
template int launch_block_update<double>(double*, double*, const double*, const double*, const double*,
                                         const int*, int64_t, int, double*, hipStream_t);
template int launch_block_direction<float>(float*, const float*, const double*, const int*, int64_t, int,
                                           hipStream_t);
template int launch_block_direction<double>(double*, const double*, const double*, const int*, int64_t, int,
                                            hipStream_t);

}  // namespace vxl
