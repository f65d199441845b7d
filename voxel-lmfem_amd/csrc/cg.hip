// CG / iterative-refinement vector kernels, gfx950.
//
// Vectors are nrhs consecutive fields of `len` entries (len = 3N); blockIdx.y
// selects the right-hand side and a per-RHS `active` flag lets converged
// systems skip all work.  Dot products accumulate in FP64 and reduce in two
// deterministic stages: per-workgroup partials (wavefront shuffles, then LDS),
// then one workgroup per RHS sums its partials in a fixed order and applies a
// scalar update (alpha, beta, ...) in the same launch.
#include "vxl_cg.h"

namespace vxl {
namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Block sum of v; result valid in thread 0.
__device__ __forceinline__ double block_sum(double v) {
  __shared__ double red[kBlock / 64];
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double s = 0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) s += red[w];
  }
  __syncthreads();
  return s;
}

template <typename A, typename B>
__global__ void __launch_bounds__(kBlock)
dot_kernel(const A* a, const B* b, int64_t len, double* partial, const int* active) {
  const int r = blockIdx.y;
  if (active && !active[r]) return;
  const A* ar = a + int64_t(r) * len;
  const B* br = b + int64_t(r) * len;
  double acc = 0;
  for (int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x; i < len; i += int64_t(gridDim.x) * kBlock)
    acc += double(ar[i]) * double(br[i]);
  acc = block_sum(acc);
  if (threadIdx.x == 0) partial[int64_t(r) * gridDim.x + blockIdx.x] = acc;
}

// x += alpha p; r -= alpha q; partial ||r||^2.
template <typename T>
__global__ void __launch_bounds__(kBlock)
update_xr_kernel(T* x, T* r, const T* p, const T* q, int64_t len, const double* alpha,
                 double* partial, const int* active) {
  const int k = blockIdx.y;
  if (active && !active[k]) return;
  const T a = T(alpha[k]);
  const int64_t off = int64_t(k) * len;
  double acc = 0;
  for (int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x; i < len; i += int64_t(gridDim.x) * kBlock) {
    x[off + i] = fma(a, p[off + i], x[off + i]);
    const T rn = fma(-a, q[off + i], r[off + i]);
    r[off + i] = rn;
    acc += double(rn) * double(rn);
  }
  acc = block_sum(acc);
  if (threadIdx.x == 0) partial[int64_t(k) * gridDim.x + blockIdx.x] = acc;
}

// r -= alpha q; partial ||r||^2 (x is advanced later from the stored
// directions, see flush_kernel).
template <typename T>
__global__ void __launch_bounds__(kBlock)
update_r_kernel(T* r, const T* q, int64_t len, const double* alpha, double* partial, const int* active) {
  const int k = blockIdx.y;
  if (active && !active[k]) return;
  const T a = T(alpha[k]);
  const int64_t off = int64_t(k) * len;
  double acc = 0;
  for (int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x; i < len; i += int64_t(gridDim.x) * kBlock) {
    const T rn = fma(-a, q[off + i], r[off + i]);
    r[off + i] = rn;
    acc += double(rn) * double(rn);
  }
  acc = block_sum(acc);
  if (threadIdx.x == 0) partial[int64_t(k) * gridDim.x + blockIdx.x] = acc;
}

// x64 += sum_t alpha_t p_t over the stored search directions of every system,
// accumulated in FP64 (kAssign: x64 = sum_t alpha_t p_t, x64 not read).
// alpha_t = 0 where a system was not active; such terms are skipped, so a
// direction slot that a system never wrote does not reach its x64.
template <typename T, bool kAssign>
__global__ void __launch_bounds__(kBlock)
flush_kernel(double* x64, DirectionSet<T> d, const double* alpha, int nrhs, int64_t len) {
  const int k = blockIdx.y;
  const int64_t off = int64_t(k) * len;
  double a[kMaxDirections];
#pragma unroll
  for (int t = 0; t < kMaxDirections; ++t) a[t] = t < d.count ? alpha[t * nrhs + k] : 0.0;
  if constexpr (!kAssign) {  // no step pending for this system: x64 stays
    bool any = false;
#pragma unroll
    for (int t = 0; t < kMaxDirections; ++t) any = any || a[t] != 0.0;
    if (!any) return;
  }
  for (int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x; i < len; i += int64_t(gridDim.x) * kBlock) {
    double acc = kAssign ? 0.0 : x64[off + i];
#pragma unroll
    for (int t = 0; t < kMaxDirections; ++t)
      if (t < d.count && a[t] != 0.0) acc = fma(a[t], double(d.p[t][off + i]), acc);
    x64[off + i] = acc;
  }
}

// flush_kernel for len % 4 == 0: four consecutive entries per thread, so each
// wave moves 1-2 KB per stream and instruction over the up to ten streams.
__device__ __forceinline__ void load4(const float* p, double (&o)[4]) {
  const float4 v = *reinterpret_cast<const float4*>(p);
  o[0] = v.x;
  o[1] = v.y;
  o[2] = v.z;
  o[3] = v.w;
}
__device__ __forceinline__ void load4(const double* p, double (&o)[4]) {
  const double2 v0 = reinterpret_cast<const double2*>(p)[0], v1 = reinterpret_cast<const double2*>(p)[1];
  o[0] = v0.x;
  o[1] = v0.y;
  o[2] = v1.x;
  o[3] = v1.y;
}

template <typename T, bool kAssign>
__global__ void __launch_bounds__(kBlock)
flush4_kernel(double* x64, DirectionSet<T> d, const double* alpha, int nrhs, int64_t len) {
  const int k = blockIdx.y;
  const int64_t off = int64_t(k) * len;
  double a[kMaxDirections];
#pragma unroll
  for (int t = 0; t < kMaxDirections; ++t) a[t] = t < d.count ? alpha[t * nrhs + k] : 0.0;
  if constexpr (!kAssign) {  // no step pending for this system (converged or inactive): x64 stays
    bool any = false;
#pragma unroll
    for (int t = 0; t < kMaxDirections; ++t) any = any || a[t] != 0.0;
    if (!any) return;
  }
  const int64_t len4 = len / 4;
  for (int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x; i < len4; i += int64_t(gridDim.x) * kBlock) {
    double2* xp = reinterpret_cast<double2*>(x64 + off + 4 * i);
    double2 x0 = {0.0, 0.0}, x1 = {0.0, 0.0};
    if constexpr (!kAssign) {
      x0 = xp[0];
      x1 = xp[1];
    }
#pragma unroll
    for (int t = 0; t < kMaxDirections; ++t)
      if (t < d.count && a[t] != 0.0) {
        double v[4];
        load4(d.p[t] + off + 4 * i, v);
        x0.x = fma(a[t], v[0], x0.x);
        x0.y = fma(a[t], v[1], x0.y);
        x1.x = fma(a[t], v[2], x1.x);
        x1.y = fma(a[t], v[3], x1.y);
      }
    xp[0] = x0;
    xp[1] = x1;
  }
}

// Block sums of NV values per thread; partial[v * gridDim.x + blockIdx.x].
template <int NV>
__device__ __forceinline__ void block_sums(double (&v)[NV], double* partial) {
  __shared__ double red[kBlock / 64][NV];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    const double w = wave_sum(v[i]);
    if (lane == 0) red[wave][i] = w;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    double s = 0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) s += red[w][threadIdx.x];
    partial[int64_t(threadIdx.x) * gridDim.x + blockIdx.x] = s;
  }
}

// Gram matrix of six stacked fields f and u (len entries each):
// partial[(i * 6 + j) * gridDim.x + block] = sum f_i . u_j over the block.
__global__ void __launch_bounds__(kBlock)
gram6_kernel(const double* f, const double* u, int64_t len, double* partial) {
  double acc[36];
#pragma unroll
  for (int i = 0; i < 36; ++i) acc[i] = 0.0;
  for (int64_t e = int64_t(blockIdx.x) * kBlock + threadIdx.x; e < len; e += int64_t(gridDim.x) * kBlock) {
    double a[6], b[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      a[i] = f[int64_t(i) * len + e];
      b[i] = u[int64_t(i) * len + e];
    }
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
      for (int j = 0; j < 6; ++j) acc[i * 6 + j] = fma(a[i], b[j], acc[i * 6 + j]);
  }
  block_sums<36>(acc, partial);
}

// r64 = f - q64; partial ||r64||^2; optional r32 = float(r64).
__global__ void __launch_bounds__(kBlock)
residual_kernel(double* r, const double* f, const double* q, float* r32, int64_t len,
                double* partial, const int* active) {
  const int k = blockIdx.y;
  if (active && !active[k]) return;
  const int64_t off = int64_t(k) * len;
  double acc = 0;
  for (int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x; i < len; i += int64_t(gridDim.x) * kBlock) {
    const double v = f[off + i] - q[off + i];
    if (r) r[off + i] = v;
    if (r32) r32[off + i] = float(v);
    acc += v * v;
  }
  acc = block_sum(acc);
  if (threadIdx.x == 0) partial[int64_t(k) * gridDim.x + blockIdx.x] = acc;
}

// x64 += x32; x32 = 0.
__global__ void __launch_bounds__(kBlock)
accumulate_kernel(double* x64, float* x32, int64_t len, const int* active) {
  const int k = blockIdx.y;
  if (active && !active[k]) return;
  const int64_t off = int64_t(k) * len;
  for (int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x; i < len; i += int64_t(gridDim.x) * kBlock) {
    x64[off + i] += double(x32[off + i]);
    x32[off + i] = 0.0f;
  }
}

template <typename A, typename B>
__global__ void __launch_bounds__(kBlock)
convert_kernel(B* dst, const A* src, int64_t len, const int* active) {
  const int k = blockIdx.y;
  if (active && !active[k]) return;
  const int64_t off = int64_t(k) * len;
  for (int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x; i < len; i += int64_t(gridDim.x) * kBlock)
    dst[off + i] = B(src[off + i]);
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
fill_kernel(T* dst, int64_t len, const int* active) {
  const int k = blockIdx.y;
  if (active && !active[k]) return;
  const int64_t off = int64_t(k) * len;
  for (int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x; i < len; i += int64_t(gridDim.x) * kBlock)
    dst[off + i] = T(0);
}

// Per-(RHS, component) sums of an FP64 field: grid.y = 3 * nrhs rows of n.
__global__ void __launch_bounds__(kBlock)
row_sum_kernel(const double* u, int64_t n, double* partial) {
  const int row = blockIdx.y;
  const double* ur = u + int64_t(row) * n;
  double acc = 0;
  for (int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += int64_t(gridDim.x) * kBlock)
    acc += ur[i];
  acc = block_sum(acc);
  if (threadIdx.x == 0) partial[int64_t(row) * gridDim.x + blockIdx.x] = acc;
}

// One workgroup per row: mean[row] = inv_n * sum of its partials (fixed order).
__global__ void __launch_bounds__(kBlock)
row_mean_kernel(const double* partial, int nblk, double inv_n, double* mean) {
  const int row = blockIdx.x;
  double acc = 0;
  for (int b = threadIdx.x; b < nblk; b += kBlock) acc += partial[int64_t(row) * nblk + b];
  acc = block_sum(acc);
  if (threadIdx.x == 0) mean[row] = acc * inv_n;
}

__global__ void __launch_bounds__(kBlock)
subtract_mean_kernel(double* u, int64_t n, const double* mean, const int* active) {
  const int row = blockIdx.y;
  if (active && !active[row / 3]) return;
  const double m = mean[row];
  double* ur = u + int64_t(row) * n;
  for (int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += int64_t(gridDim.x) * kBlock)
    ur[i] -= m;
}

// f -= mean (per RHS and component), r = TR(f), partial ||f||^2: the RHS
// projection, its norm and the first residual in one pass.
template <typename TR>
__global__ void __launch_bounds__(kBlock)
center_copy_kernel(double* f, TR* r, const double* mean, int64_t n, double* partial) {
  const int k = blockIdx.y;
  const int64_t len = 3 * n, off = int64_t(k) * len;
  double acc = 0;
  for (int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x; i < len; i += int64_t(gridDim.x) * kBlock) {
    const double v = f[off + i] - mean[k * 3 + int(i / n)];
    f[off + i] = v;
    r[off + i] = TR(v);
    acc += v * v;
  }
  acc = block_sum(acc);
  if (threadIdx.x == 0) partial[int64_t(k) * gridDim.x + blockIdx.x] = acc;
}

// One workgroup per RHS: total = sum partial[k][0..nblk) in a fixed order,
// then the scalar update selected by op.
__global__ void __launch_bounds__(kBlock)
finalize_kernel(const double* partial, int nblk, int op, double* scal, int nrhs, int* active,
                int* restart, int iteration, double tol, int* stop, double* hist) {
  const int k = blockIdx.x;
  if (active && !active[k]) {
    if (hist && threadIdx.x == 0) hist[k] = 0.0;  // no step for an inactive system
    return;
  }
  double acc = 0;
  for (int b = threadIdx.x; b < nblk; b += kBlock) acc += partial[int64_t(k) * nblk + b];
  acc = block_sum(acc);
  if (threadIdx.x != 0) return;
  double* S = scal;
  switch (op) {
    case kOpStore:
      S[kTmp * nrhs + k] = acc;
      break;
    case kOpRho:
      S[kRho * nrhs + k] = acc;
      break;
    case kOpFF:
      S[kFF * nrhs + k] = acc;
      break;
    case kOpRR:
      S[kRR * nrhs + k] = acc;
      // Convergence latch (femlib/batched/pcg.py:676-685 semantics): the first
      // iteration whose recursive residual is below tol is recorded and the
      // system is parked (u frozen) until the host certifies it.
      if (stop && sqrt(acc / S[kFF * nrhs + k]) < tol && stop[k] == 0) {
        stop[k] = iteration;
        active[k] = 0;
      }
      break;
    case kOpAlpha: {
      S[kPQ * nrhs + k] = acc;
      const bool bad = !isfinite(acc) || acc <= 0.0;
      S[kAlpha * nrhs + k] = bad ? 0.0 : S[kRho * nrhs + k] / fmax(acc, 1e-300);
      if (hist) hist[k] = S[kAlpha * nrhs + k];
      break;
    }
    case kOpBeta: {
      const double rho_old = S[kRho * nrhs + k];
      double beta = rho_old > 0.0 ? acc / rho_old : 0.0;
      if (restart && restart[k]) {  // consumed here, so a captured block needs no host reset
        beta = 0.0;
        restart[k] = 0;
      }
      if (!isfinite(beta)) beta = 0.0;
      S[kBeta * nrhs + k] = beta;
      S[kRho * nrhs + k] = acc;
      break;
    }
  }
}

inline dim3 vgrid(int64_t len, int nrhs) {
  int64_t b = (len + kBlock - 1) / kBlock;
  if (b > kVecBlocks) b = kVecBlocks;
  if (b < 1) b = 1;
  return dim3(unsigned(b), unsigned(nrhs));
}

}  // namespace

int vec_blocks(int64_t len) {
  int64_t b = (len + kBlock - 1) / kBlock;
  return int(b > kVecBlocks ? kVecBlocks : (b < 1 ? 1 : b));
}

template <typename A, typename B>
int launch_dot(const A* a, const B* b, int64_t len, int nrhs, double* partial, const int* active,
               hipStream_t s) {
  hipLaunchKernelGGL((dot_kernel<A, B>), vgrid(len, nrhs), dim3(kBlock), 0, s, a, b, len, partial, active);
  VXL_HIP(hipGetLastError());
  return VXL_OK;
}
template int launch_dot<float, float>(const float*, const float*, int64_t, int, double*, const int*, hipStream_t);
template int launch_dot<double, double>(const double*, const double*, int64_t, int, double*, const int*, hipStream_t);

int launch_finalize(const double* partial, int nblk, int op, double* scal, int nrhs, int* active,
                    int* restart, hipStream_t s, int iteration, double tol, int* stop, double* hist) {
  hipLaunchKernelGGL(finalize_kernel, dim3(nrhs), dim3(kBlock), 0, s, partial, nblk, op, scal, nrhs,
                     active, restart, iteration, tol, stop, hist);
  VXL_HIP(hipGetLastError());
  return VXL_OK;
}

int launch_gram6(const double* f, const double* u, int64_t len, double* partial, double scale, double* out,
                 hipStream_t s) {
  hipLaunchKernelGGL(gram6_kernel, dim3(kGramBlocks), dim3(kBlock), 0, s, f, u, len, partial);
  VXL_HIP(hipGetLastError());
  hipLaunchKernelGGL(row_mean_kernel, dim3(36), dim3(kBlock), 0, s, partial, kGramBlocks, scale, out);
  VXL_HIP(hipGetLastError());
  return VXL_OK;
}

int launch_row_totals(const double* partial, int rows, int nblk, double scale, double* out, hipStream_t s) {
  hipLaunchKernelGGL(row_mean_kernel, dim3(rows), dim3(kBlock), 0, s, partial, nblk, scale, out);
  VXL_HIP(hipGetLastError());
  return VXL_OK;
}

template <typename T>
int launch_update_r(T* r, const T* q, int64_t len, int nrhs, const double* alpha, double* partial,
                    const int* active, hipStream_t s) {
  hipLaunchKernelGGL((update_r_kernel<T>), vgrid(len, nrhs), dim3(kBlock), 0, s, r, q, len, alpha, partial, active);
  VXL_HIP(hipGetLastError());
  return VXL_OK;
}
template int launch_update_r<float>(float*, const float*, int64_t, int, const double*, double*, const int*,
                                    hipStream_t);
template int launch_update_r<double>(double*, const double*, int64_t, int, const double*, double*, const int*,
                                     hipStream_t);

template <typename T>
int launch_flush(double* x64, const DirectionSet<T>& d, const double* alpha, int nrhs, int64_t len, bool assign,
                 hipStream_t s) {
  if (d.count <= 0) return VXL_OK;
  if (d.count > kMaxDirections) return fail("too many stored search directions");
  const bool v4 = len % 4 == 0;
  const dim3 g = vgrid(v4 ? len / 4 : len, nrhs);
  if (v4 && assign)
    hipLaunchKernelGGL((flush4_kernel<T, true>), g, dim3(kBlock), 0, s, x64, d, alpha, nrhs, len);
  else if (v4)
    hipLaunchKernelGGL((flush4_kernel<T, false>), g, dim3(kBlock), 0, s, x64, d, alpha, nrhs, len);
  else if (assign)
    hipLaunchKernelGGL((flush_kernel<T, true>), g, dim3(kBlock), 0, s, x64, d, alpha, nrhs, len);
  else
    hipLaunchKernelGGL((flush_kernel<T, false>), g, dim3(kBlock), 0, s, x64, d, alpha, nrhs, len);
  VXL_HIP(hipGetLastError());
  return VXL_OK;
}
template int launch_flush<float>(double*, const DirectionSet<float>&, const double*, int, int64_t, bool,
                                 hipStream_t);
template int launch_flush<double>(double*, const DirectionSet<double>&, const double*, int, int64_t, bool,
                                  hipStream_t);

template <typename T>
int launch_update_xr(T* x, T* r, const T* p, const T* q, int64_t len, int nrhs, const double* alpha,
                     double* partial, const int* active, hipStream_t s) {
  hipLaunchKernelGGL((update_xr_kernel<T>), vgrid(len, nrhs), dim3(kBlock), 0, s, x, r, p, q, len,
                     alpha, partial, active);
  VXL_HIP(hipGetLastError());
  return VXL_OK;
}
template int launch_update_xr<float>(float*, float*, const float*, const float*, int64_t, int, const double*, double*, const int*, hipStream_t);
template int launch_update_xr<double>(double*, double*, const double*, const double*, int64_t, int, const double*, double*, const int*, hipStream_t);


int launch_residual(double* r, const double* f, const double* q, float* r32, int64_t len, int nrhs,
                    double* partial, const int* active, hipStream_t s) {
  hipLaunchKernelGGL(residual_kernel, vgrid(len, nrhs), dim3(kBlock), 0, s, r, f, q, r32, len, partial, active);
  VXL_HIP(hipGetLastError());
  return VXL_OK;
}

int launch_accumulate(double* x64, float* x32, int64_t len, int nrhs, const int* active, hipStream_t s) {
  hipLaunchKernelGGL(accumulate_kernel, vgrid(len, nrhs), dim3(kBlock), 0, s, x64, x32, len, active);
  VXL_HIP(hipGetLastError());
  return VXL_OK;
}

template <typename A, typename B>
int launch_convert(B* dst, const A* src, int64_t len, int nrhs, const int* active, hipStream_t s) {
  hipLaunchKernelGGL((convert_kernel<A, B>), vgrid(len, nrhs), dim3(kBlock), 0, s, dst, src, len, active);
  VXL_HIP(hipGetLastError());
  return VXL_OK;
}
template int launch_convert<double, float>(float*, const double*, int64_t, int, const int*, hipStream_t);
template int launch_convert<double, double>(double*, const double*, int64_t, int, const int*, hipStream_t);
template int launch_convert<float, float>(float*, const float*, int64_t, int, const int*, hipStream_t);

template <typename T>
int launch_fill_zero(T* dst, int64_t len, int nrhs, const int* active, hipStream_t s) {
  hipLaunchKernelGGL((fill_kernel<T>), vgrid(len, nrhs), dim3(kBlock), 0, s, dst, len, active);
  VXL_HIP(hipGetLastError());
  return VXL_OK;
}
template int launch_fill_zero<float>(float*, int64_t, int, const int*, hipStream_t);
template int launch_fill_zero<double>(double*, int64_t, int, const int*, hipStream_t);

template <typename TR>
int launch_center_copy(double* f, TR* r, int64_t n, int nrhs, double* partial, int* nblk_out, hipStream_t s) {
  const dim3 g = vgrid(n, 3 * nrhs);
  double* mean = partial + int64_t(3) * nrhs * g.x;  // 3*nrhs slots after the row partials
  hipLaunchKernelGGL(row_sum_kernel, g, dim3(kBlock), 0, s, f, n, partial);
  VXL_HIP(hipGetLastError());
  hipLaunchKernelGGL(row_mean_kernel, dim3(3 * nrhs), dim3(kBlock), 0, s, partial, int(g.x), 1.0 / double(n), mean);
  VXL_HIP(hipGetLastError());
  // The norm partials go after the means (the row partials are consumed).
  double* norms = mean + 3 * nrhs;
  const dim3 gv = vgrid(3 * n, nrhs);
  hipLaunchKernelGGL((center_copy_kernel<TR>), gv, dim3(kBlock), 0, s, f, r, mean, n, norms);
  VXL_HIP(hipGetLastError());
  *nblk_out = int(gv.x);
  return VXL_OK;
}
template int launch_center_copy<float>(double*, float*, int64_t, int, double*, int*, hipStream_t);
template int launch_center_copy<double>(double*, double*, int64_t, int, double*, int*, hipStream_t);

int launch_project_zero_mean(double* u, int64_t n, int nrhs, double* partial, const int* active,
                             hipStream_t s) {
  const dim3 g = vgrid(n, 3 * nrhs);
  double* mean = partial + int64_t(3) * nrhs * g.x;  // 3*nrhs slots after the partials
  hipLaunchKernelGGL(row_sum_kernel, g, dim3(kBlock), 0, s, u, n, partial);
  VXL_HIP(hipGetLastError());
  hipLaunchKernelGGL(row_mean_kernel, dim3(3 * nrhs), dim3(kBlock), 0, s, partial, int(g.x),
                     1.0 / double(n), mean);
  VXL_HIP(hipGetLastError());
  hipLaunchKernelGGL(subtract_mean_kernel, g, dim3(kBlock), 0, s, u, n, mean, active);
  VXL_HIP(hipGetLastError());
  return VXL_OK;
}

}  // namespace vxl
