// Element-energy contraction of converged load-case fields: the adjoint of
// C_eff with respect to a per-voxel stiffness scale.
//
// With C_e = s_e Chat_e and u_j the discrete periodic fluctuation of macro
// strain E_j, at the discrete solution
//   C_ij = (1/|Omega|) sum_e s_e int_e (E_i + B u_i)^T Chat_e (E_j + B u_j) dV,
// and because every u_j makes the energy stationary, the derivative of the
// fields drops out (the problem is self-adjoint, no second solve):
//   dC_ij / ds_e = (1/|Omega|) int_e (E_i + B u_i)^T Chat_e (E_j + B u_j) dV.
// For a functional J = sum_ij W_ij C_ij the contraction is evaluated as
//   sum_i int_e (E_i + B u_i)^T Chat_e (Et_i + B ut_i) dV,
// with Et = W E and ut = W u formed first (one vector pass), using the same
// 2x2x2 Gauss rule as the operator so it is exact for the discrete energy.
#include "vxl_internal.h"

namespace vxl {
namespace {

// ut[i][c][n] = sum_j W[i][j] u[j][c][n]
__global__ void __launch_bounds__(256)
combine_kernel(const double* u, const double* W, int nload, int64_t len, double* ut) {
  const int i = blockIdx.y;
  for (int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < len; k += int64_t(gridDim.x) * blockDim.x) {
    double acc = 0;
    for (int j = 0; j < nload; ++j) acc = fma(W[i * nload + j], u[int64_t(j) * len + k], acc);
    ut[int64_t(i) * len + k] = acc;
  }
}

// d21 != nullptr: contract with the per-element tensor d21[e] (21 coefficients)
// instead of the cell's coefficient row.
template <typename Idx>
__global__ void __launch_bounds__(128)
energy_gradient_kernel(const void* index, const double* table, const double* d21, const double* u,
                       const double* ut, const double* E, const double* Et, int nload, double* out, int nx, int ny,
                       int nz, double hx, double hy, double hz) {
  const int64_t nxy = int64_t(nx) * ny, n = nxy * nz;
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int x = int(e % nx), y = int((e / nx) % ny), z = int(e / nxy);
  int64_t node[8];
#pragma unroll
  for (int a = 0; a < 8; ++a)
    node[a] = int64_t((z + (a >> 2)) % nz) * nxy + int64_t((y + ((a >> 1) & 1)) % ny) * nx + (x + (a & 1)) % nx;
  const double* c = d21 ? d21 + e * 21 : table + int64_t(load_row<Idx>(index, e)) * kRowWidth;
  double C[21];
#pragma unroll
  for (int k = 0; k < 21; ++k) C[k] = c[k];
  const double gp = 0.57735026918962576451;  // 1/sqrt(3)
  const double wv = hx * hy * hz / 8.0;
  double total = 0;
  for (int i = 0; i < nload; ++i) {
    double a1[8][3], a2[8][3];
    const double* ui = u + int64_t(i) * 3 * n;
    const double* vi = ut + int64_t(i) * 3 * n;
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
      for (int cc = 0; cc < 3; ++cc) {
        a1[a][cc] = ui[cc * n + node[a]];
        a2[a][cc] = vi[cc * n + node[a]];
      }
#pragma unroll
    for (int g = 0; g < 8; ++g) {
      const double xi = (g & 1) ? gp : -gp, eta = ((g >> 1) & 1) ? gp : -gp, zeta = (g >> 2) ? gp : -gp;
      double d1[3][3] = {}, d2[3][3] = {};  // [component][derivative axis]
#pragma unroll
      for (int a = 0; a < 8; ++a) {
        const double sx = (a & 1) ? 1.0 : -1.0, sy = ((a >> 1) & 1) ? 1.0 : -1.0, sz = (a >> 2) ? 1.0 : -1.0;
        const double fx = 1 + sx * xi, fy = 1 + sy * eta, fz = 1 + sz * zeta;
        const double gx = sx * fy * fz / (4 * hx), gy = sy * fx * fz / (4 * hy), gz = sz * fx * fy / (4 * hz);
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) {
          d1[cc][0] = fma(gx, a1[a][cc], d1[cc][0]);
          d1[cc][1] = fma(gy, a1[a][cc], d1[cc][1]);
          d1[cc][2] = fma(gz, a1[a][cc], d1[cc][2]);
          d2[cc][0] = fma(gx, a2[a][cc], d2[cc][0]);
          d2[cc][1] = fma(gy, a2[a][cc], d2[cc][1]);
          d2[cc][2] = fma(gz, a2[a][cc], d2[cc][2]);
        }
      }
      const double e1[6] = {E[6 * i + 0] + d1[0][0], E[6 * i + 1] + d1[1][1], E[6 * i + 2] + d1[2][2],
                            E[6 * i + 3] + d1[1][2] + d1[2][1], E[6 * i + 4] + d1[0][2] + d1[2][0],
                            E[6 * i + 5] + d1[0][1] + d1[1][0]};
      const double e2[6] = {Et[6 * i + 0] + d2[0][0], Et[6 * i + 1] + d2[1][1], Et[6 * i + 2] + d2[2][2],
                            Et[6 * i + 3] + d2[1][2] + d2[2][1], Et[6 * i + 4] + d2[0][2] + d2[2][0],
                            Et[6 * i + 5] + d2[0][1] + d2[1][0]};
      double w = 0;
#pragma unroll
      for (int r = 0; r < 6; ++r) {
        double s = 0;
#pragma unroll
        for (int k = 0; k < 6; ++k) s = fma(C[pidx(r, k)], e2[k], s);
        w = fma(e1[r], s, w);
      }
      total = fma(wv, w, total);
    }
  }
  out[e] = total;
}

}  // namespace

int energy_gradient(const vxl_cell* cell, int nload, const double* u, const double* macro_host,
                    const double* w_host, const double* d21, double* out, hipStream_t s) {
  const int64_t n = cell->grid.n, len = 3 * n;
  std::vector<double> Et(6 * nload, 0.0);
  for (int i = 0; i < nload; ++i)
    for (int j = 0; j < nload; ++j)
      for (int k = 0; k < 6; ++k) Et[6 * i + k] += w_host[i * nload + j] * macro_host[6 * j + k];
  // small: W, E, Et | large: ut
  void* small = nullptr;
  void* ut = nullptr;
  VXL_HIP(dev_malloc(&small, sizeof(double) * (nload * nload + 12 * nload)));
  const hipError_t ea = dev_malloc(&ut, sizeof(double) * size_t(len) * nload);
  if (ea != hipSuccess) {
    dev_free(small);
    return fail("hipMalloc failed for the adjoint work fields");
  }
  double* W = static_cast<double*>(small);
  double* E = W + nload * nload;
  double* Etd = E + 6 * nload;
  auto body = [&]() -> int {
    VXL_HIP(hipMemcpyAsync(W, w_host, sizeof(double) * nload * nload, hipMemcpyHostToDevice, s));
    VXL_HIP(hipMemcpyAsync(E, macro_host, sizeof(double) * 6 * nload, hipMemcpyHostToDevice, s));
    VXL_HIP(hipMemcpyAsync(Etd, Et.data(), sizeof(double) * 6 * nload, hipMemcpyHostToDevice, s));
    const unsigned vb = unsigned(std::min<int64_t>((len + 255) / 256, 4096));
    hipLaunchKernelGGL(combine_kernel, dim3(vb, nload), dim3(256), 0, s, u, W, nload, len,
                       static_cast<double*>(ut));
    VXL_HIP(hipGetLastError());
    const unsigned eb = unsigned((n + 127) / 128);
    const auto& g = cell->grid;
#define VXL_EG(I)                                                                                           \
  hipLaunchKernelGGL((energy_gradient_kernel<I>), dim3(eb), dim3(128), 0, s, cell->mat.index, cell->mat.c64, d21, u, \
                     static_cast<const double*>(ut), E, Etd, nload, out, g.nx, g.ny, g.nz, cell->h[0],        \
                     cell->h[1], cell->h[2])
    switch (cell->mat.index_kind) {
      case VXL_INDEX_U8: VXL_EG(uint8_t); break;
      case VXL_INDEX_U16: VXL_EG(uint16_t); break;
      default: VXL_EG(int32_t); break;
    }
#undef VXL_EG
    VXL_HIP(hipGetLastError());
    VXL_HIP(hipStreamSynchronize(s));
    return VXL_OK;
  };
  const int rc = body();
  (void)hipStreamSynchronize(s);
  dev_free(ut);
  dev_free(small);
  return rc;
}

}  // namespace vxl
