// Material front end on the device: voxel label -> coefficient row.
//
// The label half of the reference's pool construction (rows [0, L) are the
// per-label C21 rows in ascending label order, hybrid_material.py:38-48, with
// phase_table's label order, material_field.py:57-68): every voxel of label l
// gets row rank(l).  Here it is one pass over the resident label array with a
// dense lookup table over [label_lo, label_lo + span), fused with what the
// host needs next: which labels occur (the phases the reference-medium design
// looks at) and the rows at the Zeman sampling lattice
// (hybrid_adapter.py:101-125: every max(1, n/16)-th voxel per axis).
#include "vxl_internal.h"

namespace vxl {
namespace {

struct Lattice {
  int sz, sy, sx;  // strides
  int cz, cy, cx;  // counts
};

// lut[l - lo] >= 0: the label's row; -1: no material; kLocalRow: a local-frame
// phase whose voxels get their own row local_base + e.
constexpr int kLocalRow = -2;

template <typename L, typename Idx>
__global__ void __launch_bounds__(256)
label_rows_kernel(const L* __restrict__ labels, int nx, int ny, int nz, long long lo, int span,
                  const int* __restrict__ lut, Idx* __restrict__ index, int* __restrict__ present,
                  int* __restrict__ undefined, Lattice lat, int* __restrict__ samples, int64_t local_base) {
  const int64_t nxy = int64_t(nx) * ny, n = nxy * nz;
  for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < n; e += int64_t(gridDim.x) * blockDim.x) {
    const int x = int(e % nx), y = int((e / nx) % ny), z = int(e / nxy);
    const long long l = static_cast<long long>(labels[e]);
    const long long off = l - lo;
    const int entry = (off >= 0 && off < span) ? lut[off] : -1;
    int64_t row = entry;
    if (entry == kLocalRow) row = local_base + e;
    if (entry == -1) {
      undefined[0] = 1;
      row = 0;
    } else if (x == 0 || static_cast<long long>(labels[e - 1]) != l) {
      present[off] = 1;  // first voxel of each x-run marks its label
    }
    index[e] = static_cast<Idx>(row);
    if (z % lat.sz == 0 && y % lat.sy == 0 && x % lat.sx == 0)
      samples[(int64_t(z / lat.sz) * lat.cy + y / lat.sy) * lat.cx + x / lat.sx] = int(row);
  }
}

template <typename L, typename Idx>
void launch_label_rows(const void* labels, int nx, int ny, int nz, long long lo, int span, const int* lut,
                       void* index, int* present, int* undefined, Lattice lat, int* samples, int64_t local_base,
                       hipStream_t s) {
  const int64_t n = int64_t(nx) * ny * nz;
  const int blocks = int(std::min<int64_t>((n + 255) / 256, 65536));
  hipLaunchKernelGGL((label_rows_kernel<L, Idx>), dim3(blocks), dim3(256), 0, s, static_cast<const L*>(labels), nx,
                     ny, nz, lo, span, lut, static_cast<Idx*>(index), present, undefined, lat, samples, local_base);
}

// Bond stress matrix T of a rotation R (C' = T C T^T), engineering Voigt
// order 11, 22, 33, 23, 13, 12; the pairs (i, j) of shear rows/columns are
// (1, 2), (0, 2), (0, 1).
__device__ __forceinline__ void bond_matrix(const double (&R)[3][3], double (&T)[6][6]) {
#pragma clang fp contract(off)
  const int pi[3] = {1, 0, 0}, pj[3] = {2, 2, 1};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
#pragma unroll
    for (int b = 0; b < 3; ++b) T[a][b] = R[a][b] * R[a][b];
#pragma unroll
    for (int k = 0; k < 3; ++k) T[a][3 + k] = 2.0 * R[a][pi[k]] * R[a][pj[k]];
  }
#pragma unroll
  for (int m = 0; m < 3; ++m) {
#pragma unroll
    for (int b = 0; b < 3; ++b) T[3 + m][b] = R[pi[m]][b] * R[pj[m]][b];
#pragma unroll
    for (int k = 0; k < 3; ++k)
      T[3 + m][3 + k] = R[pi[m]][pi[k]] * R[pj[m]][pj[k]] + R[pi[m]][pj[k]] * R[pj[m]][pi[k]];
  }
}

// Rows of local-frame voxels: frame from an axis pair (columns a, c x a, c with
// a = axis1/|axis1|, c = a x b / |a x b|, b = axis2/|axis2|) or a rotation
// matrix, then C' = T C_phase T^T written at rows[local_base + e].  bad[k] = 1
// flags: 0 nonfinite, 1 zero axis, 2 collinear axes, 3 not orthonormal,
// 4 not right-handed (the host path's checks and tolerances).
template <typename L>
__global__ void __launch_bounds__(128)
local_rows_kernel(const L* __restrict__ labels, int64_t n, long long lo, int span, const int* __restrict__ slot,
                  const double* __restrict__ cphase, int mode, const double* __restrict__ A, const double* __restrict__ B,
                  int64_t stride, double* __restrict__ rows, int64_t local_base, int* __restrict__ bad) {
  // Every product is rounded on its own (no fused multiply-add) and every sum
  // runs in ascending index order: the host front end (materials.frames,
  // materials.bond_rotate, NumPy elementwise operations) evaluates the same
  // expressions in the same order, so the rows agree bit for bit.
#pragma clang fp contract(off)
  for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < n; e += int64_t(gridDim.x) * blockDim.x) {
    const long long off = static_cast<long long>(labels[e]) - lo;
    if (off < 0 || off >= span) continue;
    const int ph = slot[off];
    if (ph < 0) continue;
    double R[3][3];
    bool ok = true;
    if (mode == 1) {
      double a[3], b[3];
      for (int i = 0; i < 3; ++i) {
        a[i] = A[e * stride + i];
        b[i] = B[e * stride + i];
      }
      if (!(isfinite(a[0]) && isfinite(a[1]) && isfinite(a[2]) && isfinite(b[0]) && isfinite(b[1]) &&
            isfinite(b[2]))) {
        bad[0] = 1;
        ok = false;
      }
      const double na = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
      const double nb = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
      if (ok && (!(na > 0) || !(nb > 0))) {
        bad[1] = 1;
        ok = false;
      }
      if (ok) {
        for (int i = 0; i < 3; ++i) {
          a[i] /= na;
          b[i] /= nb;
        }
        double c[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
        const double nc = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
        if (!(nc > 1e-12)) {
          bad[2] = 1;
          ok = false;
        } else {
          for (int i = 0; i < 3; ++i) c[i] /= nc;
          const double d[3] = {c[1] * a[2] - c[2] * a[1], c[2] * a[0] - c[0] * a[2], c[0] * a[1] - c[1] * a[0]};
          for (int i = 0; i < 3; ++i) {
            R[i][0] = a[i];
            R[i][1] = d[i];
            R[i][2] = c[i];
          }
        }
      }
    } else {
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) R[i][j] = A[e * stride + 3 * i + j];
      for (int i = 0; i < 3 && ok; ++i)
        for (int j = 0; j < 3; ++j)
          if (!isfinite(R[i][j])) {
            bad[0] = 1;
            ok = false;
            break;
          }
      if (ok) {
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) {
            const double g = R[0][i] * R[0][j] + R[1][i] * R[1][j] + R[2][i] * R[2][j];
            if (fabs(g - (i == j ? 1.0 : 0.0)) > 1e-10) ok = false;
          }
        if (!ok) bad[3] = 1;
      }
      if (ok) {
        const double det = R[0][0] * (R[1][1] * R[2][2] - R[1][2] * R[2][1]) -
                           R[0][1] * (R[1][0] * R[2][2] - R[1][2] * R[2][0]) +
                           R[0][2] * (R[1][0] * R[2][1] - R[1][1] * R[2][0]);
        if (fabs(det - 1.0) > 1e-10) {
          bad[4] = 1;
          ok = false;
        }
      }
    }
    double* out = rows + (local_base + e) * 21;
    if (!ok) {
      for (int k = 0; k < 21; ++k) out[k] = 0.0;
      continue;
    }
    double T[6][6];
    bond_matrix(R, T);
    const double* c21 = cphase + int64_t(ph) * 21;
    double TC[6][6];
    for (int i = 0; i < 6; ++i)
      for (int j = 0; j < 6; ++j) {
        double acc = T[i][0] * c21[pidx(0, j)];
        for (int k = 1; k < 6; ++k) acc = acc + T[i][k] * c21[pidx(k, j)];
        TC[i][j] = acc;
      }
    int w = 0;
    for (int i = 0; i < 6; ++i)
      for (int j = i; j < 6; ++j, ++w) {
        double acc = TC[i][0] * T[j][0];
        for (int k = 1; k < 6; ++k) acc = acc + TC[i][k] * T[j][k];
        out[w] = acc;
      }
  }
}

}  // namespace

int label_rows(int nz, int ny, int nx, int label_bytes, const void* labels_dev, long long label_lo, int span,
               const int* lut_host, int index_kind, void* index_dev, int* present_host, int* undefined_host,
               int* samples_host, int64_t local_base, hipStream_t s) {
  Lattice lat;
  lat.sz = std::max(1, nz / 16);
  lat.sy = std::max(1, ny / 16);
  lat.sx = std::max(1, nx / 16);
  lat.cz = (nz + lat.sz - 1) / lat.sz;
  lat.cy = (ny + lat.sy - 1) / lat.sy;
  lat.cx = (nx + lat.sx - 1) / lat.sx;
  const size_t nsamp = size_t(lat.cz) * lat.cy * lat.cx;
  // One block: lut | present | undefined | samples.
  const size_t words = size_t(span) * 2 + 1 + nsamp;
  void* buf = nullptr;
  VXL_HIP(dev_malloc(&buf, sizeof(int) * words));
  int* lut = static_cast<int*>(buf);
  int* present = lut + span;
  int* undefined = present + span;
  int* samples = undefined + 1;
  int rc = VXL_OK;
  auto body = [&]() -> int {
    VXL_HIP(hipMemcpyAsync(lut, lut_host, sizeof(int) * span, hipMemcpyHostToDevice, s));
    VXL_HIP(hipMemsetAsync(present, 0, sizeof(int) * (size_t(span) + 1), s));
#define VXL_LR(L, I) launch_label_rows<L, I>(labels_dev, nx, ny, nz, label_lo, span, lut, index_dev, present, \
                                             undefined, lat, samples, local_base, s)
    if (label_bytes == 4) {
      if (index_kind == VXL_INDEX_U8) VXL_LR(int32_t, uint8_t);
      else if (index_kind == VXL_INDEX_U16) VXL_LR(int32_t, uint16_t);
      else VXL_LR(int32_t, int32_t);
    } else {
      if (index_kind == VXL_INDEX_U8) VXL_LR(int64_t, uint8_t);
      else if (index_kind == VXL_INDEX_U16) VXL_LR(int64_t, uint16_t);
      else VXL_LR(int64_t, int32_t);
    }
#undef VXL_LR
    VXL_HIP(hipGetLastError());
    VXL_HIP(hipMemcpyAsync(present_host, present, sizeof(int) * span, hipMemcpyDeviceToHost, s));
    VXL_HIP(hipMemcpyAsync(undefined_host, undefined, sizeof(int), hipMemcpyDeviceToHost, s));
    VXL_HIP(hipMemcpyAsync(samples_host, samples, sizeof(int) * nsamp, hipMemcpyDeviceToHost, s));
    VXL_HIP(hipStreamSynchronize(s));
    return VXL_OK;
  };
  rc = body();
  (void)hipStreamSynchronize(s);
  dev_free(buf);
  return rc;
}

int local_rows(int64_t n, int label_bytes, const void* labels_dev, long long label_lo, int span, const int* slot_host,
               int phases, const double* c21_host, int mode, const double* a_dev, const double* b_dev, int64_t stride,
               double* rows_dev, int64_t local_base, int* bad_host, hipStream_t s) {
  const size_t words = size_t(span) + 8;
  void* ibuf = nullptr;
  void* cbuf = nullptr;
  VXL_HIP(dev_malloc(&ibuf, sizeof(int) * words));
  if (dev_malloc(&cbuf, sizeof(double) * 21 * size_t(phases)) != hipSuccess) {
    dev_free(ibuf);
    return fail("hipMalloc failed for the local-frame tables");
  }
  int* slot = static_cast<int*>(ibuf);
  int* bad = slot + span;
  auto body = [&]() -> int {
    VXL_HIP(hipMemcpyAsync(slot, slot_host, sizeof(int) * span, hipMemcpyHostToDevice, s));
    VXL_HIP(hipMemsetAsync(bad, 0, sizeof(int) * 8, s));
    VXL_HIP(hipMemcpyAsync(cbuf, c21_host, sizeof(double) * 21 * phases, hipMemcpyHostToDevice, s));
    const int blocks = int(std::min<int64_t>((n + 127) / 128, 65536));
    if (label_bytes == 4)
      hipLaunchKernelGGL((local_rows_kernel<int32_t>), dim3(blocks), dim3(128), 0, s,
                         static_cast<const int32_t*>(labels_dev), n, label_lo, span, slot,
                         static_cast<const double*>(cbuf), mode, a_dev, b_dev, stride, rows_dev, local_base, bad);
    else
      hipLaunchKernelGGL((local_rows_kernel<int64_t>), dim3(blocks), dim3(128), 0, s,
                         static_cast<const int64_t*>(labels_dev), n, label_lo, span, slot,
                         static_cast<const double*>(cbuf), mode, a_dev, b_dev, stride, rows_dev, local_base, bad);
    VXL_HIP(hipGetLastError());
    VXL_HIP(hipMemcpyAsync(bad_host, bad, sizeof(int) * 5, hipMemcpyDeviceToHost, s));
    VXL_HIP(hipStreamSynchronize(s));
    return VXL_OK;
  };
  const int rc = body();
  (void)hipStreamSynchronize(s);
  dev_free(cbuf);
  dev_free(ibuf);
  return rc;
}

}  // namespace vxl
