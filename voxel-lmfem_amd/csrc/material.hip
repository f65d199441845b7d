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

template <typename L, typename Idx>
__global__ void __launch_bounds__(256)
label_rows_kernel(const L* __restrict__ labels, int nx, int ny, int nz, long long lo, int span,
                  const int* __restrict__ lut, Idx* __restrict__ index, int* __restrict__ present,
                  int* __restrict__ undefined, Lattice lat, int* __restrict__ samples) {
  const int64_t nxy = int64_t(nx) * ny, n = nxy * nz;
  for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < n; e += int64_t(gridDim.x) * blockDim.x) {
    const int x = int(e % nx), y = int((e / nx) % ny), z = int(e / nxy);
    const long long l = static_cast<long long>(labels[e]);
    const long long off = l - lo;
    int row = (off >= 0 && off < span) ? lut[off] : -1;
    if (row < 0) {
      undefined[0] = 1;
      row = 0;
    } else if (x == 0 || static_cast<long long>(labels[e - 1]) != l) {
      present[off] = 1;  // first voxel of each x-run marks its label
    }
    index[e] = static_cast<Idx>(row);
    if (z % lat.sz == 0 && y % lat.sy == 0 && x % lat.sx == 0)
      samples[(int64_t(z / lat.sz) * lat.cy + y / lat.sy) * lat.cx + x / lat.sx] = row;
  }
}

template <typename L, typename Idx>
void launch_label_rows(const void* labels, int nx, int ny, int nz, long long lo, int span, const int* lut,
                       void* index, int* present, int* undefined, Lattice lat, int* samples, hipStream_t s) {
  const int64_t n = int64_t(nx) * ny * nz;
  const int blocks = int(std::min<int64_t>((n + 255) / 256, 65536));
  hipLaunchKernelGGL((label_rows_kernel<L, Idx>), dim3(blocks), dim3(256), 0, s, static_cast<const L*>(labels), nx,
                     ny, nz, lo, span, lut, static_cast<Idx*>(index), present, undefined, lat, samples);
}

}  // namespace

int label_rows(int nz, int ny, int nx, int label_bytes, const void* labels_dev, long long label_lo, int span,
               const int* lut_host, int index_kind, void* index_dev, int* present_host, int* undefined_host,
               int* samples_host, hipStream_t s) {
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
                                             undefined, lat, samples, s)
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

}  // namespace vxl
