// Label-grouped dense element stiffness on the gfx950 matrix cores: the FP32
// operator q = K p as f_e = K_L (u_e - u_e0) per element, with K_L the dense
// 24x24 element matrix of coefficient row L, run as
// v_mfma_f32_16x16x32_f16 products.  The reference's counterpart is its
// grouped element kernel (_core/runtime/fi/grouped_kernels.cu:12-74, one dense
// Ke per label group); here the grouping happens per wave and per layer.
//
// Precision.  FP32 operands are split into two FP16 terms, x = hi + lo with
// hi = f16(x) and lo = f16(x - hi), and the product keeps the three leading
// terms hi.hi + hi.lo + lo.hi with FP32 accumulation (about 2^-21 relative to
// the largest term of a column).  K_L is scaled by a power of two per row so
// that its largest entry lies in [0.5, 1); each element column is scaled by a
// power of two from its own largest entry.  K annihilates rigid translations,
// so the element's node-0 displacement is subtracted first: the operands
// are displacement differences, as in the modal form.
//
// Operand layout (16x16x32 f16): lane l holds A[l % 16][k] and B[k][l % 16]
// for the eight k of its lane group l / 16, and D rows 4 (l / 16) + i, i < 4,
// of column l % 16.  The k index is 4 a + c (node a = ax + 2 ay + 4 az of the
// element, component c, c = 3 a zero pad), so lane group g supplies the
// nodes 2g and 2g + 1, two neighbours along x (one 32-byte LDS read); the A
// table is stored in the same lane order, so any fixed k order the hardware
// uses is common to A and B and leaves the sums unchanged.
//
// Schedule: a workgroup owns a 32 x 7 tile of output nodes and marches along
// z through a chunk of planes.  Each step stages one node plane (float4 per
// node) in LDS, computes the 33 x 8 elements of one layer (four column groups
// of 16 elements per wave; per wave one product pass per distinct row among
// its 64 elements), writes the 8 corner forces of each element to LDS, and
// every output node gathers its four bottom and four top contributions in a
// fixed order (no atomics, deterministic).
#include "vxl_internal.h"

#include <cmath>
#include <vector>

namespace vxl {
namespace {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kMX = 32, kMY = 7;
constexpr int kPX = kMX + 2, kPY = kMY + 2, kPlane = kPX * kPY;  // staged node plane with halo
constexpr int kEX = kMX + 1, kEY = kMY + 1, kEL = kEX * kEY;    // elements of one layer
constexpr int kNodes = kMX * kMY;
constexpr int kNT = (kEL + 63) / 64 * 64;  // 5 waves, one element slot per (wave, column group, column)
constexpr int kWaves = kNT / 64;
static_assert(kPlane <= kNT && kNodes <= kNT, "one staged node and one output node per thread");

struct MfmaArgs {
  const float* u;
  float* out;
  int64_t rhs_stride;
  const void* index;
  const uint4* ktab;  // rows x [tile 2][term 2][lane 64] x 8 halves
  const int* kexp;    // rows: power-of-two exponent of each K_L
  const float* scale;
  int nx, ny, nz;
  int zchunk, tiles_x;
  const int* active;  // kFuse: skipped systems
  double* partial;    // kFuse: p . q per workgroup
};

__device__ __forceinline__ int wrapm(int v, int n) {
  v %= n;
  return v < 0 ? v + n : v;
}

template <typename Idx, bool kScale, bool kFuse>
__global__ void __launch_bounds__(kNT)
mfma_apply_kernel(MfmaArgs a) {
  __shared__ floatx4 sp[2][kPlane];
  __shared__ floatx4 sf[8][kEL];
  __shared__ double red[kWaves];

  const int tid = threadIdx.x;
  const uint3 b = xcd_block3();
  const int tile = int(b.x), chunk = int(b.y), rhs = int(b.z);
  const int x0 = (tile % a.tiles_x) * kMX;
  const int y0 = (tile / a.tiles_x) * kMY;
  const int z0 = chunk * a.zchunk;
  const int z1 = min(z0 + a.zchunk, a.nz);
  if (z0 >= a.nz) return;
  if constexpr (kFuse) {
    if (a.active && !a.active[rhs]) return;
  }
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const int64_t nxy = int64_t(nx) * ny;
  const int64_t n = nxy * nz;
  const float* u = a.u + int64_t(rhs) * a.rhs_stride;
  float* out = a.out + int64_t(rhs) * a.rhs_stride;

  // Plane staging: thread tid copies halo-tile node tid.
  const bool stager = tid < kPlane;
  const int ld_off = wrapm(y0 - 1 + tid / kPX, ny) * nx + wrapm(x0 - 1 + tid % kPX, nx);
  auto load_plane = [&](int z, int buf) {
    if (!stager) return;
    const int64_t node = int64_t(wrapm(z, nz)) * nxy + ld_off;
    floatx4 v;
    v.x = u[node];
    v.y = u[n + node];
    v.z = u[2 * n + node];
    v.w = 0.f;
    sp[buf][tid] = v;
  };

  // Element columns of this lane: e = wave * 64 + cg * 16 + (lane % 16).
  const int lane = tid & 63, wave = tid >> 6, g = lane >> 4, col = lane & 15;
  int e_xy[4], p0[4];
  bool ev[4];
#pragma unroll
  for (int cg = 0; cg < 4; ++cg) {
    const int e = wave * 64 + cg * 16 + col;
    ev[cg] = e < kEL;
    const int ex = e % kEX, ey = e / kEX;
    e_xy[cg] = wrapm(y0 - 1 + ey, ny) * nx + wrapm(x0 - 1 + ex, nx);
    p0[cg] = ev[cg] ? ey * kPX + ex : 0;  // node (0, 0) of the element in a staged plane
  }
  const int gy = g & 1, gz = g >> 1;  // nodes 2g, 2g + 1 of the element: ay = g & 1, az = g >> 1

  auto compute_layer = [&](int z, int bottom, int top) {
    const int64_t zoff = int64_t(wrapm(z, nz)) * nxy;
    int rows[4], ecol[4];
    half8 bhi[4], blo[4];
    int done = 0;
#pragma unroll
    for (int cg = 0; cg < 4; ++cg) {
      rows[cg] = ev[cg] ? load_row<Idx>(a.index, zoff + e_xy[cg]) : -1;
      if (!ev[cg]) done |= 1 << cg;
    }
#pragma unroll
    for (int cg = 0; cg < 4; ++cg) {
      const floatx4 v0 = sp[bottom][p0[cg]];
      const int pp = p0[cg] + gy * kPX;
      const floatx4 va = sp[gz ? top : bottom][pp];
      const floatx4 vb = sp[gz ? top : bottom][pp + 1];
      float d[8] = {va.x - v0.x, va.y - v0.y, va.z - v0.z, 0.f, vb.x - v0.x, vb.y - v0.y, vb.z - v0.z, 0.f};
      float m = 0.f;
#pragma unroll
      for (int j = 0; j < 8; ++j) m = fmaxf(m, fabsf(d[j]));
      m = fmaxf(m, __shfl_xor(m, 16, 64));
      m = fmaxf(m, __shfl_xor(m, 32, 64));
      int ex = 0;
      if (m > 0.f) (void)frexpf(m, &ex);
      ecol[cg] = ex;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const float s = ldexpf(d[j], -ex);
        const _Float16 h = static_cast<_Float16>(s);
        bhi[cg][j] = h;
        blo[cg][j] = static_cast<_Float16>(s - static_cast<float>(h));
      }
    }
    floatx4 D[4][2];
#pragma unroll
    for (int cg = 0; cg < 4; ++cg)
#pragma unroll
      for (int t = 0; t < 2; ++t) D[cg][t] = floatx4{0.f, 0.f, 0.f, 0.f};
    // One product pass per distinct coefficient row among the wave's columns.
    while (true) {
      int cand = -1;
#pragma unroll
      for (int cg = 3; cg >= 0; --cg)
        if (!((done >> cg) & 1)) cand = rows[cg];
      const unsigned long long pend = __ballot(cand >= 0);
      if (!pend) break;
      const int L = __shfl(cand, __ffsll(pend) - 1, 64);
      half8 ahi[2], alo[2];
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        const uint4 h = a.ktab[((int64_t(L) * 2 + t) * 2 + 0) * 64 + lane];
        const uint4 l = a.ktab[((int64_t(L) * 2 + t) * 2 + 1) * 64 + lane];
        ahi[t] = __builtin_bit_cast(half8, h);
        alo[t] = __builtin_bit_cast(half8, l);
      }
      bool mine[4], any[4];
      half8 bh[4], bl[4];
#pragma unroll
      for (int cg = 0; cg < 4; ++cg) {
        mine[cg] = !((done >> cg) & 1) && rows[cg] == L;
        any[cg] = __ballot(mine[cg]) != 0;
        const half8 zero = {};
        bh[cg] = mine[cg] ? bhi[cg] : zero;
        bl[cg] = mine[cg] ? blo[cg] : zero;
        if (mine[cg]) done |= 1 << cg;
      }
      // Term-major order: consecutive products write different accumulators,
      // so each accumulator's chain of three products is spread out.
#pragma unroll
      for (int term = 0; term < 3; ++term)
#pragma unroll
        for (int cg = 0; cg < 4; ++cg) {
          if (!any[cg]) continue;
#pragma unroll
          for (int t = 0; t < 2; ++t)
            D[cg][t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(term == 2 ? alo[t] : ahi[t],
                                                              term == 1 ? bl[cg] : bh[cg], D[cg][t], 0, 0, 0);
        }
    }
    // Corner forces: D rows 16 t + 4 g + i are node a = 4 t + g, component i.
#pragma unroll
    for (int cg = 0; cg < 4; ++cg) {
      if (!ev[cg]) continue;
      const int e = wave * 64 + cg * 16 + col;
      const int sh = ecol[cg] + a.kexp[rows[cg]];
      float s = 1.f;
      if constexpr (kScale) s = a.scale[zoff + e_xy[cg]];
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        floatx4 f;
        f.x = ldexpf(D[cg][t].x, sh) * s;
        f.y = ldexpf(D[cg][t].y, sh) * s;
        f.z = ldexpf(D[cg][t].z, sh) * s;
        f.w = 0.f;
        sf[4 * t + g][e] = f;
      }
    }
  };

  const bool is_node = tid < kNodes;
  const int xl = tid % kMX, yl = tid / kMX;
  const bool node_valid = is_node && (x0 + xl) < nx && (y0 + yl) < ny;
  const int n_xy = (y0 + yl) * nx + (x0 + xl);
  const int np = (yl + 1) * kPX + xl + 1;  // the node in a staged plane
  auto gather = [&](int zbit, float (&acc)[3]) {
#pragma unroll
    for (int dy = 0; dy < 2; ++dy)
#pragma unroll
      for (int dx = 0; dx < 2; ++dx) {
        const int e = (yl + dy) * kEX + xl + dx;
        const int corner = (1 - dx) + 2 * (1 - dy) + 4 * zbit;
        const floatx4 f = sf[corner][e];
        acc[0] += f.x;
        acc[1] += f.y;
        acc[2] += f.z;
      }
  };

  float acc[3] = {0.f, 0.f, 0.f};
  float pq = 0.f;
  load_plane(z0 - 1, 0);
  load_plane(z0, 1);
  __syncthreads();
  compute_layer(z0 - 1, 0, 1);
  __syncthreads();
  if (is_node) gather(1, acc);
  for (int k = 0; z0 + k < z1; ++k) {
    const int z = z0 + k;
    load_plane(z + 1, k & 1);
    __syncthreads();
    floatx4 pn = {0.f, 0.f, 0.f, 0.f};
    if constexpr (kFuse) {
      if (node_valid) pn = sp[(k + 1) & 1][np];  // plane z: the bottom of layer z
    }
    compute_layer(z, (k + 1) & 1, k & 1);
    __syncthreads();
    if (is_node) {
      gather(0, acc);
      if (node_valid) {
        const int64_t node = int64_t(z) * nxy + n_xy;
        out[node] = acc[0];
        out[n + node] = acc[1];
        out[2 * n + node] = acc[2];
        if constexpr (kFuse) pq = fmaf(pn.z, acc[2], fmaf(pn.y, acc[1], fmaf(pn.x, acc[0], pq)));
      }
      acc[0] = acc[1] = acc[2] = 0.f;
      gather(1, acc);
    }
  }
  if constexpr (kFuse) {
    double v = double(pq);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) red[wave] = v;
    __syncthreads();
    if (tid == 0) {
      double s = 0.0;
#pragma unroll
      for (int w = 0; w < kWaves; ++w) s += red[w];
      a.partial[(int64_t(rhs) * gridDim.y + chunk) * gridDim.x + tile] = s;
    }
  }
}

// Dense element stiffness of the fully integrated HEX8 (2x2x2 Gauss rule)
// for the Voigt stiffness c21 (PAIRS order) and voxel size h, in the DOF
// order 3 a + c (node a = ax + 2 ay + 4 az): K = sum_gp B^T C B det J.
void dense_element_matrix(const double* c21, const double h[3], double (&K)[24][24]) {
  double C[6][6];
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) C[i][j] = c21[pidx(i, j)];
  for (int r = 0; r < 24; ++r)
    for (int s = 0; s < 24; ++s) K[r][s] = 0.0;
  const double gp = 1.0 / std::sqrt(3.0);
  const double detj = h[0] * h[1] * h[2] / 8.0;
  for (int q = 0; q < 8; ++q) {
    const double xi[3] = {(q & 1) ? gp : -gp, (q & 2) ? gp : -gp, (q & 4) ? gp : -gp};
    double B[6][24] = {};
    for (int na = 0; na < 8; ++na) {
      const double s[3] = {(na & 1) ? 1.0 : -1.0, (na & 2) ? 1.0 : -1.0, (na & 4) ? 1.0 : -1.0};
      double dn[3];
      for (int d = 0; d < 3; ++d) {
        double p = s[d] / 8.0;
        for (int o = 0; o < 3; ++o)
          if (o != d) p *= 1.0 + s[o] * xi[o];
        dn[d] = p * 2.0 / h[d];
      }
      const int c = 3 * na;
      B[0][c] = dn[0];
      B[1][c + 1] = dn[1];
      B[2][c + 2] = dn[2];
      B[3][c + 1] = dn[2];
      B[3][c + 2] = dn[1];
      B[4][c] = dn[2];
      B[4][c + 2] = dn[0];
      B[5][c] = dn[1];
      B[5][c + 1] = dn[0];
    }
    double CB[6][24];
    for (int i = 0; i < 6; ++i)
      for (int s = 0; s < 24; ++s) {
        double acc = 0.0;
        for (int j = 0; j < 6; ++j) acc += C[i][j] * B[j][s];
        CB[i][s] = acc;
      }
    for (int r = 0; r < 24; ++r)
      for (int s = 0; s < 24; ++s) {
        double acc = 0.0;
        for (int i = 0; i < 6; ++i) acc += B[i][r] * CB[i][s];
        K[r][s] += acc * detj;
      }
  }
}

}  // namespace

int mfma_max_rows() { return 4096; }

// Dense K_L of every coefficient row, scaled by 2^-E_L and split into FP16
// hi/lo A operands in lane order; uploaded once per cell.
int mfma_prepare(vxl_cell* cell) {
  if (cell->mfma_table) return VXL_OK;
  if (cell->element != VXL_ELEMENT_FI) return fail("the matrix-core stencil implements the FI element only");
  if (cell->rows > mfma_max_rows())
    return fail("the matrix-core stencil groups elements by coefficient row: at most 4096 rows");
  const int rows = cell->rows;
  std::vector<double> c64(size_t(rows) * kRowWidth);
  VXL_HIP(hipMemcpyAsync(c64.data(), cell->mat.c64, c64.size() * sizeof(double), hipMemcpyDeviceToHost,
                         cell->stream));
  VXL_HIP(hipStreamSynchronize(cell->stream));
  std::vector<_Float16> tab(size_t(rows) * 2 * 2 * 64 * 8);
  std::vector<int> kexp(rows);
  for (int r = 0; r < rows; ++r) {
    double K[24][24];
    dense_element_matrix(&c64[size_t(r) * kRowWidth], cell->h, K);
    double m = 0.0;
    for (int i = 0; i < 24; ++i)
      for (int j = 0; j < 24; ++j) m = std::fmax(m, std::fabs(K[i][j]));
    int ex = 0;
    if (m > 0.0) (void)std::frexp(m, &ex);
    kexp[r] = ex;
    for (int t = 0; t < 2; ++t)
      for (int lane = 0; lane < 64; ++lane)
        for (int j = 0; j < 8; ++j) {
          const int row = 16 * t + (lane & 15), k = 8 * (lane >> 4) + j;
          const int ar = row >> 2, cr = row & 3, ak = k >> 2, ck = k & 3;
          const double v = (cr < 3 && ck < 3) ? K[3 * ar + cr][3 * ak + ck] : 0.0;
          const float s = static_cast<float>(std::ldexp(v, -ex));
          const _Float16 hi = static_cast<_Float16>(s);
          const _Float16 lo = static_cast<_Float16>(s - static_cast<float>(hi));
          tab[((size_t(r) * 2 + t) * 2 + 0) * 512 + lane * 8 + j] = hi;
          tab[((size_t(r) * 2 + t) * 2 + 1) * 512 + lane * 8 + j] = lo;
        }
  }
  void* dt = nullptr;
  void* de = nullptr;
  VXL_HIP(dev_malloc(&dt, tab.size() * sizeof(_Float16)));
  if (dev_malloc(&de, kexp.size() * sizeof(int)) != hipSuccess) {
    dev_free(dt);
    return fail("HIP error: device allocation of the matrix-core exponents failed");
  }
  hipError_t e1 = hipMemcpyAsync(dt, tab.data(), tab.size() * sizeof(_Float16), hipMemcpyHostToDevice, cell->stream);
  hipError_t e2 = hipMemcpyAsync(de, kexp.data(), kexp.size() * sizeof(int), hipMemcpyHostToDevice, cell->stream);
  hipError_t e3 = hipStreamSynchronize(cell->stream);
  if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) {
    dev_free(dt);
    dev_free(de);
    return fail("HIP error: upload of the matrix-core element matrices failed");
  }
  cell->mfma_table = dt;
  cell->mfma_exp = static_cast<int*>(de);
  return VXL_OK;
}

void mfma_release(vxl_cell* cell) {
  if (cell->mfma_table) dev_free(cell->mfma_table);
  if (cell->mfma_exp) dev_free(cell->mfma_exp);
  cell->mfma_table = nullptr;
  cell->mfma_exp = nullptr;
  cell->mfma = 0;
}

void mfma_grid(const vxl_cell* cell, int nrhs, int& tiles_x, int& tiles, int& chunks, int& zchunk) {
  const int nx = cell->grid.nx, ny = cell->grid.ny, nz = cell->grid.nz;
  tiles_x = (nx + kMX - 1) / kMX;
  tiles = tiles_x * ((ny + kMY - 1) / kMY);
  const int zmin = std::max(kMinChunkPlanes, cell->zmin[0]);
  chunks = (4096 + tiles * nrhs - 1) / (tiles * nrhs);
  chunks = std::max(1, std::min(chunks, nz / zmin));
  zchunk = (nz + chunks - 1) / chunks;
  chunks = (nz + zchunk - 1) / zchunk;
}

int mfma_blocks(const vxl_cell* cell, int nrhs) {
  int tx, tiles, chunks, zc;
  mfma_grid(cell, nrhs, tx, tiles, chunks, zc);
  return tiles * chunks;
}

int launch_mfma_apply(const vxl_cell* cell, int nrhs, const float* u, float* out, const int* active,
                      double* partial, int* blocks, hipStream_t s) {
  if (!cell->mfma_table) return fail("matrix-core stencil not prepared");
  MfmaArgs a{};
  a.u = u;
  a.out = out;
  a.rhs_stride = 3 * cell->grid.n;
  a.index = cell->mat.index;
  a.ktab = static_cast<const uint4*>(cell->mfma_table);
  a.kexp = cell->mfma_exp;
  a.scale = cell->mat.scale;
  a.nx = cell->grid.nx;
  a.ny = cell->grid.ny;
  a.nz = cell->grid.nz;
  a.active = active;
  a.partial = partial;
  int tiles = 0, chunks = 0;
  mfma_grid(cell, nrhs, a.tiles_x, tiles, chunks, a.zchunk);
  if (blocks) *blocks = tiles * chunks;
  const dim3 grid(tiles, chunks, nrhs);
  const bool sc = cell->mat.scale != nullptr, fuse = partial != nullptr;
#define VXL_MFMA(IDX)                                                                                         \
  do {                                                                                                        \
    if (sc && fuse)                                                                                           \
      hipLaunchKernelGGL((mfma_apply_kernel<IDX, true, true>), grid, dim3(kNT), 0, s, a);                     \
    else if (sc)                                                                                              \
      hipLaunchKernelGGL((mfma_apply_kernel<IDX, true, false>), grid, dim3(kNT), 0, s, a);                    \
    else if (fuse)                                                                                            \
      hipLaunchKernelGGL((mfma_apply_kernel<IDX, false, true>), grid, dim3(kNT), 0, s, a);                    \
    else                                                                                                      \
      hipLaunchKernelGGL((mfma_apply_kernel<IDX, false, false>), grid, dim3(kNT), 0, s, a);                   \
  } while (0)
  switch (cell->mat.index_kind) {
    case VXL_INDEX_U8:
      VXL_MFMA(uint8_t);
      break;
    case VXL_INDEX_U16:
      VXL_MFMA(uint16_t);
      break;
    case VXL_INDEX_I32:
      VXL_MFMA(int32_t);
      break;
    default:
      return fail("unknown material index kind");
  }
#undef VXL_MFMA
  VXL_HIP(hipGetLastError());
  return VXL_OK;
}

}  // namespace vxl
