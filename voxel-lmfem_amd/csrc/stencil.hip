// Matrix-free periodic HEX8 (fully integrated) stiffness action, RHS and mean
// stress for gfx950.
//
// Element action (exact 2x2x2 Gauss integration of a trilinear element with a
// constant 6x6 stiffness C).  With node a = ax + 2ay + 4az and sign
// s_d(a) = +-1, the Walsh-Hadamard coefficients q[m] = sum_a u_a prod_{d in m}
// s_d(a) give u(xi) = 1/8 sum_m q[m] prod_{d in m} xi_d.  The displacement
// gradient is a polynomial in {1, xi, eta, zeta, eta*zeta, xi*zeta, xi*eta};
// these seven monomials are orthogonal under the Gauss rule with weights
// {1, 1/3, 1/3, 1/3, 1/9, 1/9, 1/9}, so the element energy splits into seven
// independent strain modes and K_e = H^T G^T W C G H with H the 8-point
// Hadamard butterfly network.  This is the same element as the reference's
// modal form (_core/hybrid_operator/fi_modal.cuh:2-255); the code below is
// written from the derivation above, with the weights folded symmetrically
// into the geometric factors.
//
// Stencil schedule: a workgroup owns a TX x TY tile of output nodes and marches
// along z through a chunk of planes.  Each step stages one node plane (with a
// one-node halo) in LDS, computes the (TX+1)(TY+1) elements of one element
// layer (halo elements are recomputed instead of exchanged), writes their 24
// nodal forces to LDS, and every output node gathers its four bottom-layer and
// four top-layer contributions in a fixed order.  No atomics: the result is
// bitwise deterministic.
#include "vxl_internal.h"

#include <cstdlib>

namespace vxl {

ModalFactors modal_factors(const double h[3]) {
  ModalFactors f;
  const double V = h[0] * h[1] * h[2];
  const double w[3] = {1.0, 1.0 / 3.0, 1.0 / 9.0};
  for (int k = 0; k < 3; ++k)
    for (int a = 0; a < 3; ++a) f.g[k][a] = std::sqrt(w[k] * V) / (4.0 * h[a]);
  return f;
}

namespace {

template <typename T>
struct G9 {
  T g[3][3];  // [class][axis]
};

// s = C e restricted to the strain components in MASK (bit k = Voigt k).
template <int MASK, typename T>
__device__ __forceinline__ void cmul(const T* c, const T (&e)[6], T (&s)[6]) {
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    if (!((MASK >> i) & 1)) continue;
    T acc = T(0);
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      if (!((MASK >> j) & 1)) continue;
      acc = fma(c[pidx(i, j)], e[j], acc);
    }
    s[i] = acc;
  }
}

// One strain mode.  MX/MY/MZ name the Hadamard mode whose coefficient vector
// forms gradient column x/y/z (-1: column absent); CLS selects the weight
// class.  Accumulates the back-projected modal forces into P.
// kIso (cubic voxels, gx = gy = gz within a class): the gradient scale is
// moved onto the back-projection, P += g^2 C e(q), and g holds g^2.
template <int MX, int MY, int MZ, int CLS, bool kIso, typename T>
__device__ __forceinline__ void strain_mode(const T* c, const G9<T>& g, const T (&q)[3][8],
                                            T (&P)[3][8]) {
  constexpr bool HX = MX >= 0, HY = MY >= 0, HZ = MZ >= 0;
  constexpr int MASK = (HX ? 1 : 0) | (HY ? 2 : 0) | (HZ ? 4 : 0) | ((HY || HZ) ? 8 : 0) |
                       ((HX || HZ) ? 16 : 0) | ((HX || HY) ? 32 : 0);
  const T bx = g.g[CLS][0], by = g.g[CLS][1], bz = g.g[CLS][2];
  const T gx = kIso ? T(1) : bx, gy = kIso ? T(1) : by, gz = kIso ? T(1) : bz;
  // Gradient columns col_j[i] = d u_i / d x_j.
  T cx[3], cy[3], cz[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    cx[i] = HX ? (kIso ? q[i][HX ? MX : 0] : gx * q[i][HX ? MX : 0]) : T(0);
    cy[i] = HY ? (kIso ? q[i][HY ? MY : 0] : gy * q[i][HY ? MY : 0]) : T(0);
    cz[i] = HZ ? (kIso ? q[i][HZ ? MZ : 0] : gz * q[i][HZ ? MZ : 0]) : T(0);
  }
  T e[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
  if constexpr (HX) e[0] = cx[0];
  if constexpr (HY) e[1] = cy[1];
  if constexpr (HZ) e[2] = cz[2];
  if constexpr (HY && HZ) e[3] = cz[1] + cy[2];
  else if constexpr (HY) e[3] = cy[2];
  else if constexpr (HZ) e[3] = cz[1];
  if constexpr (HX && HZ) e[4] = cz[0] + cx[2];
  else if constexpr (HX) e[4] = cx[2];
  else if constexpr (HZ) e[4] = cz[0];
  if constexpr (HX && HY) e[5] = cy[0] + cx[1];
  else if constexpr (HX) e[5] = cx[1];
  else if constexpr (HY) e[5] = cy[0];
  T s[6];
  if constexpr (kIso && CLS == 0) {
    // Constant mode of a cubic voxel: one factor b0 = g^2 on every axis.
    // Scaling the six strains by it (6 products) instead of the nine
    // back-projected stresses makes the back-projection plain assignments;
    // only this mode writes the slots MX = 1, MY = 2, MZ = 4, so P starts at
    // zero there.
    static_assert(MX == 1 && MY == 2 && MZ == 4, "the constant mode");
    const T b0 = bx;
#pragma unroll
    for (int i = 0; i < 6; ++i) e[i] *= b0;
    cmul<MASK>(c, e, s);
    P[0][MX] = s[0];
    P[1][MX] = s[5];
    P[2][MX] = s[4];
    P[0][MY] = s[5];
    P[1][MY] = s[1];
    P[2][MY] = s[3];
    P[0][MZ] = s[4];
    P[1][MZ] = s[3];
    P[2][MZ] = s[2];
    return;
  }
  cmul<MASK>(c, e, s);
  // Back-projection: P[m_j] += g_j * (sigma_0j, sigma_1j, sigma_2j).
  if constexpr (HX) {
    P[0][MX] = fma(bx, s[0], P[0][MX]);
    P[1][MX] = fma(bx, s[5], P[1][MX]);
    P[2][MX] = fma(bx, s[4], P[2][MX]);
  }
  if constexpr (HY) {
    P[0][MY] = fma(by, s[5], P[0][MY]);
    P[1][MY] = fma(by, s[1], P[1][MY]);
    P[2][MY] = fma(by, s[3], P[2][MY]);
  }
  if constexpr (HZ) {
    P[0][MZ] = fma(bz, s[4], P[0][MZ]);
    P[1][MZ] = fma(bz, s[3], P[1][MZ]);
    P[2][MZ] = fma(bz, s[2], P[2][MZ]);
  }
}

template <typename T>
__device__ __forceinline__ void hadamard_forward(T (&v)[8]) {
#pragma unroll
  for (int k = 1; k < 8; k <<= 1) {
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      if (a & k) continue;
      const T lo = v[a], hi = v[a + k];
      v[a] = hi + lo;
      v[a + k] = hi - lo;
    }
  }
}

template <typename T>
__device__ __forceinline__ void hadamard_transpose(T (&v)[8]) {
#pragma unroll
  for (int k = 1; k < 8; k <<= 1) {
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      if (a & k) continue;
      const T s = v[a], d = v[a + k];
      v[a] = s - d;
      v[a + k] = s + d;
    }
  }
}

// Modal element action: Hadamard coefficients q[c][m] in, modal forces P[c][m]
// out (nodal forces f = H^T P).  c: 21 packed stiffness.
template <bool kIso = false, typename T>
__device__ __forceinline__ void element_modal(const T (&q)[3][8], const T* c, const G9<T>& g,
                                              T (&P)[3][8]) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int m = 0; m < 8; ++m) P[i][m] = T(0);
  // Mode numbering: bit 0 = x, bit 1 = y, bit 2 = z.
  strain_mode<1, 2, 4, 0, kIso>(c, g, q, P);    // constant
  strain_mode<-1, 3, 5, 1, kIso>(c, g, q, P);   // xi
  strain_mode<3, -1, 6, 1, kIso>(c, g, q, P);   // eta
  strain_mode<5, 6, -1, 1, kIso>(c, g, q, P);   // zeta
  strain_mode<7, -1, -1, 2, kIso>(c, g, q, P);  // eta*zeta
  strain_mode<-1, 7, -1, 2, kIso>(c, g, q, P);  // xi*zeta
  strain_mode<-1, -1, 7, 2, kIso>(c, g, q, P);  // xi*eta
}

// C3D8R-HG element (femlib/c3d8r.py:62-78): one-point (centroid) integration,
// V B0^T C B0, which is exactly the constant strain mode above, plus the
// hourglass penalty s(C) H H^T.  H holds, per displacement component, the
// corner patterns xy, yz, zx and xyz over sqrt(8); in Hadamard coefficients
// the penalty is P[c][m] += s / 8 q[c][m] for the modes m = 3, 6, 5, 7.
// hg = coef V / hbar^2 / 8; s / 8 = hg * max((I1 + 7 I2) / 60, 1e-12).
template <bool kIso = false, typename T>
__device__ __forceinline__ void element_hg(const T (&q)[3][8], const T* c, const G9<T>& g, T hg, T (&P)[3][8]) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int m = 0; m < 8; ++m) P[i][m] = T(0);
  strain_mode<1, 2, 4, 0, kIso>(c, g, q, P);
  const T normal = c[pidx(0, 0)] + c[pidx(1, 1)] + c[pidx(2, 2)];
  const T i1 = normal + T(2) * (c[pidx(0, 1)] + c[pidx(0, 2)] + c[pidx(1, 2)]);
  const T i2 = normal + T(2) * (c[pidx(3, 3)] + c[pidx(4, 4)] + c[pidx(5, 5)]);
  const T k = hg * fmax((i1 + T(7) * i2) * T(1.0 / 60.0), T(1e-12));
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    P[i][3] = k * q[i][3];
    P[i][5] = k * q[i][5];
    P[i][6] = k * q[i][6];
    P[i][7] = k * q[i][7];
  }
}

// v[c][a]: nodal displacements in, nodal forces out.
template <typename T>
__device__ __forceinline__ void element_action(T (&v)[3][8], const T* c, const G9<T>& g) {
#pragma unroll
  for (int i = 0; i < 3; ++i) hadamard_forward(v[i]);
  T P[3][8];
  element_modal(v, c, g, P);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    hadamard_transpose(P[i]);
#pragma unroll
    for (int a = 0; a < 8; ++a) v[i][a] = P[i][a];
  }
}

// Two-point Hadamard of one face (4 nodes a2 = ax + 2 ay) and its transpose.
template <typename T>
__device__ __forceinline__ void face_forward(T (&v)[4]) {
  T a = v[1] + v[0], b = v[1] - v[0], c = v[3] + v[2], d = v[3] - v[2];
  v[0] = c + a;
  v[2] = c - a;
  v[1] = d + b;
  v[3] = d - b;
}
template <typename T>
__device__ __forceinline__ void face_transpose(T (&v)[4]) {
  T a = v[0] - v[2], c = v[0] + v[2], b = v[1] - v[3], d = v[1] + v[3];
  v[0] = a - b;
  v[1] = a + b;
  v[2] = c - d;
  v[3] = c + d;
}

template <typename T>
struct ApplyArgs {
  const T* u;
  T* out;
  int64_t rhs_stride;  // elements between right-hand sides
  const void* index;
  const T* table;      // rows x 24
  const float* scale;  // optional
  const float4* frames;  // kFrames: unit quaternion (w, x, y, z) of every voxel's material frame
  // Stacked cells: right-hand side r reads index block r / group_rhs, which
  // starts group_stride entries after the previous one (0 for one cell).
  int64_t group_stride;
  int group_rhs;
  int nx, ny, nz;
  int zchunk;
  int tiles_x;
  G9<T> g;
  T hg;  // C3D8R-HG: coef * V / hbar^2 / 8 (hourglass penalty per unit material scale)
  // kResid: r = f - K u written as r32 (if set) or r64, ||r||^2 per workgroup.
  const double* f;
  float* r32;
  double* r64;
  // kFuse: systems with active[r] == 0 are skipped and u . (K u) of each
  // workgroup is written to partial[(r * chunks + chunk) * tiles + tile].
  const int* active;
  double* partial;
};

template <typename T>
void set_groups(ApplyArgs<T>& a, const vxl_cell* cell, int nrhs) {
  const int groups = cell->groups > 1 ? cell->groups : 1;
  a.group_stride = groups > 1 ? cell->grid.n : 0;
  a.group_rhs = groups > 1 ? std::max(1, nrhs / groups) : std::max(1, nrhs);
}

__device__ __forceinline__ int wrapmod(int v, int n) {
  v %= n;
  return v < 0 ? v + n : v;
}

// Workgroups are dealt to the 8 XCDs round-robin by linear id, and each XCD
// has its own L2.  Remap ids so every XCD owns a contiguous run of (tile,
// chunk, rhs) in tile-fastest order: x-neighbouring tiles, whose halo rows
// share cache lines, then run on the same XCD at about the same time.
struct BlockId {
  int tile, chunk, rhs;
};
__device__ __forceinline__ BlockId xcd_block() {
  const unsigned gx = gridDim.x, gy = gridDim.y;
  const unsigned total = gx * gy * gridDim.z;
  unsigned lin = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
  const unsigned full = total & ~7u;
  if (lin < full) lin = (lin & 7u) * (full >> 3) + (lin >> 3);
  BlockId b;
  b.tile = int(lin % gx);
  b.chunk = int((lin / gx) % gy);
  b.rhs = int(lin / (gx * gy));
  return b;
}

// Output-node tile of one workgroup: 32 x TY, TY from a short list per
// precision (the first entry is the default; the autotuner picks per cell).
// FP32 32 x 14 nodes = 33 x 15 elements = 8 waves, so two workgroups (16
// waves, 4 per SIMD) share a CU at up to 128 VGPRs; 32 x 16 (9 waves) fits
// only one at the 96 VGPRs that pairing would need, and ran 1.55x slower.
// Smaller tiles give small cells more workgroups.  FP64 default 32 x 6
// (4 waves): measured 2.90 ms against 3.66 ms for 32 x 8 per six-RHS launch
// at 256^3 (schedule 3).
constexpr int kTilesF32[3] = {14, 8, 6};
constexpr int kTilesF64[3] = {6, 8, 4};

template <int TX, int TY>
struct Tile {
  static constexpr int PX = TX + 2, PY = TY + 2, PLANE = PX * PY;
  static constexpr int EX = TX + 1, EY = TY + 1, EL = EX * EY;
  static constexpr int NODES = TX * TY;
  static constexpr int NT0 = EL > NODES ? EL : NODES;
  static constexpr int NT = (NT0 + 63) / 64 * 64;
  static constexpr int ITEMS = (PLANE + NT - 1) / NT;
};

template <typename T, typename Idx, bool kScale, int TX, int TY>
__global__ void __launch_bounds__((Tile<TX, TY>::NT))
apply_kernel(ApplyArgs<T> a) {
  using TL = Tile<TX, TY>;
  __shared__ T su[2][3][TL::PLANE];
  __shared__ T sf[24][TL::EL];

  const int tid = threadIdx.x;
  const BlockId bid = xcd_block();
  const int tile = bid.tile;
  const int x0 = (tile % a.tiles_x) * TX;
  const int y0 = (tile / a.tiles_x) * TY;
  const int z0 = bid.chunk * a.zchunk;
  const int z1 = min(z0 + a.zchunk, a.nz);
  if (z0 >= a.nz) return;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const int64_t nxy = int64_t(nx) * ny;
  const int64_t n = nxy * nz;
  const T* u = a.u + int64_t(bid.rhs) * a.rhs_stride;
  T* out = a.out + int64_t(bid.rhs) * a.rhs_stride;

  // Plane staging: each thread copies up to ITEMS halo-tile nodes.
  int ld_lds[TL::ITEMS];
  int ld_off[TL::ITEMS];  // offsets inside one z plane (nx * ny < 2^31)
#pragma unroll
  for (int k = 0; k < TL::ITEMS; ++k) {
    const int idx = tid + k * TL::NT;
    ld_lds[k] = idx < TL::PLANE ? idx : -1;
    const int px = idx % TL::PX, py = idx / TL::PX;
    ld_off[k] = wrapmod(y0 - 1 + py, ny) * nx + wrapmod(x0 - 1 + px, nx);
  }
  auto load_plane = [&](int z, int buf) {
    const int64_t base = int64_t(wrapmod(z, nz)) * nxy;
#pragma unroll
    for (int k = 0; k < TL::ITEMS; ++k) {
      if (ld_lds[k] < 0) continue;
      const int64_t node = base + ld_off[k];
#pragma unroll
      for (int c = 0; c < 3; ++c) su[buf][c][ld_lds[k]] = u[c * n + node];
    }
  };

  // Element ownership.
  const bool is_elem = tid < TL::EL;
  const int exl = tid % TL::EX, eyl = tid / TL::EX;
  const int e_xy = wrapmod(y0 - 1 + eyl, ny) * nx + wrapmod(x0 - 1 + exl, nx);
  auto compute_layer = [&](int z, int bottom, int top) {
    if (!is_elem) return;
    const int64_t voxel = int64_t(wrapmod(z, nz)) * nxy + e_xy;
    const int row = load_row<Idx>(a.index, int64_t(bid.rhs / a.group_rhs) * a.group_stride + voxel);
    const T* crow = a.table + int64_t(row) * kRowWidth;
    T c[21];
#pragma unroll
    for (int k = 0; k < 21; ++k) c[k] = crow[k];
    if constexpr (kScale) {
      const T s = T(a.scale[voxel]);
#pragma unroll
      for (int k = 0; k < 21; ++k) c[k] *= s;
    }
    T v[3][8];
#pragma unroll
    for (int cc = 0; cc < 3; ++cc)
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int ax = q & 1, ay = (q >> 1) & 1, az = q >> 2;
        v[cc][q] = su[az ? top : bottom][cc][(eyl + ay) * TL::PX + exl + ax];
      }
    element_action(v, c, a.g);
#pragma unroll
    for (int q = 0; q < 8; ++q)
#pragma unroll
      for (int cc = 0; cc < 3; ++cc) sf[q * 3 + cc][tid] = v[cc][q];
  };

  // Node ownership.
  const bool is_node = tid < TL::NODES;
  const int xl = tid % TX, yl = tid / TX;
  const bool node_valid = is_node && (x0 + xl) < nx && (y0 + yl) < ny;
  const int n_xy = (y0 + yl) * nx + (x0 + xl);
  // The four elements (ex, ey) in {xl, xl+1} x {yl, yl+1}; corner index
  // (ax, ay) = (xl+1-ex, yl+1-ey).
  auto gather = [&](int zbit, T (&acc)[3]) {
#pragma unroll
    for (int dy = 0; dy < 2; ++dy)
#pragma unroll
      for (int dx = 0; dx < 2; ++dx) {
        const int e = (yl + dy) * TL::EX + xl + dx;
        const int corner = (1 - dx) + 2 * (1 - dy) + 4 * zbit;
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) acc[cc] += sf[corner * 3 + cc][e];
      }
  };

  T acc[3] = {T(0), T(0), T(0)};
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
    compute_layer(z, (k + 1) & 1, k & 1);
    __syncthreads();
    if (is_node) {
      gather(0, acc);
      if (node_valid) {
        const int64_t node = int64_t(z) * nxy + n_xy;
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) out[cc * n + node] = acc[cc];
      }
      acc[0] = acc[1] = acc[2] = T(0);
      gather(1, acc);
    }
  }
}

// ---- material frames rebuilt in the stencil ---------------------------------
// With per-voxel frames the FP32 stencil reads a phase row (the phase's C21 in
// its local frame; slot 21 holds the kind: 0 global frame, 1 local frame, 2
// local-frame cubic, 3 local-frame orthotropic) and the voxel's unit
// quaternion q = (w, x, y, z), 16 B, and rotates the row in registers: kind 1
// by the Bond stress matrix, C' = T C T^T (the reference's _rotate_packed,
// material_field.py:105-114, in engineering Voigt order 11, 22, 33, 23, 13,
// 12); kind 2 by the cubic closed form C' = C12 d(x)d + C44 (shear part) +
// a sum_m p_m p_m^T, where p_m holds the Voigt pair products of local axis m
// (column m of R) and a = C11 - C12 - 2 C44, which equals T C T^T for a cubic
// C; kind 3 by T C T^T restricted to the nine orthotropic constants: the
// columns of T are p_0, p_1, p_2 (normal) and q_s = (r_a r_b^T + r_b r_a^T)
// in Voigt form for the shear pairs s = (1, 2), (0, 2), (0, 1), so
// C' = sum_{m,n<3} C_mn p_m p_n^T + sum_s G_s q_s q_s^T.
__device__ __forceinline__ void frame_matrix(float4 q, float (&R)[3][3]) {
  const float w = q.x, x = q.y, y = q.z, z = q.w;
  R[0][0] = 1.f - 2.f * (y * y + z * z);
  R[0][1] = 2.f * (x * y - w * z);
  R[0][2] = 2.f * (x * z + w * y);
  R[1][0] = 2.f * (x * y + w * z);
  R[1][1] = 1.f - 2.f * (x * x + z * z);
  R[1][2] = 2.f * (y * z - w * x);
  R[2][0] = 2.f * (x * z - w * y);
  R[2][1] = 2.f * (y * z + w * x);
  R[2][2] = 1.f - 2.f * (x * x + y * y);
}

// Row i of the Bond stress matrix of R; shear rows/columns pair (1, 2),
// (0, 2), (0, 1).
__device__ __forceinline__ void bond_row(const float (&R)[3][3], int i, float (&t)[6]) {
  constexpr int pi[3] = {1, 0, 0}, pj[3] = {2, 2, 1};
  if (i < 3) {
#pragma unroll
    for (int b = 0; b < 3; ++b) t[b] = R[i][b] * R[i][b];
#pragma unroll
    for (int k = 0; k < 3; ++k) t[3 + k] = 2.f * R[i][pi[k]] * R[i][pj[k]];
  } else {
    const int m = i - 3;
#pragma unroll
    for (int b = 0; b < 3; ++b) t[b] = R[pi[m]][b] * R[pj[m]][b];
#pragma unroll
    for (int k = 0; k < 3; ++k) t[3 + k] = R[pi[m]][pi[k]] * R[pj[m]][pj[k]] + R[pi[m]][pj[k]] * R[pj[m]][pi[k]];
  }
}

// kForms: the kinds compiled in, 1 every kind, 2 cubic only, 3 cubic and
// orthotropic (the branches a cell does not need are left out, which keeps
// the register count down).
template <int kForms>
__device__ __forceinline__ void rotate_row(float (&c)[21], float4 q, int kind) {
  float R[3][3];
  frame_matrix(q, R);
  if (kForms == 2 || kind == 2) {
    const float c12 = c[pidx(0, 1)], c44 = c[pidx(3, 3)];
    const float an = c[pidx(0, 0)] - c12 - 2.f * c44;
    float p[3][6];
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      p[m][0] = R[0][m] * R[0][m];
      p[m][1] = R[1][m] * R[1][m];
      p[m][2] = R[2][m] * R[2][m];
      p[m][3] = R[1][m] * R[2][m];
      p[m][4] = R[0][m] * R[2][m];
      p[m][5] = R[0][m] * R[1][m];
    }
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
      for (int j = i; j < 6; ++j) {
        float v = (i < 3 && j < 3) ? c12 : 0.f;
        if (i == j) v += i < 3 ? 2.f * c44 : c44;
        const float sum = p[0][i] * p[0][j] + p[1][i] * p[1][j] + p[2][i] * p[2][j];
        c[pidx(i, j)] = fmaf(an, sum, v);
      }
    return;
  }
  if constexpr (kForms == 2) return;
  if (kForms == 3 || kind == 3) {
    constexpr int sa[3] = {1, 0, 0}, sb[3] = {2, 2, 1};  // shear pair s = (sa, sb)
    constexpr int vi[6] = {0, 1, 2, 1, 0, 0}, vj[6] = {0, 1, 2, 2, 2, 1};  // Voigt I = (vi, vj)
    float p[3][6], qs[3][6];
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
      for (int I = 0; I < 6; ++I) p[m][I] = R[vi[I]][m] * R[vj[I]][m];
#pragma unroll
    for (int s = 0; s < 3; ++s)
#pragma unroll
      for (int I = 0; I < 6; ++I)
        qs[s][I] = fmaf(R[vi[I]][sa[s]], R[vj[I]][sb[s]], R[vi[I]][sb[s]] * R[vj[I]][sa[s]]);
    float nrm[3][3], g[3];
#pragma unroll
    for (int m = 0; m < 3; ++m) {
#pragma unroll
      for (int n = 0; n < 3; ++n) nrm[m][n] = c[pidx(m, n)];
      g[m] = c[pidx(3 + m, 3 + m)];
    }
#pragma unroll
    for (int I = 0; I < 6; ++I) {
      float u[3], v[3];
#pragma unroll
      for (int n = 0; n < 3; ++n) {
        u[n] = fmaf(nrm[2][n], p[2][I], fmaf(nrm[1][n], p[1][I], nrm[0][n] * p[0][I]));
        v[n] = g[n] * qs[n][I];
      }
#pragma unroll
      for (int J = I; J < 6; ++J) {
        float acc = u[0] * p[0][J];
        acc = fmaf(u[1], p[1][J], acc);
        acc = fmaf(u[2], p[2][J], acc);
        acc = fmaf(v[0], qs[0][J], acc);
        acc = fmaf(v[1], qs[1][J], acc);
        c[pidx(I, J)] = fmaf(v[2], qs[2][J], acc);
      }
    }
    return;
  }
  if constexpr (kForms == 3) return;
  float out[21];
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    float ti[6], w[6];
    bond_row(R, i, ti);
#pragma unroll
    for (int l = 0; l < 6; ++l) {
      float acc = 0.f;
#pragma unroll
      for (int k = 0; k < 6; ++k) acc = fmaf(ti[k], c[pidx(k, l)], acc);
      w[l] = acc;
    }
#pragma unroll
    for (int j = i; j < 6; ++j) {
      float tj[6];
      bond_row(R, j, tj);
      float acc = 0.f;
#pragma unroll
      for (int l = 0; l < 6; ++l) acc = fmaf(w[l], tj[l], acc);
      out[pidx(i, j)] = acc;
    }
  }
#pragma unroll
  for (int k = 0; k < 21; ++k) c[k] = out[k];
}

// Stencil v2: element columns persist across the z march.  A column keeps the
// face (xy) Hadamard modes of its bottom node face and the force its previous
// layer put on that face in registers, so each layer transforms one new face
// (4 nodes), runs the modal element action, and exchanges only the 4 corner
// forces of one face through LDS.  kPF (schedule 3): the plane after next is
// prefetched into registers and the staged planes and forces are double
// buffered by layer parity, one barrier per layer; otherwise one staged plane
// and two barriers per layer.  Same arithmetic as the full 8-point
// transform, regrouped.
// kIso: cubic voxels, a.g holds the squared modal factors (strain_mode).
// kHG: the C3D8R-HG element instead of the fully integrated one.
// FP32: the register budget lets two workgroups share a CU (waves per SIMD).
// kResid: the certification form r = f - K u with ||r||^2 (FP64 operator).
template <typename T, typename Idx, bool kScale, int TX, int TY, bool kFuse, bool kPF, bool kIso, bool kHG = false,
          bool kResid = false, int kFrames = 0>
__global__ void __launch_bounds__((Tile<TX, TY>::NT), (sizeof(T) == 4 ? (2 * Tile<TX, TY>::NT / 64 + 3) / 4 : 0))
apply_kernel_v2(ApplyArgs<T> a) {
  using TL = Tile<TX, TY>;
  // kPF: two staged planes and two force buffers (layer parity), so one
  // barrier per layer separates the writers of one buffer from its readers.
  constexpr int NB = kPF ? 2 : 1;
  __shared__ T su[NB][3][TL::PLANE];
  __shared__ T sc[NB][12][TL::EL];
  __shared__ double red[TL::NT / 64];

  const int tid = threadIdx.x;
  const BlockId bid = xcd_block();
  const int tile = bid.tile;
  const int x0 = (tile % a.tiles_x) * TX;
  const int y0 = (tile / a.tiles_x) * TY;
  const int z0 = bid.chunk * a.zchunk;
  const int z1 = min(z0 + a.zchunk, a.nz);
  if (z0 >= a.nz) return;
  if constexpr (kFuse || kResid) {
    if (a.active && !a.active[bid.rhs]) return;
  }
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const int64_t nxy = int64_t(nx) * ny;
  const int64_t n = nxy * nz;
  const T* u = a.u + int64_t(bid.rhs) * a.rhs_stride;
  T* out = a.out + int64_t(bid.rhs) * a.rhs_stride;

  int ld_lds[TL::ITEMS];
  int ld_off[TL::ITEMS];  // offsets inside one z plane (nx * ny < 2^31)
#pragma unroll
  for (int k = 0; k < TL::ITEMS; ++k) {
    const int idx = tid + k * TL::NT;
    ld_lds[k] = idx < TL::PLANE ? idx : -1;
    const int px = idx % TL::PX, py = idx / TL::PX;
    ld_off[k] = wrapmod(y0 - 1 + py, ny) * nx + wrapmod(x0 - 1 + px, nx);
  }
  auto load_plane = [&](int z, int buf) {
    const int64_t base = int64_t(wrapmod(z, nz)) * nxy;
#pragma unroll
    for (int k = 0; k < TL::ITEMS; ++k) {
      if (ld_lds[k] < 0) continue;
      const int64_t node = base + ld_off[k];
#pragma unroll
      for (int c = 0; c < 3; ++c) su[buf][c][ld_lds[k]] = u[c * n + node];
    }
  };

  const bool is_elem = tid < TL::EL;
  const int exl = tid % TL::EX, eyl = tid / TL::EX;
  const int e_xy = wrapmod(y0 - 1 + eyl, ny) * nx + wrapmod(x0 - 1 + exl, nx);
  auto face_modes = [&](T (&f)[3][4], int buf) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
#pragma unroll
      for (int a2 = 0; a2 < 4; ++a2) f[c][a2] = su[buf][c][(eyl + (a2 >> 1)) * TL::PX + exl + (a2 & 1)];
      face_forward(f[c]);
    }
  };
  T Bf[3][4], Ftp[3][4];
  // Element of layer z from its bottom (Bb) and top (Tf) face modes; returns the
  // bottom-face force (Fb) and top-face force (Ft) in face-mode form.
  const int64_t gbase = int64_t(bid.rhs / a.group_rhs) * a.group_stride;  // stacked cells
  auto row_of = [&](int z) { return load_row<Idx>(a.index, gbase + int64_t(wrapmod(z, nz)) * nxy + e_xy); };
  static_assert(kFrames == 0 || (std::is_same<T, float>::value && !kScale && !kHG && !kResid),
                "per-voxel frames: FP32 FI operator without a scale field");
  // Coefficients of the element of layer z: its row, times the voxel's scale
  // (kScale), rotated into the voxel's material frame (kFrames: 1 any
  // local-frame phase, 2 cubic local-frame phases only, 3 cubic and
  // orthotropic local-frame phases only).
  auto coeffs = [&](int z, int row, T (&c)[21]) {
    const int64_t voxel = int64_t(wrapmod(z, nz)) * nxy + e_xy;
    const T* crow = a.table + int64_t(row) * kRowWidth;
#pragma unroll
    for (int k = 0; k < 21; ++k) c[k] = crow[k];
    if constexpr (kScale) {
      const T sc_ = T(a.scale[voxel]);
#pragma unroll
      for (int k = 0; k < 21; ++k) c[k] *= sc_;
    }
    if constexpr (kFrames != 0) {
      const int kind = int(crow[21]);
      if (kind != 0) rotate_row<kFrames>(c, a.frames[voxel], kind);
    }
  };
  auto element = [&](const T (&c)[21], const T (&Bb)[3][4], const T (&Tf)[3][4], T (&Fb)[3][4], T (&Ft)[3][4]) {
    T q[3][8], P[3][8];
#pragma unroll
    for (int cc = 0; cc < 3; ++cc)
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        q[cc][m] = Tf[cc][m] + Bb[cc][m];
        q[cc][m + 4] = Tf[cc][m] - Bb[cc][m];
      }
    if constexpr (kHG)
      element_hg<kIso>(q, c, a.g, a.hg, P);
    else
      element_modal<kIso>(q, c, a.g, P);
#pragma unroll
    for (int cc = 0; cc < 3; ++cc) {
      // Mode 0 (a rigid translation) carries no force: P[cc][0] is 0.
      Fb[cc][0] = -P[cc][4];
      Ft[cc][0] = P[cc][4];
#pragma unroll
      for (int m = 1; m < 4; ++m) {
        Fb[cc][m] = P[cc][m] - P[cc][m + 4];
        Ft[cc][m] = P[cc][m] + P[cc][m + 4];
      }
    }
  };

  const bool is_node = tid < TL::NODES;
  const int xl = tid % TX, yl = tid / TX;
  const bool node_valid = is_node && (x0 + xl) < nx && (y0 + yl) < ny;
  const int n_xy = (y0 + yl) * nx + (x0 + xl);
  // kFuse: u . (K u) as the sum over the nodes this thread writes of
  // u_node . (K u)_node; u at the node of plane z is read from LDS while that
  // plane is resident (one layer earlier) and kept in registers.
  double rr = 0.0;  // kResid: sum of r^2 over the nodes this thread writes
  T pq = T(0);  // at most zchunk nodes per thread; FP64 from the block sum on
  auto node_value = [&](T (&v)[3], int buf) {
#pragma unroll
    for (int c = 0; c < 3; ++c) v[c] = su[buf][c][(yl + 1) * TL::PX + xl + 1];
  };

  // Next-plane prefetch: raw values of the plane after next are loaded into
  // registers before the element computation and stored to LDS after it, so
  // their HBM latency overlaps the arithmetic.
  T pfz[TL::ITEMS][3];
  auto prefetch = [&](int z) {
    const int64_t base = int64_t(wrapmod(z, nz)) * nxy;
#pragma unroll
    for (int k = 0; k < TL::ITEMS; ++k) {
      if (ld_lds[k] < 0) continue;
      const int64_t node = base + ld_off[k];
#pragma unroll
      for (int c = 0; c < 3; ++c) pfz[k][c] = u[c * n + node];
    }
  };
  auto commit = [&](int buf) {
#pragma unroll
    for (int k = 0; k < TL::ITEMS; ++k) {
      if (ld_lds[k] < 0) continue;
#pragma unroll
      for (int c = 0; c < 3; ++c) su[buf][c][ld_lds[k]] = pfz[k][c];
    }
  };

  // Prologue: bottom face of layer z0-1, layer z0-1 itself, then plane z0+1.
  load_plane(z0 - 1, 0);
  __syncthreads();
  if (is_elem) face_modes(Bf, 0);
  __syncthreads();
  load_plane(z0, 0);
  __syncthreads();
  T pnA[3] = {T(0), T(0), T(0)}, pnB[3] = {T(0), T(0), T(0)};  // kFuse: u at this thread's node
  if constexpr (kFuse) {
    if (node_valid) node_value(pnA, 0);
  }
  if (is_elem) {
    T Tf[3][4], Fb[3][4];
    face_modes(Tf, 0);
    T c[21];
    coeffs(z0 - 1, row_of(z0 - 1), c);
    element(c, Bf, Tf, Fb, Ftp);
#pragma unroll
    for (int cc = 0; cc < 3; ++cc)
#pragma unroll
      for (int m = 0; m < 4; ++m) Bf[cc][m] = Tf[cc][m];
  }
  __syncthreads();
  if constexpr (kPF) {
    load_plane(z0 + 1, 0);  // the first layer's buffer
    __syncthreads();
  }
  // The material row of the next layer is fetched one step ahead, so the
  // label -> row -> coefficients chain never stalls a step on two latencies.
  // (Plain form only: the CG form is at its register limit.)
  int row_next = (!kFuse && is_elem) ? row_of(z0) : 0;
  // One layer.  Bc/Fp: this layer's bottom face modes and the force the layer
  // below put on that face; Bn/Fn receive the top face modes and top force,
  // which are the next layer's Bc/Fp.  The loop alternates two register sets,
  // so no copies are needed between layers.
  // par: layer parity (kPF: this layer's plane and force buffer; the plane
  // after next goes to the other one).
  auto layer = [&](auto par, int z, const T (&Bc)[3][4], const T (&Fp)[3][4], T (&Bn)[3][4], T (&Fn)[3][4],
                   const T (&pc)[3], T (&pn)[3]) {
    constexpr int P = kPF ? decltype(par)::value : 0;
    int row_cur = row_next;
    if constexpr (kFuse) {
      if (is_elem) row_cur = row_of(z);
    } else {
      if (is_elem && z + 1 < z1) row_next = row_of(z + 1);
    }
    const bool more = kPF && z + 1 < z1;  // another step follows: it needs plane z+2
    // kResid: the right-hand side at this thread's output node, loaded before
    // the element work so that its latency is hidden.
    double fpre[3] = {0.0, 0.0, 0.0};
    if constexpr (kResid) {
      if (node_valid) {
        const int64_t base = int64_t(bid.rhs) * a.rhs_stride + int64_t(z) * nxy + n_xy;
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) fpre[cc] = a.f[base + cc * n];
      }
    }
    // kFrames: the rotation runs before this step's face modes and prefetch
    // are live, which keeps the register count inside the launch bound.
    T c[21];
    if constexpr (kFrames != 0) {
      if (is_elem) coeffs(z, row_cur, c);
    }
    if constexpr (!kPF) {
      load_plane(z + 1, 0);
      __syncthreads();
    }
    // LDS holds plane z+1: read everything this step needs from it first.
    if (is_elem) face_modes(Bn, P);
    if constexpr (kFuse) {
      if (node_valid && z + 1 < z1) node_value(pn, P);
    }
    if (more) prefetch(z + 2);
    if (is_elem) {
      T Fb[3][4];
      if constexpr (kFrames == 0) coeffs(z, row_cur, c);
      element(c, Bc, Bn, Fb, Fn);
#pragma unroll
      for (int cc = 0; cc < 3; ++cc) {
        T F[4];
#pragma unroll
        for (int m = 0; m < 4; ++m) F[m] = Fp[cc][m] + Fb[cc][m];
        face_transpose(F);
#pragma unroll
        for (int a2 = 0; a2 < 4; ++a2) sc[P][a2 * 3 + cc][tid] = F[a2];
      }
    }
    if constexpr (kPF) {
      if (more) commit(1 - P);  // the other plane buffer: last read one layer ago
    }
    __syncthreads();
    if (node_valid) {
      T acc[3] = {T(0), T(0), T(0)};
#pragma unroll
      for (int dy = 0; dy < 2; ++dy)
#pragma unroll
        for (int dx = 0; dx < 2; ++dx) {
          const int e = (yl + dy) * TL::EX + xl + dx;
          const int corner = (1 - dx) + 2 * (1 - dy);
#pragma unroll
          for (int cc = 0; cc < 3; ++cc) acc[cc] += sc[P][corner * 3 + cc][e];
        }
      const int64_t node = int64_t(z) * nxy + n_xy;
      if constexpr (kResid) {
        const int64_t base = int64_t(bid.rhs) * a.rhs_stride;
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) {
          const double v = fpre[cc] - double(acc[cc]);
          if (a.r32)
            a.r32[base + cc * n + node] = float(v);
          else
            a.r64[base + cc * n + node] = v;
          rr = fma(v, v, rr);
        }
      } else {
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) out[cc * n + node] = acc[cc];
        if constexpr (kFuse) pq = fma(pc[2], acc[2], fma(pc[1], acc[1], fma(pc[0], acc[0], pq)));
      }
    }
    // kPF: no barrier here.  The next layer writes the other force buffer
    // and reads the plane committed above; this layer's buffers are written
    // again two layers on, after the next layer's barrier.
  };
  T Bf2[3][4], Ftp2[3][4];
  for (int z = z0; z < z1; z += 2) {
    layer(std::integral_constant<int, 0>(), z, Bf, Ftp, Bf2, Ftp2, pnA, pnB);
    if (z + 1 < z1) layer(std::integral_constant<int, 1>(), z + 1, Bf2, Ftp2, Bf, Ftp, pnB, pnA);
  }
  if constexpr (kFuse || kResid) {
    const double total = block_reduce_sum<TL::NT>(kResid ? rr : double(pq), red);
    if (tid == 0)
      a.partial[(int64_t(bid.rhs) * gridDim.y + bid.chunk) * gridDim.x + bid.tile] = total;
  }
}

// Stencil schedule: 2 (persistent face columns), 3 (the same with the plane
// after next prefetched into registers) or 1 (per-layer full 8-node
// transform); VXL_STENCIL_VARIANT overrides the per-precision default.
int env_stencil_variant() {
  static const int v = [] {
    const char* e = std::getenv("VXL_STENCIL_VARIANT");
    if (e && e[0] >= '1' && e[0] <= '3') return e[0] - '0';
    return 0;
  }();
  return v;
}

template <typename T>
G9<T> cast_g(const ModalFactors& f) {
  G9<T> g;
  for (int k = 0; k < 3; ++k)
    for (int a = 0; a < 3; ++a) g.g[k][a] = T(f.g[k][a]);
  return g;
}

template <typename T, int TX, int TY>
void apply_grid(const vxl_cell* cell, int nrhs, int& tiles_x, int& tiles, int& chunks, int& zchunk,
                int zmin = 0) {
  const int nx = cell->grid.nx, ny = cell->grid.ny, nz = cell->grid.nz;
  tiles_x = (nx + TX - 1) / TX;
  const int tiles_y = (ny + TY - 1) / TY;
  tiles = tiles_x * tiles_y;
  // Enough workgroups to fill 256 CUs several times; chunks of >= zmin
  // planes (default 8: the recomputed layer per chunk stays <= 1/8).
  if (zmin <= 0) zmin = cell->zmin[std::is_same<T, double>::value ? 1 : 0];
  chunks = (4096 + tiles * nrhs - 1) / (tiles * nrhs);
  chunks = std::max(1, std::min(chunks, nz / std::max(1, zmin)));
  zchunk = (nz + chunks - 1) / chunks;
  chunks = (nz + zchunk - 1) / zchunk;
}

template <typename T, typename Idx, bool kScale, int TX, int TY, bool kHG>
int launch_apply_impl(const vxl_cell* cell, int nrhs, const T* u, T* out, hipStream_t s,
                      const int* active = nullptr, double* partial = nullptr, int* blocks = nullptr) {
  using TL = Tile<TX, TY>;
  ApplyArgs<T> a{};
  set_groups(a, cell, nrhs);
  a.u = u;
  a.out = out;
  a.active = active;
  a.partial = partial;
  a.rhs_stride = 3 * cell->grid.n;
  a.index = cell->mat.index;
  a.table = std::is_same<T, double>::value ? reinterpret_cast<const T*>(cell->mat.c64)
                                           : reinterpret_cast<const T*>(cell->mat.c32);
  a.scale = cell->mat.scale;
  a.nx = cell->grid.nx;
  a.ny = cell->grid.ny;
  a.nz = cell->grid.nz;
  int tiles = 0, chunks = 0;
  apply_grid<T, TX, TY>(cell, nrhs, a.tiles_x, tiles, chunks, a.zchunk);
  const ModalFactors mf = modal_factors(cell->h);
  a.g = cast_g<T>(mf);
  dim3 grid(tiles, chunks, nrhs);
  if (blocks) *blocks = tiles * chunks;
  // The C3D8R-HG element runs schedule 3 only.
  const int v = kHG ? 3 : cell->variant[std::is_same<T, double>::value ? 1 : 0];
  a.hg = T(kHG ? hg_factor(cell) / 8.0 : 0.0);
  // Cubic voxels: the squared factors go to the back-projection (one multiply
  // per gradient entry less); schedule 1 keeps the general form.
  const bool iso = cell->h[0] == cell->h[1] && cell->h[1] == cell->h[2] && v != 1;
  if (iso)
    for (int k = 0; k < 3; ++k)
      for (int d = 0; d < 3; ++d) a.g.g[k][d] = T(mf.g[k][d] * mf.g[k][d]);
#define VXL_V2(FUSE, PF)                                                                                   \
  do {                                                                                                     \
    if (iso)                                                                                               \
      hipLaunchKernelGGL((apply_kernel_v2<T, Idx, kScale, TX, TY, FUSE, PF, true, kHG>), grid, dim3(TL::NT), 0, s, \
                         a);                                                                               \
    else                                                                                                   \
      hipLaunchKernelGGL((apply_kernel_v2<T, Idx, kScale, TX, TY, FUSE, PF, false, kHG>), grid, dim3(TL::NT), 0, \
                         s, a);                                                                            \
  } while (0)
  if constexpr (kHG) {
    if (partial)
      VXL_V2(true, true);
    else
      VXL_V2(false, true);
  } else if (partial) {
    if (v == 3)
      VXL_V2(true, true);
    else
      VXL_V2(true, false);
  } else if (v == 1) {
    hipLaunchKernelGGL((apply_kernel<T, Idx, kScale, TX, TY>), grid, dim3(TL::NT), 0, s, a);
  } else if (v == 3) {
    VXL_V2(false, true);
  } else {
    VXL_V2(false, false);
  }
#undef VXL_V2
  VXL_HIP(hipGetLastError());
  return VXL_OK;
}

// FP32 stencil with per-voxel material frames (vxl_cell_set_frames): phase
// index (u8), phase rows and quaternions; always the schedule-3 form, on the
// cell's tile.  With the Bond rotation the 32 x 14 tile spills (128 VGPRs,
// 18-27 registers to scratch) and is still faster than the spill-free 32 x 8
// (146 VGPRs, fewer waves per CU and more halo elements): 3.8 against 5.2 ms
// for six right-hand sides at 256^3, random frames.
template <typename F>
int frames_tile(const vxl_cell* cell, F&& f) {
  const int ty = cell->tile_y[0];
  if (ty == kTilesF32[0]) return f(std::integral_constant<int, kTilesF32[0]>());
  if (ty == kTilesF32[2]) return f(std::integral_constant<int, kTilesF32[2]>());
  return f(std::integral_constant<int, kTilesF32[1]>());
}

template <int TX, int TY>
int launch_apply_frames(const vxl_cell* cell, int nrhs, const float* u, float* out, hipStream_t s,
                        const int* active = nullptr, double* partial = nullptr, int* blocks = nullptr) {
  using TL = Tile<TX, TY>;
  ApplyArgs<float> a{};
  set_groups(a, cell, nrhs);
  a.u = u;
  a.out = out;
  a.active = active;
  a.partial = partial;
  a.rhs_stride = 3 * cell->grid.n;
  a.index = cell->frame_index;
  a.table = cell->frame_table;
  a.frames = reinterpret_cast<const float4*>(cell->frame_quat);
  a.nx = cell->grid.nx;
  a.ny = cell->grid.ny;
  a.nz = cell->grid.nz;
  int tiles = 0, chunks = 0;
  apply_grid<float, TX, TY>(cell, nrhs, a.tiles_x, tiles, chunks, a.zchunk);
  const ModalFactors mf = modal_factors(cell->h);
  a.g = cast_g<float>(mf);
  dim3 grid(tiles, chunks, nrhs);
  if (blocks) *blocks = tiles * chunks;
  const bool iso = cell->h[0] == cell->h[1] && cell->h[1] == cell->h[2];
  if (iso)
    for (int k = 0; k < 3; ++k)
      for (int d = 0; d < 3; ++d) a.g.g[k][d] = float(mf.g[k][d] * mf.g[k][d]);
#define VXL_FR(FUSE, ISO, KIND)                                                                              \
  hipLaunchKernelGGL((apply_kernel_v2<float, uint8_t, false, TX, TY, FUSE, true, ISO, false, false, KIND>), grid, \
                     dim3(TL::NT), 0, s, a)
#define VXL_FR2(FUSE, ISO)          \
  do {                              \
    if (cell->frame_forms == 2)     \
      VXL_FR(FUSE, ISO, 2);         \
    else if (cell->frame_forms == 3) \
      VXL_FR(FUSE, ISO, 3);         \
    else                            \
      VXL_FR(FUSE, ISO, 1);         \
  } while (0)
  if (partial) {
    if (iso)
      VXL_FR2(true, true);
    else
      VXL_FR2(true, false);
  } else {
    if (iso)
      VXL_FR2(false, true);
    else
      VXL_FR2(false, false);
  }
#undef VXL_FR2
#undef VXL_FR
  VXL_HIP(hipGetLastError());
  return VXL_OK;
}

template <typename T, int TX, int TY, bool kHG = false, typename... Extra>
int dispatch_apply(const vxl_cell* cell, int nrhs, const T* u, T* out, hipStream_t s, Extra... extra) {
  const bool sc = cell->mat.scale != nullptr;
  switch (cell->mat.index_kind) {
    case VXL_INDEX_U8:
      return sc ? launch_apply_impl<T, uint8_t, true, TX, TY, kHG>(cell, nrhs, u, out, s, extra...)
                : launch_apply_impl<T, uint8_t, false, TX, TY, kHG>(cell, nrhs, u, out, s, extra...);
    case VXL_INDEX_U16:
      return sc ? launch_apply_impl<T, uint16_t, true, TX, TY, kHG>(cell, nrhs, u, out, s, extra...)
                : launch_apply_impl<T, uint16_t, false, TX, TY, kHG>(cell, nrhs, u, out, s, extra...);
    case VXL_INDEX_I32:
      return sc ? launch_apply_impl<T, int32_t, true, TX, TY, kHG>(cell, nrhs, u, out, s, extra...)
                : launch_apply_impl<T, int32_t, false, TX, TY, kHG>(cell, nrhs, u, out, s, extra...);
  }
  return fail("unknown material index kind");
}

// ---- RHS: f = -V sum_e B0^T C_e E, node-owned gather (no atomics) ----------
// sig[(row * nload + r) * 6 + i] = (C_row E_r)_i, once per coefficient row.
__global__ void __launch_bounds__(256)
row_stress_kernel(const double* table, int rows, const double* macro, int nload, double* sig) {
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= int64_t(rows) * nload) return;
  const int64_t row = t / nload;
  const int r = int(t % nload);
  const double* c = table + row * kRowWidth;
  const double* E = macro + 6 * r;
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    double acc = 0;
#pragma unroll
    for (int j = 0; j < 6; ++j) acc = fma(c[pidx(i, j)], E[j], acc);
    sig[t * 6 + i] = acc;
  }
}

// Direct form for large row tables (one row per voxel): C_e E per element.
template <typename Idx, bool kScale>
__global__ void __launch_bounds__(256)
rhs_direct_kernel(const void* index, const double* table, const float* scale, const double* macro, int nload,
                  double* f, int nx, int ny, int nz, double gx, double gy, double gz, double volume) {
  const int64_t nxy = int64_t(nx) * ny, n = nxy * nz;
  const int64_t node = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (node >= n) return;
  const int x = int(node % nx), y = int((node / nx) % ny), z = int(node / nxy);
  for (int r = 0; r < nload; ++r) {
    const double* E = macro + 6 * r;
    double fx = 0, fy = 0, fz = 0;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int ax = q & 1, ay = (q >> 1) & 1, az = q >> 2;
      const int64_t e = int64_t(wrapmod(z - az, nz)) * nxy + int64_t(wrapmod(y - ay, ny)) * nx +
                        wrapmod(x - ax, nx);
      const double* c = table + int64_t(load_row<Idx>(index, e)) * kRowWidth;
      const double sc = kScale ? double(scale[e]) : 1.0;
      double s[6];
#pragma unroll
      for (int i = 0; i < 6; ++i) {
        double acc = 0;
#pragma unroll
        for (int j = 0; j < 6; ++j) acc = fma(c[pidx(i, j)], E[j], acc);
        s[i] = kScale ? acc * sc : acc;
      }
      const double dx = ax ? gx : -gx, dy = ay ? gy : -gy, dz = az ? gz : -gz;
      fx -= volume * (dx * s[0] + dy * s[5] + dz * s[4]);
      fy -= volume * (dx * s[5] + dy * s[1] + dz * s[3]);
      fz -= volume * (dx * s[4] + dy * s[3] + dz * s[2]);
    }
    double* fr = f + int64_t(r) * 3 * n;
    fr[node] = fx;
    fr[n + node] = fy;
    fr[2 * n + node] = fz;
  }
}

// Node-owned gather of the 8 precomputed element stresses around each node.
// Grid (ceil(nx ny / 256), min(nz, 65535)): planes come from blockIdx.y (in
// steps of gridDim.y) and the in-plane position from 32-bit arithmetic
// (nx ny < 2^31), so no 64-bit division is left.
template <typename Idx, bool kScale>
__global__ void __launch_bounds__(256)
rhs_kernel(const void* index, const double* sig, const float* scale, int nload, double* f, int nx, int ny,
           int nz, double gx, double gy, double gz, double volume) {
  const int64_t nxy = int64_t(nx) * ny, n = nxy * nz;
  const int pxy = int(blockIdx.x) * int(blockDim.x) + int(threadIdx.x);
  if (pxy >= int(nxy)) return;
  const int y = pxy / nx, x = pxy - y * nx;
  const int ym = y == 0 ? ny - 1 : y - 1, xm = x == 0 ? nx - 1 : x - 1;
  for (int z = int(blockIdx.y); z < nz; z += int(gridDim.y)) {
    const int64_t node = int64_t(z) * nxy + pxy;
    const int zm = z == 0 ? nz - 1 : z - 1;
    int64_t base[8];
    double sc[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int ax = q & 1, ay = (q >> 1) & 1, az = q >> 2;
      const int64_t e = int64_t(az ? zm : z) * nxy + int64_t(ay ? ym : y) * nx + (ax ? xm : x);
      base[q] = int64_t(load_row<Idx>(index, e)) * nload * 6;
      sc[q] = kScale ? double(scale[e]) : 1.0;
    }
    for (int r = 0; r < nload; ++r) {
      double fx = 0, fy = 0, fz = 0;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int ax = q & 1, ay = (q >> 1) & 1, az = q >> 2;
        const double* sp = sig + base[q] + 6 * r;
        double s[6];
#pragma unroll
        for (int i = 0; i < 6; ++i) s[i] = kScale ? sp[i] * sc[q] : sp[i];
        const double dx = ax ? gx : -gx, dy = ay ? gy : -gy, dz = az ? gz : -gz;
        fx -= volume * (dx * s[0] + dy * s[5] + dz * s[4]);
        fy -= volume * (dx * s[5] + dy * s[1] + dz * s[3]);
        fz -= volume * (dx * s[4] + dy * s[3] + dz * s[2]);
      }
      double* fr = f + int64_t(r) * 3 * n;
      fr[node] = fx;
      fr[n + node] = fy;
      fr[2 * n + node] = fz;
    }
  }
}

// ---- mean stress <C (E + B0 u_e)>: per-block partial sums + final reduction --
constexpr int kStressBlock = 256;

// kFields: also write the element-centre fields as planes [r][i][e] (strain,
// stress) and [r][e] (energy = strain . stress / 2); null outputs are skipped.
// partial == nullptr skips the block sums.
template <typename Idx, bool kScale, bool kFields>
__global__ void __launch_bounds__(kStressBlock)
stress_kernel(const void* index, const double* table, const float* scale, const double* u,
              const double* macro, int nload, double* partial, double* strain, double* stress,
              double* energy, int nx, int ny, int nz, double gx, double gy, double gz) {
  __shared__ double red[kStressBlock];
  const int64_t nxy = int64_t(nx) * ny, n = nxy * nz;
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const bool valid = e < n;
  int x = 0, y = 0, z = 0;
  const double* c = table;
  double sc = 1.0;
  if (valid) {
    x = int(e % nx);
    y = int((e / nx) % ny);
    z = int(e / nxy);
    c = table + int64_t(load_row<Idx>(index, e)) * kRowWidth;
    if constexpr (kScale) sc = double(scale[e]);
  }
  for (int r = 0; r < nload; ++r) {
    double s[6] = {0, 0, 0, 0, 0, 0};
    if (valid) {
      double eps[6];
#pragma unroll
      for (int j = 0; j < 6; ++j) eps[j] = macro[6 * r + j];
      const double* ur = u + int64_t(r) * 3 * n;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int ax = q & 1, ay = (q >> 1) & 1, az = q >> 2;
        const int64_t node = int64_t(wrapmod(z + az, nz)) * nxy + int64_t(wrapmod(y + ay, ny)) * nx +
                             wrapmod(x + ax, nx);
        const double ux = ur[node], uy = ur[n + node], uz = ur[2 * n + node];
        const double dx = ax ? gx : -gx, dy = ay ? gy : -gy, dz = az ? gz : -gz;
        eps[0] += dx * ux;
        eps[1] += dy * uy;
        eps[2] += dz * uz;
        eps[3] += dz * uy + dy * uz;
        eps[4] += dz * ux + dx * uz;
        eps[5] += dy * ux + dx * uy;
      }
#pragma unroll
      for (int i = 0; i < 6; ++i) {
        double acc = 0;
#pragma unroll
        for (int j = 0; j < 6; ++j) acc = fma(c[pidx(i, j)], eps[j], acc);
        s[i] = acc * sc;
      }
      if constexpr (kFields) {
        if (strain)
          for (int i = 0; i < 6; ++i) strain[(int64_t(r) * 6 + i) * n + e] = eps[i];
        if (stress)
          for (int i = 0; i < 6; ++i) stress[(int64_t(r) * 6 + i) * n + e] = s[i];
        if (energy) {
          double w = 0;
#pragma unroll
          for (int i = 0; i < 6; ++i) w += s[i] * eps[i];
          energy[int64_t(r) * n + e] = 0.5 * w;
        }
      }
    }
    if (!partial) continue;
    for (int i = 0; i < 6; ++i) {
      red[threadIdx.x] = s[i];
      __syncthreads();
      for (int w = kStressBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
      }
      if (threadIdx.x == 0) partial[(int64_t(r) * 6 + i) * gridDim.x + blockIdx.x] = red[0];
      __syncthreads();
    }
  }
}

// Sums partial[row][0..blocks) in a fixed order; one workgroup per row.
__global__ void __launch_bounds__(256)
rows_sum_kernel(const double* partial, int64_t blocks, double scale, double* out) {
  __shared__ double red[256];
  const double* p = partial + int64_t(blockIdx.x) * blocks;
  double acc = 0;
  for (int64_t i = threadIdx.x; i < blocks; i += 256) acc += p[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0] * scale;
}

template <typename F>
int by_index(const vxl_cell* cell, F&& f) {
  switch (cell->mat.index_kind) {
    case VXL_INDEX_U8: return f(uint8_t());
    case VXL_INDEX_U16: return f(uint16_t());
    case VXL_INDEX_I32: return f(int32_t());
  }
  return fail("unknown material index kind");
}

}  // namespace

// Measured at 256^3 (six RHS, default tiles): the plane prefetch (3) is the
// fastest schedule in both precisions (FP32 1.37 against 1.42 ms per launch).
int default_stencil_variant(int precision) {
  if (const int v = env_stencil_variant()) return v;
  (void)precision;
  return 3;
}

// Certification residual r = f - K u (FP64 operator, default FP64 tile,
// schedule 3, chunks of kDefaultChunkPlanes planes): u FP64, r into r32 or
// r64, ||r||^2 partials per workgroup; systems with active == 0 are skipped.
template <typename Idx, bool kScale>
int launch_resid_impl(const vxl_cell* cell, int nrhs, const double* u, const double* f, float* r32, double* r64,
                      const int* active, double* partial, int* blocks, hipStream_t s) {
  constexpr int TY = kTilesF64[0];
  using TL = Tile<32, TY>;
  ApplyArgs<double> a{};
  set_groups(a, cell, nrhs);
  a.u = u;
  a.out = nullptr;
  a.f = f;
  a.r32 = r32;
  a.r64 = r64;
  a.active = active;
  a.partial = partial;
  a.rhs_stride = 3 * cell->grid.n;
  a.index = cell->mat.index;
  a.table = cell->mat.c64;
  a.scale = cell->mat.scale;
  a.nx = cell->grid.nx;
  a.ny = cell->grid.ny;
  a.nz = cell->grid.nz;
  int tiles = 0, chunks = 0;
  apply_grid<double, 32, TY>(cell, nrhs, a.tiles_x, tiles, chunks, a.zchunk, kDefaultChunkPlanes);
  const ModalFactors mf = modal_factors(cell->h);
  a.g = cast_g<double>(mf);
  const bool hg = cell->element == VXL_ELEMENT_C3D8R_HG;
  a.hg = hg ? hg_factor(cell) / 8.0 : 0.0;
  const bool iso = cell->h[0] == cell->h[1] && cell->h[1] == cell->h[2];
  if (iso)
    for (int k = 0; k < 3; ++k)
      for (int d = 0; d < 3; ++d) a.g.g[k][d] = mf.g[k][d] * mf.g[k][d];
  const dim3 grid = dim3(tiles, chunks, nrhs);
  if (blocks) *blocks = tiles * chunks;
#define VXL_R(ISO, HG) \
  hipLaunchKernelGGL((apply_kernel_v2<double, Idx, kScale, 32, TY, false, true, ISO, HG, true>), grid, dim3(TL::NT), 0, s, a)
  if (iso && hg)
    VXL_R(true, true);
  else if (iso)
    VXL_R(true, false);
  else if (hg)
    VXL_R(false, true);
  else
    VXL_R(false, false);
#undef VXL_R
  VXL_HIP(hipGetLastError());
  return VXL_OK;
}

int launch_residual_apply(const vxl_cell* cell, int nrhs, const double* u, const double* f, float* r32, double* r64,
                          const int* active, double* partial, int* blocks, hipStream_t s) {
  const bool sc = cell->mat.scale != nullptr;
  switch (cell->mat.index_kind) {
    case VXL_INDEX_U8:
      return sc ? launch_resid_impl<uint8_t, true>(cell, nrhs, u, f, r32, r64, active, partial, blocks, s)
                : launch_resid_impl<uint8_t, false>(cell, nrhs, u, f, r32, r64, active, partial, blocks, s);
    case VXL_INDEX_U16:
      return sc ? launch_resid_impl<uint16_t, true>(cell, nrhs, u, f, r32, r64, active, partial, blocks, s)
                : launch_resid_impl<uint16_t, false>(cell, nrhs, u, f, r32, r64, active, partial, blocks, s);
    case VXL_INDEX_I32:
      return sc ? launch_resid_impl<int32_t, true>(cell, nrhs, u, f, r32, r64, active, partial, blocks, s)
                : launch_resid_impl<int32_t, false>(cell, nrhs, u, f, r32, r64, active, partial, blocks, s);
  }
  return fail("unknown material index kind");
}

int residual_apply_blocks(const vxl_cell* cell, int nrhs) {
  int tx, tiles, chunks, zc;
  apply_grid<double, 32, kTilesF64[0]>(cell, nrhs, tx, tiles, chunks, zc, kDefaultChunkPlanes);
  return tiles * chunks;
}

// Runtime tile choice -> template instance (the C3D8R-HG element runs the
// default tile only).
template <typename T, typename F>
int with_tile(const vxl_cell* cell, F&& f) {
  constexpr bool d = std::is_same<T, double>::value;
  const int ty = cell->element == VXL_ELEMENT_C3D8R_HG ? (d ? kTilesF64[0] : kTilesF32[0]) : cell->tile_y[d ? 1 : 0];
  const auto& list = d ? kTilesF64 : kTilesF32;
  if (ty == list[0]) return f(std::integral_constant<int, (d ? kTilesF64[0] : kTilesF32[0])>());
  if (ty == list[1]) return f(std::integral_constant<int, (d ? kTilesF64[1] : kTilesF32[1])>());
  if (ty == list[2]) return f(std::integral_constant<int, (d ? kTilesF64[2] : kTilesF32[2])>());
  return fail("unsupported stencil tile height");
}

template <typename T>
constexpr int default_tile_y() {
  return std::is_same<T, double>::value ? kTilesF64[0] : kTilesF32[0];
}

template <typename T>
int launch_apply_t(const vxl_cell* cell, int nrhs, const T* u, T* out, hipStream_t s) {
  if constexpr (std::is_same<T, float>::value) {
    if (cell->mfma) return launch_mfma_apply(cell, nrhs, u, out, nullptr, nullptr, nullptr, s);
    if (cell->frame_index)
      return frames_tile(cell, [&](auto ty) {
        return launch_apply_frames<32, decltype(ty)::value>(cell, nrhs, u, out, s);
      });
  }
  if (cell->element == VXL_ELEMENT_C3D8R_HG)
    return dispatch_apply<T, 32, default_tile_y<T>(), true>(cell, nrhs, u, out, s);
  return with_tile<T>(cell, [&](auto ty) { return dispatch_apply<T, 32, decltype(ty)::value>(cell, nrhs, u, out, s); });
}
template <typename T>
int launch_apply_cg_t(const vxl_cell* cell, int nrhs, const T* p, T* q, const int* active, double* partial,
                      int* blocks, hipStream_t s) {
  if constexpr (std::is_same<T, float>::value) {
    if (cell->mfma) return launch_mfma_apply(cell, nrhs, p, q, active, partial, blocks, s);
    if (cell->frame_index)
      return frames_tile(cell, [&](auto ty) {
        return launch_apply_frames<32, decltype(ty)::value>(cell, nrhs, p, q, s, active, partial, blocks);
      });
  }
  if (cell->element == VXL_ELEMENT_C3D8R_HG)
    return dispatch_apply<T, 32, default_tile_y<T>(), true>(cell, nrhs, p, q, s, active, partial, blocks);
  return with_tile<T>(cell, [&](auto ty) {
    return dispatch_apply<T, 32, decltype(ty)::value>(cell, nrhs, p, q, s, active, partial, blocks);
  });
}
template <typename T>
int cg_blocks_t(const vxl_cell* cell, int nrhs, int zmin) {
  if constexpr (std::is_same<T, float>::value) {
    if (cell->mfma) return mfma_blocks(cell, nrhs);
    if (cell->frame_index)
      return frames_tile(cell, [&](auto ty) {
        int tx, tiles, chunks, zc;
        apply_grid<T, 32, decltype(ty)::value>(cell, nrhs, tx, tiles, chunks, zc, zmin);
        return tiles * chunks;
      });
  }
  return with_tile<T>(cell, [&](auto ty) {
    int tx, tiles, chunks, zc;
    apply_grid<T, 32, decltype(ty)::value>(cell, nrhs, tx, tiles, chunks, zc, zmin);
    return tiles * chunks;
  });
}

namespace {
__global__ void zero_flags_kernel(int* flag, int n) {
  if (int(threadIdx.x) < n) flag[threadIdx.x] = 0;
}

__global__ void __launch_bounds__(256)
frames_check_kernel(const uint8_t* __restrict__ index, const float4* __restrict__ quat, int64_t n, int nphase,
                    int* __restrict__ flag) {
  for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < n; e += int64_t(gridDim.x) * blockDim.x) {
    if (int(index[e]) >= nphase) flag[0] = 1;
    const float4 q = quat[e];
    const float nq = q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w;
    if (!(fabsf(nq - 1.f) <= 1e-5f)) flag[1] = 1;
  }
}
}  // namespace

void frames_release(vxl_cell* cell) {
  if (cell->frame_index) dev_free(cell->frame_index);
  if (cell->frame_table) dev_free(cell->frame_table);
  if (cell->frame_quat) dev_free(cell->frame_quat);
  cell->frame_index = nullptr;
  cell->frame_table = nullptr;
  cell->frame_quat = nullptr;
}

int launch_frames_check(const uint8_t* index, const float* quat, int64_t n, int nphase, int* flag, hipStream_t s) {
  hipLaunchKernelGGL(zero_flags_kernel, dim3(1), dim3(64), 0, s, flag, 2);
  VXL_HIP(hipGetLastError());
  const int blocks = int(std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 8192)));
  hipLaunchKernelGGL(frames_check_kernel, dim3(blocks), dim3(256), 0, s, index,
                     reinterpret_cast<const float4*>(quat), n, nphase, flag);
  VXL_HIP(hipGetLastError());
  return VXL_OK;
}

template <>
int launch_apply<float>(const vxl_cell* cell, int nrhs, const float* u, float* out, hipStream_t s) {
  return launch_apply_t<float>(cell, nrhs, u, out, s);
}
template <>
int launch_apply<double>(const vxl_cell* cell, int nrhs, const double* u, double* out, hipStream_t s) {
  return launch_apply_t<double>(cell, nrhs, u, out, s);
}

template <>
int launch_apply_cg<float>(const vxl_cell* cell, int nrhs, const float* p, float* q, const int* active,
                           double* partial, int* blocks, hipStream_t s) {
  return launch_apply_cg_t<float>(cell, nrhs, p, q, active, partial, blocks, s);
}
template <>
int launch_apply_cg<double>(const vxl_cell* cell, int nrhs, const double* p, double* q, const int* active,
                            double* partial, int* blocks, hipStream_t s) {
  return launch_apply_cg_t<double>(cell, nrhs, p, q, active, partial, blocks, s);
}
template <>
int apply_cg_blocks<float>(const vxl_cell* cell, int nrhs) {
  return cg_blocks_t<float>(cell, nrhs, 0);
}
template <>
int apply_cg_blocks<double>(const vxl_cell* cell, int nrhs) {
  return cg_blocks_t<double>(cell, nrhs, 0);
}

// Partial sums the CG form can need for any tile and chunk depth the cell may
// be switched to later (the solver allocates them once).
int apply_cg_blocks_max(const vxl_cell* cell, int nrhs, int precision) {
  vxl_cell probe = *cell;
  int most = 0;
  const bool d = precision == VXL_F64;
  for (int t = 0; t < 3; ++t) {
    probe.tile_y[d ? 1 : 0] = d ? kTilesF64[t] : kTilesF32[t];
    const int b = d ? cg_blocks_t<double>(&probe, nrhs, kMinChunkPlanes) : cg_blocks_t<float>(&probe, nrhs, kMinChunkPlanes);
    most = std::max(most, b);
  }
  if (!d) {  // the matrix-core form, which the cell may be switched to later
    probe.zmin[0] = kMinChunkPlanes;
    most = std::max(most, mfma_blocks(&probe, nrhs));
  }
  return most;
}

bool stencil_tile_supported(int precision, int tile_y) {
  const auto& list = precision == VXL_F64 ? kTilesF64 : kTilesF32;
  return tile_y == list[0] || tile_y == list[1] || tile_y == list[2];
}
int default_stencil_tile(int precision) { return precision == VXL_F64 ? kTilesF64[0] : kTilesF32[0]; }
int stencil_tile_option(int precision, int k) { return precision == VXL_F64 ? kTilesF64[k] : kTilesF32[k]; }

// One cell of a stacked cell: the same grid and rows, index block g.
vxl_cell group_view(const vxl_cell* cell, int g) {
  vxl_cell v = *cell;
  const size_t es = cell->mat.index_kind == VXL_INDEX_U8 ? 1 : cell->mat.index_kind == VXL_INDEX_U16 ? 2 : 4;
  v.mat.index = static_cast<const char*>(cell->mat.index) + size_t(g) * size_t(cell->grid.n) * es;
  v.groups = 1;
  return v;
}

int launch_rhs(const vxl_cell* cell, int nload, const double* macro_dev, double* f, double* sig, hipStream_t s) {
  if (cell->groups > 1) {  // stacked cells: each group's loads over its own index block
    if (nload % cell->groups) return fail("stacked cells: the load count must be a multiple of the cell count");
    const int per = nload / cell->groups;
    for (int g = 0; g < cell->groups; ++g) {
      const vxl_cell v = group_view(cell, g);
      VXL_TRY(launch_rhs(&v, per, macro_dev + int64_t(6) * per * g, f + int64_t(3) * cell->grid.n * per * g, sig, s));
    }
    return VXL_OK;
  }
  const auto& g = cell->grid;
  const int blocks = int((g.n + 255) / 256);
  const double gx = 0.25 / cell->h[0], gy = 0.25 / cell->h[1], gz = 0.25 / cell->h[2];
  const dim3 gplane(unsigned((int64_t(g.nx) * g.ny + 255) / 256), unsigned(std::min(g.nz, 65535)));
  if (sig) {
    const int64_t work = int64_t(cell->rows) * nload;
    hipLaunchKernelGGL(row_stress_kernel, dim3(unsigned((work + 255) / 256)), dim3(256), 0, s, cell->mat.c64,
                       cell->rows, macro_dev, nload, sig);
    VXL_HIP(hipGetLastError());
  }
  return by_index(cell, [&](auto tag) -> int {
    using Idx = decltype(tag);
    if (!sig) {
      if (cell->mat.scale)
        hipLaunchKernelGGL((rhs_direct_kernel<Idx, true>), dim3(blocks), dim3(256), 0, s, cell->mat.index,
                           cell->mat.c64, cell->mat.scale, macro_dev, nload, f, g.nx, g.ny, g.nz, gx, gy, gz,
                           cell->volume);
      else
        hipLaunchKernelGGL((rhs_direct_kernel<Idx, false>), dim3(blocks), dim3(256), 0, s, cell->mat.index,
                           cell->mat.c64, cell->mat.scale, macro_dev, nload, f, g.nx, g.ny, g.nz, gx, gy, gz,
                           cell->volume);
    } else if (cell->mat.scale)
      hipLaunchKernelGGL((rhs_kernel<Idx, true>), gplane, dim3(256), 0, s, cell->mat.index, sig,
                         cell->mat.scale, nload, f, g.nx, g.ny, g.nz, gx, gy, gz, cell->volume);
    else
      hipLaunchKernelGGL((rhs_kernel<Idx, false>), gplane, dim3(256), 0, s, cell->mat.index, sig,
                         cell->mat.scale, nload, f, g.nx, g.ny, g.nz, gx, gy, gz, cell->volume);
    VXL_HIP(hipGetLastError());
    return VXL_OK;
  });
}

// Per-row stresses pay off while the row table is small next to the grid;
// one-row-per-voxel tables use the direct form (scratch size 0).
int64_t rhs_scratch_size(const vxl_cell* cell, int nload) {
  if (int64_t(cell->rows) * 64 > cell->grid.n && cell->rows > 4096) return 0;
  return int64_t(cell->rows) * nload * 6;
}

// Per-block sums of the 21 coefficients (times the per-voxel scale) over
// the elements: the volume average of C is their total / N.
template <typename Idx, bool kScale>
__global__ void __launch_bounds__(256)
stiffness_sums_kernel(const void* index, const double* table, const float* scale, int64_t n, double* partial) {
  double acc[21];
#pragma unroll
  for (int c = 0; c < 21; ++c) acc[c] = 0.0;
  for (int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x; e < n; e += int64_t(gridDim.x) * 256) {
    const double* row = table + int64_t(load_row<Idx>(index, e)) * kRowWidth;
    const double sc = kScale ? double(scale[e]) : 1.0;
#pragma unroll
    for (int c = 0; c < 21; ++c) acc[c] = fma(sc, row[c], acc[c]);
  }
  __shared__ double red[4][21];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < 21; ++c) {
    double v = acc[c];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) red[wave][c] = v;
  }
  __syncthreads();
  if (threadIdx.x < 21)
    partial[int64_t(threadIdx.x) * gridDim.x + blockIdx.x] =
        red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
}

int64_t stress_partial_size(const vxl_cell* cell, int nload) {
  const int64_t blocks = (cell->grid.n + kStressBlock - 1) / kStressBlock;
  return blocks * 6 * nload;
}

static int stress_common(const vxl_cell* cell, int nload, const double* u, const double* macro_dev,
                         double* partial, double* out_dev, double* strain, double* stress, double* energy,
                         hipStream_t s) {
  if (cell->groups > 1) {  // stacked cells: mean stress per group (no centre fields)
    if (strain || stress || energy) return fail("centre fields take one cell, not a stacked cell");
    if (nload % cell->groups) return fail("stacked cells: the load count must be a multiple of the cell count");
    const int per = nload / cell->groups;
    for (int g = 0; g < cell->groups; ++g) {
      const vxl_cell v = group_view(cell, g);
      VXL_TRY(stress_common(&v, per, u + int64_t(3) * cell->grid.n * per * g, macro_dev + int64_t(6) * per * g,
                            partial, out_dev ? out_dev + int64_t(6) * per * g : nullptr, nullptr, nullptr, nullptr, s));
    }
    return VXL_OK;
  }
  const auto& g = cell->grid;
  const int64_t blocks = (g.n + kStressBlock - 1) / kStressBlock;
  const double gx = 0.25 / cell->h[0], gy = 0.25 / cell->h[1], gz = 0.25 / cell->h[2];
  const bool fields = strain || stress || energy;
  int rc = by_index(cell, [&](auto tag) -> int {
    using Idx = decltype(tag);
    const bool sc = cell->mat.scale != nullptr;
#define VXL_STRESS(SC, F)                                                                                \
  hipLaunchKernelGGL((stress_kernel<Idx, SC, F>), dim3(unsigned(blocks)), dim3(kStressBlock), 0, s,        \
                     cell->mat.index, cell->mat.c64, cell->mat.scale, u, macro_dev, nload, partial, strain, \
                     stress, energy, g.nx, g.ny, g.nz, gx, gy, gz)
    if (fields) {
      if (sc)
        VXL_STRESS(true, true);
      else
        VXL_STRESS(false, true);
    } else {
      if (sc)
        VXL_STRESS(true, false);
      else
        VXL_STRESS(false, false);
    }
#undef VXL_STRESS
    VXL_HIP(hipGetLastError());
    return VXL_OK;
  });
  if (rc != VXL_OK) return rc;
  if (out_dev && partial) {
    hipLaunchKernelGGL(rows_sum_kernel, dim3(6 * nload), dim3(256), 0, s, partial, blocks,
                       1.0 / double(g.n), out_dev);
    VXL_HIP(hipGetLastError());
  }
  return VXL_OK;
}

int launch_stiffness_sums(const vxl_cell* cell, double* partial, int blocks, hipStream_t s) {
  return by_index(cell, [&](auto tag) -> int {
    using Idx = decltype(tag);
    if (cell->mat.scale)
      hipLaunchKernelGGL((stiffness_sums_kernel<Idx, true>), dim3(unsigned(blocks)), dim3(256), 0, s, cell->mat.index,
                         cell->mat.c64, cell->mat.scale, cell->grid.n, partial);
    else
      hipLaunchKernelGGL((stiffness_sums_kernel<Idx, false>), dim3(unsigned(blocks)), dim3(256), 0, s, cell->mat.index,
                         cell->mat.c64, cell->mat.scale, cell->grid.n, partial);
    VXL_HIP(hipGetLastError());
    return VXL_OK;
  });
}

int launch_avg_stress(const vxl_cell* cell, int nload, const double* u, const double* macro_dev,
                      double* partial, double* out_dev, hipStream_t s) {
  return stress_common(cell, nload, u, macro_dev, partial, out_dev, nullptr, nullptr, nullptr, s);
}

int launch_center_fields(const vxl_cell* cell, int nload, const double* u, const double* macro_dev,
                         double* strain, double* stress, double* energy, hipStream_t s) {
  return stress_common(cell, nload, u, macro_dev, nullptr, nullptr, strain, stress, energy, s);
}

}  // namespace vxl
