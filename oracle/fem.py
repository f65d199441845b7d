"""Oracle periodic HEX8 (C3D8, 2x2x2 Gauss) operators, FP64 NumPy/SciPy.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates
  * ``src/voxel_lmfem/_core/materials/independent_validation.py:77-121``
    (eight Gauss-point B matrices, COO assembly, ``splu(K[3:,3:])``, zero-mean
    shift, eight-point stress recovery);
  * ``src/voxel_lmfem/_core/materials/hybrid_kernels.cu:49-106`` semantics of
    the node-owned RHS (centroid B0) and the mean stress (centroid strain);
  * ``src/voxel_lmfem/_core/hybrid_operator/background.py:10-43`` (uniform
    27-point stencil from tensor-product 1-D integrals);
  * ``src/voxel_lmfem/_core/runtime/fi/source/femlib/batched/
    c3d8r_hourglass_operator.py:752-758`` (``recover_center_fields``: centroid
    strain E + B0 u_e and stress C_e strain per element) and the energy
    0.5 strain . stress of ``_core/runtime/benchmark.py:101``.

Node numbering inside an element is a = ax + 2*ay + 4*az with node offset
(ax, ay, az); global node id (z*Ny + y)*Nx + x, periodic.  Fields are stored
node-major (N, 3) here; the HIP path stores component planes (3, N).

Pinning
  * ``apply`` equals the assembled sparse K on random fields (two independent
    formulations: per-element Gauss B vs the COO matrix);
  * rigid translations are annihilated, K is symmetric, u.Ku > 0 off the
    null space;
  * ``uniform_stencil`` equals the assembled K rows of a uniform cell;
  * RHS and stress recovery: a uniform cell has f = 0 and <sigma> = C E;
    the laminate Schur formula and homogeneous/rotated KATs of
    tests/test_solver.py:17-56 fix the full chain through ``assemble``;
  * ``energy_contraction``: central differences of ``assemble``'s C_eff
    with respect to one element's stiffness scale (tests/test_oracle.py);
  * C3D8R-HG control (``hg_element_matrices``, ``assemble(..., element="hg")``,
    restating ``femlib/c3d8r.py:18-78``): the element matrix is symmetric,
    annihilates every linear displacement field (rigid and constant-strain
    modes reach it only through V B0^T C B0), has each hourglass pattern as an
    eigenvector with eigenvalue s(C); a uniform cell returns C_eff = C and a
    z-laminate the Schur formula; the penalty scale of an isotropic medium is
    coef (lam + 3 mu)/2 V / hbar^2;
  * ``center_fields``: a periodic affine-free field (u = 0) gives strain = E
    and stress = C_e E in every element; a rigid translation leaves both
    unchanged; a node-local displacement enters the 8 adjacent elements with
    the centroid derivatives +-1/(4h); the element mean of the stress equals
    ``avg_stress``.
"""
import itertools

import numpy as np
from scipy import sparse
from scipy.sparse.linalg import splu

OFFSETS = np.array([(x, y, z) for z in (0, 1) for y in (0, 1) for x in (0, 1)])
SIGNS = 2 * OFFSETS - 1


def gauss_B(h):
    """(8, 6, 24) strain-displacement matrices at the 2x2x2 Gauss points."""
    h = np.asarray(h, dtype=np.float64)
    Bs = []
    for point in itertools.product((-1 / np.sqrt(3), 1 / np.sqrt(3)), repeat=3):
        point = np.array(point)
        d = np.empty((8, 3))
        for j in range(3):
            others = [k for k in range(3) if k != j]
            d[:, j] = SIGNS[:, j] * np.prod(1 + SIGNS[:, others] * point[others], axis=1) / (4 * h[j])
        B = np.zeros((6, 24))
        for a, (dx, dy, dz) in enumerate(d):
            B[0, 3 * a], B[1, 3 * a + 1], B[2, 3 * a + 2] = dx, dy, dz
            B[3, 3 * a + 1], B[3, 3 * a + 2] = dz, dy
            B[4, 3 * a], B[4, 3 * a + 2] = dz, dx
            B[5, 3 * a], B[5, 3 * a + 1] = dy, dx
        Bs.append(B)
    return np.array(Bs)


def centroid_B(h):
    """(6, 24) centroid strain-displacement matrix (average of gauss_B)."""
    h = np.asarray(h, dtype=np.float64)
    d = SIGNS / (4 * h)
    B = np.zeros((6, 24))
    for a, (dx, dy, dz) in enumerate(d):
        B[0, 3 * a], B[1, 3 * a + 1], B[2, 3 * a + 2] = dx, dy, dz
        B[3, 3 * a + 1], B[3, 3 * a + 2] = dz, dy
        B[4, 3 * a], B[4, 3 * a + 2] = dz, dx
        B[5, 3 * a], B[5, 3 * a + 1] = dy, dx
    return B


def element_nodes(shape):
    """(N, 8) global node ids of every element's corners, periodic."""
    nz, ny, nx = shape
    z, y, x = np.indices(shape)
    origins = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
    coords = (origins[:, None] + OFFSETS) % [nx, ny, nz]
    return (coords[..., 2] * ny + coords[..., 1]) * nx + coords[..., 0]


def element_matrices(full, h):
    """(N, 24, 24) FI element stiffness Ke = sum_g Bg^T C Bg V/8."""
    Bs = gauss_B(h)
    return np.einsum("gai,eab,gbj->eij", Bs, full, Bs) * np.prod(h) / 8


def assemble_matrix(full, shape, h):
    n = int(np.prod(shape))
    nodes = element_nodes(shape)
    dofs = (3 * nodes[..., None] + np.arange(3)).reshape(n, 24)
    Ks = element_matrices(full, h)
    rows = np.broadcast_to(dofs[:, :, None], (n, 24, 24)).ravel()
    cols = np.broadcast_to(dofs[:, None, :], (n, 24, 24)).ravel()
    return sparse.coo_matrix((Ks.ravel(), (rows, cols)), shape=(3 * n, 3 * n)).tocsr()


def hourglass_basis():
    """(24, 12) C3D8R-HG basis (femlib/c3d8r.py:18-45): for displacement
    component c (DOF 3 a + c) the corner patterns xy, yz, zx, xyz of the corner
    signs, over sqrt(8)."""
    S = SIGNS.astype(np.float64)
    pats = np.stack([S[:, 0] * S[:, 1], S[:, 1] * S[:, 2], S[:, 2] * S[:, 0], S[:, 0] * S[:, 1] * S[:, 2]],
                    axis=1) / np.sqrt(8.0)
    H = np.zeros((24, 12))
    column = 0
    for c in range(3):
        for m in range(4):
            H[c::3, column] = pats[:, m]
            column += 1
    return H


def hourglass_scale(C, h, coef):
    """femlib/c3d8r.py:48-59 for one (6, 6) tensor or a stack (..., 6, 6)."""
    C = np.asarray(C, dtype=np.float64)
    h = np.asarray(h, dtype=np.float64)
    hbar = h.sum() / 3.0
    normal = C[..., 0, 0] + C[..., 1, 1] + C[..., 2, 2]
    first = normal + 2.0 * (C[..., 0, 1] + C[..., 0, 2] + C[..., 1, 2])
    second = normal + 2.0 * (C[..., 3, 3] + C[..., 4, 4] + C[..., 5, 5])
    material = (first + 7.0 * second) / 60.0
    return coef * np.maximum(material, 1e-12) * np.prod(h) / (hbar * hbar)


def hg_element_matrices(full, h, coef=0.1):
    """(N, 24, 24) C3D8R-HG element stiffness (femlib/c3d8r.py:62-78):
    V B0^T C B0 at the centroid plus s(C) H H^T, symmetrised."""
    B0 = centroid_B(h)
    K = np.einsum("ai,eab,bj->eij", B0, full, B0) * np.prod(h)
    H = hourglass_basis()
    K = K + hourglass_scale(full, h, coef)[:, None, None] * (H @ H.T)[None]
    return 0.5 * (K + K.transpose(0, 2, 1))


def assemble(full, shape, h, element="fi", coef=0.1):
    """independent_validation.py:77-121: K, rhs (3N,6), u (N,3,6), CH, recover.
    element="hg": the C3D8R-HG element (centroid rule + hourglass penalty),
    whose RHS and stress recovery use the centroid B0."""
    nz, ny, nx = shape
    n = int(np.prod(shape))
    if element == "hg":
        Bs = centroid_B(h)[None]
        Ks = hg_element_matrices(full, h, coef)
        Fs = -np.einsum("gai,eab->eib", Bs, full) * np.prod(h)
    else:
        Bs = gauss_B(h)
        Ks = np.einsum("gai,eab,gbj->eij", Bs, full, Bs) * np.prod(h) / 8
        Fs = -np.einsum("gai,eab->eib", Bs, full) * np.prod(h) / 8
    nodes = element_nodes(shape)
    dofs = (3 * nodes[..., None] + np.arange(3)).reshape(n, 24)
    rows = np.broadcast_to(dofs[:, :, None], (n, 24, 24)).ravel()
    cols = np.broadcast_to(dofs[:, None, :], (n, 24, 24)).ravel()
    K = sparse.coo_matrix((Ks.ravel(), (rows, cols)), shape=(3 * n, 3 * n)).tocsr()
    rhs = np.zeros((3 * n, 6))
    np.add.at(rhs, dofs, Fs)
    u = np.zeros_like(rhs)
    u[3:] = splu(K[3:, 3:].tocsc()).solve(rhs[3:])
    u = u.reshape(n, 3, 6)
    u -= u.mean(axis=0, keepdims=True)

    def recover(v, j):
        local = v.reshape(n, 3)[nodes].reshape(n, 24)
        strain = np.einsum("gij,ej->egi", Bs, local) + np.eye(6)[j]
        return np.einsum("eij,egj->egi", full, strain).mean(axis=(0, 1))

    CH = np.column_stack([recover(u[..., j], j) for j in range(6)])
    return K, rhs, u, CH, recover


class MatrixFree:
    """FP64 matrix-free periodic FI operator over a per-element 6x6 field.

    ``apply`` evaluates sum_e sum_g Bg^T C_e Bg u_e V/8 through Gauss-point
    strains (independent of the modal form used by the HIP kernels).
    Fields are (N, 3) node-major arrays.
    """

    def __init__(self, full, shape, h):
        self.full = np.asarray(full, dtype=np.float64)
        self.shape = tuple(int(s) for s in shape)
        self.h = np.asarray(h, dtype=np.float64)
        self.n = int(np.prod(self.shape))
        self.nodes = element_nodes(self.shape)
        self.Bs = gauss_B(self.h)
        self.B0 = centroid_B(self.h)
        self.V = float(np.prod(self.h))

    def apply(self, u):
        u = np.asarray(u, dtype=np.float64).reshape(self.n, 3)
        local = u[self.nodes].reshape(self.n, 24)
        out = np.zeros((self.n, 3))
        for B in self.Bs:
            strain = local @ B.T                                    # (N, 6)
            stress = np.einsum("eij,ej->ei", self.full, strain)     # (N, 6)
            force = (stress @ B) * (self.V / 8)                     # (N, 24)
            np.add.at(out, self.nodes, force.reshape(self.n, 8, 3))
        return out

    def rhs(self, macro):
        """f = -V sum_e B0^T C_e E (hybrid_kernels.cu:49-68 semantics)."""
        stress = np.einsum("eij,j->ei", self.full, np.asarray(macro, dtype=np.float64))
        force = -(stress @ self.B0) * self.V
        out = np.zeros((self.n, 3))
        np.add.at(out, self.nodes, force.reshape(self.n, 8, 3))
        return out

    def avg_stress(self, u, macro):
        """<C (E + B0 u_e)> (hybrid_kernels.cu:69-106 semantics)."""
        u = np.asarray(u, dtype=np.float64).reshape(self.n, 3)
        local = u[self.nodes].reshape(self.n, 24)
        strain = local @ self.B0.T + np.asarray(macro, dtype=np.float64)
        return np.einsum("eij,ej->ei", self.full, strain).mean(axis=0)

    def center_fields(self, u, macro):
        """Element-centre strain E + B0 u_e, stress C_e strain, energy strain.stress/2."""
        u = np.asarray(u, dtype=np.float64).reshape(self.n, 3)
        local = u[self.nodes].reshape(self.n, 24)
        strain = local @ self.B0.T + np.asarray(macro, dtype=np.float64)
        stress = np.einsum("eij,ej->ei", self.full, strain)
        return strain, stress, 0.5 * (strain * stress).sum(axis=1)

    @staticmethod
    def project_zero_mean(u):
        u = np.asarray(u, dtype=np.float64)
        return u - u.mean(axis=0, keepdims=True)


def energy_contraction(Chat, shape, h, u, macro, W):
    """out[e] = sum_ij W_ij sum_g (V/8) (E_i + B_g u_i)^T Chat_e (E_j + B_g u_j).

    Chat (N,6,6) per element, u (N,3,R) fluctuations of the macro strains
    macro (R,6), W (R,R).  With C_e = s_e Chat_e this is |Omega| dC_W/ds_e at
    the discrete solution (self-adjoint sensitivity; the reference has no
    gradient, so this is pinned by central differences of ``assemble``).
    """
    n = int(np.prod(shape))
    nodes = element_nodes(shape)
    R = u.shape[-1]
    local = u[nodes].transpose(0, 3, 1, 2).reshape(n, R, 24)  # (N, R, 24)
    V = float(np.prod(h))
    out = np.zeros(n)
    for B in gauss_B(h):
        eps = np.einsum("ia,era->eri", B, local) + np.asarray(macro)[None]      # (N, R, 6)
        sig = np.einsum("eij,erj->eri", Chat, eps)                              # (N, R, 6)
        out += (V / 8) * np.einsum("ij,eik,ejk->e", np.asarray(W), eps, sig)
    return out


STENCIL_OFFSETS = np.array([d for d in itertools.product((-1, 0, 1), repeat=3)
                            if d[2] > 0 or (d[2] == 0 and d[1] > 0) or (d[2] == d[1] == 0 and d[0] > 0)],
                           dtype=np.int32)
_VOIGT = np.array([[0, 5, 4], [5, 1, 3], [4, 3, 2]])


def uniform_stencil(C, h):
    """background.py:10-34: 13 antipodal-pair 3x3 blocks, offsets (x,y,z)."""
    C = np.asarray(C, dtype=np.float64)
    h = np.asarray(h, dtype=np.float64)
    result = []
    for d in STENCIL_OFFSETS:
        M = np.where(d == 0, 2 * h / 3, h / 6)
        D = np.where(d == 0, 2 / h, -1 / h)
        H = -d.astype(float) / 2
        G = np.zeros((3, 3))
        for i in range(3):
            for j in range(3):
                G[i, j] = D[i] * np.prod(np.delete(M, i)) if i == j else -H[i] * H[j] * M[3 - i - j]
        S = np.zeros((3, 3))
        for a in range(3):
            for b in range(3):
                S[a, b] = sum(C[_VOIGT[a, i], _VOIGT[b, j]] * G[i, j] for i in range(3) for j in range(3))
        result.append(S)
    return np.array(result)


def apply_uniform_stencil(u, shape, stencil):
    """background.py:37-43: sum_d S_d (u(x+d) + u(x-d) - 2u(x))."""
    field = np.asarray(u).reshape(*shape, 3)
    out = np.zeros_like(field)
    for d, S in zip(STENCIL_OFFSETS, stencil):
        shifted = (np.roll(field, tuple(d[::-1]), axis=(0, 1, 2))
                   + np.roll(field, tuple(-d[::-1]), axis=(0, 1, 2)) - 2 * field)
        out += np.einsum("ij,...j->...i", S, shifted)
    return out.reshape(-1, 3)


class MatrixFreeHG:
    """FP64 periodic C3D8R-HG operator: sum_e Ke_hg u_e with the dense
    element matrices of ``hg_element_matrices``.  Fields are (N, 3)."""

    def __init__(self, full, shape, h, coef=0.1):
        self.shape = tuple(int(s) for s in shape)
        self.n = int(np.prod(self.shape))
        self.nodes = element_nodes(self.shape)
        self.Ke = hg_element_matrices(np.asarray(full, dtype=np.float64), h, coef)

    def apply(self, u):
        u = np.asarray(u, dtype=np.float64).reshape(self.n, 3)
        local = u[self.nodes].reshape(self.n, 24)
        force = np.einsum("eij,ej->ei", self.Ke, local)
        out = np.zeros((self.n, 3))
        np.add.at(out, self.nodes, force.reshape(self.n, 8, 3))
        return out
