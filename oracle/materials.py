"""Oracle material law and orientation bookkeeping (FP64, NumPy).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates
  * ``src/voxel_lmfem/_core/runtime/repro/source/femlib/materials.py:24-49``
    (orthotropic compliance inverse, isotropic Lame form);
  * ``src/voxel_lmfem/_core/materials/material_field.py:14-114``
    (PAIRS order, ``material_stiffness``, ``phase_table``,
    ``_orientation_values``, Bond rotation ``_rotate_packed``);
  * ``src/voxel_lmfem/_core/materials/independent_validation.py:32-42``
    (independent fourth-order rotation).

Pinning
  * isotropic/orthotropic: ``C11 = 3500*0.65/(1.35*0.30)`` (tests/test_import.py:35),
    compliance round trip, orthotropic with E1=E2=E3 reducing to isotropic;
  * rotations: Bond form equals the independent fourth-order form; a 90 degree
    z rotation permutes diag(10,20,30,4,5,6) to diag(20,10,30,5,4,6)
    (tests/test_solver.py:49-56); rotating an isotropic tensor is the identity;
  * orientation: axis pairs give orthonormal right-handed frames; rejected
    inputs raise the reference's messages (tests/test_solver.py:59-71).
"""
from collections.abc import Mapping
from numbers import Integral

import numpy as np

# Upper-triangle (i <= j) row-major order of the 21 independent coefficients.
PAIRS = tuple((i, j) for i in range(6) for j in range(i, 6))
PAIR_I = np.array([i for i, _ in PAIRS])
PAIR_J = np.array([j for _, j in PAIRS])
ORTHOTROPIC_KEYS = ("E1", "E2", "E3", "nu12", "nu13", "nu23", "G12", "G13", "G23")
# Engineering Voigt order 11, 22, 33, 23, 13, 12.
VOIGT_PAIRS = ((0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1))


def isotropic_to_voigt(E, nu):
    """femlib/materials.py:39-49."""
    lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    mu = E / (2.0 * (1.0 + nu))
    C = np.zeros((6, 6))
    C[:3, :3] = lam
    C[np.arange(3), np.arange(3)] += 2.0 * mu
    C[np.arange(3, 6), np.arange(3, 6)] = mu
    return C


def engineering_to_voigt(E1, E2, E3, nu12, nu13, nu23, G12, G13, G23):
    """femlib/materials.py:24-36: invert the orthotropic compliance."""
    S = np.zeros((6, 6))
    S[0, 0], S[1, 1], S[2, 2] = 1.0 / E1, 1.0 / E2, 1.0 / E3
    S[0, 1] = S[1, 0] = -nu12 / E1
    S[0, 2] = S[2, 0] = -nu13 / E1
    S[1, 2] = S[2, 1] = -nu23 / E2
    S[3, 3], S[4, 4], S[5, 5] = 1.0 / G23, 1.0 / G13, 1.0 / G12
    return np.linalg.inv(S)


def material_stiffness(spec):
    """material_field.py:18-54: validate one phase, return its SPD 6x6."""
    kind = spec.get("kind")
    if spec.get("frame", "global") not in ("global", "local"):
        raise ValueError("Material frame must be 'global' or 'local'")
    if kind == "isotropic":
        E, nu = float(spec["E"]), float(spec["nu"])
        if not (np.isfinite([E, nu]).all() and E > 0 and -1 < nu < .5):
            raise ValueError("Isotropic elasticity requires E>0 and -1<nu<0.5")
        C = isotropic_to_voigt(E, nu)
    elif kind == "orthotropic":
        v = np.array([spec[key] for key in ORTHOTROPIC_KEYS], dtype=float)
        if not np.isfinite(v).all() or (v[[0, 1, 2, 6, 7, 8]] <= 0).any():
            raise ValueError("Orthotropic moduli must be finite and positive")
        try:
            C = engineering_to_voigt(*v)
        except np.linalg.LinAlgError as exc:
            raise ValueError("Singular orthotropic compliance") from exc
    elif kind == "anisotropic":
        C = np.asarray(spec["C"], dtype=np.float64)
    else:
        raise ValueError("kind must be isotropic, orthotropic or anisotropic")
    if C.shape != (6, 6) or not np.isfinite(C).all():
        raise ValueError("C must be a finite engineering-Voigt 6x6 matrix")
    if not np.allclose(C, C.T, rtol=0, atol=1e-12 * np.max(np.abs(C))):
        raise ValueError("Elastic stiffness must be symmetric")
    C = (C + C.T) * .5
    try:
        np.linalg.cholesky(C)
    except np.linalg.LinAlgError as exc:
        raise ValueError("Elastic stiffness must be positive definite") from exc
    return C


def phase_table(materials):
    """material_field.py:57-68: sorted int64 labels, tensors, local flags."""
    if not isinstance(materials, Mapping) or not materials:
        raise ValueError("materials must be a nonempty mapping of integer labels")
    if any(not isinstance(k, Integral) or isinstance(k, bool) for k in materials):
        raise ValueError("Material labels must be integers")
    labels = np.array(sorted(materials), dtype=np.int64)
    specs = [materials[int(k)] for k in labels]
    tensors = np.stack([material_stiffness(s) for s in specs])
    local = np.array([s["kind"] != "isotropic" and s.get("frame", "global") == "local"
                      for s in specs])
    return labels, tensors, local


def orientation_values(orientation, shape, flat_indices):
    """material_field.py:71-102: frames at the selected voxels.

    Columns of each returned R are the local axes in global coordinates.
    """
    coordinates = np.unravel_index(flat_indices, shape)

    def select(values, tail):
        try:
            view = np.broadcast_to(np.asarray(values, dtype=float), tuple(shape) + tail)
        except ValueError as exc:
            raise ValueError("Orientation shape is not broadcastable to phase_id") from exc
        selected = np.array(view[coordinates], copy=True)
        if not np.isfinite(selected).all():
            raise ValueError("Active local-frame orientations must be finite")
        return selected

    if isinstance(orientation, Mapping):
        a, b = select(orientation["axis1"], (3,)), select(orientation["axis2"], (3,))
        na, nb = np.linalg.norm(a, axis=-1), np.linalg.norm(b, axis=-1)
        if (na <= 0).any() or (nb <= 0).any():
            raise ValueError("Orientation axes must be nonzero")
        a, b = a / na[:, None], b / nb[:, None]
        c = np.cross(a, b)
        nc = np.linalg.norm(c, axis=-1)
        if (nc <= 1e-12).any():
            raise ValueError("Orientation axes must not be collinear")
        c /= nc[:, None]
        return np.stack([a, np.cross(c, a), c], axis=-1)
    R = select(orientation, (3, 3))
    if not np.allclose(R.swapaxes(-1, -2) @ R, np.eye(3), rtol=0, atol=1e-10):
        raise ValueError("Orientation must be an orthonormal rotation matrix")
    if not np.allclose(np.linalg.det(R), 1, rtol=0, atol=1e-10):
        raise ValueError("Orientation must be right-handed (det=+1)")
    return R


def bond_matrix(R):
    """Bond stress-transformation matrix T (n,6,6) for engineering Voigt.

    Same construction as material_field.py:105-114: rotated stiffness is
    T C T^T.
    """
    R = np.asarray(R, dtype=np.float64).reshape(-1, 3, 3)
    i, j = [1, 0, 0], [2, 2, 1]
    r1, r2 = R[:, i, :], R[:, j, :]
    top = np.concatenate((R * R, 2 * R[:, :, i] * R[:, :, j]), -1)
    bottom = np.concatenate((r1 * r2, r1[:, :, i] * r2[:, :, j] + r1[:, :, j] * r2[:, :, i]), -1)
    return np.concatenate((top, bottom), -2)


def rotate_packed(C, R):
    """material_field.py:105-114: Bond rotation, 21 upper-triangle outputs.

    T C T^T with the sums over k in ascending order and every product
    rounded on its own: the reference's torch matmul leaves its summation
    order to the BLAS, so one order is fixed here (pinned by
    fourth_order_rotate and the diagonal known answers in the tests)."""
    T = bond_matrix(R)
    C = np.asarray(C, dtype=np.float64)
    TC = T[:, :, 0, None] * C[0][None, None, :]
    for k in range(1, 6):
        TC = TC + T[:, :, k, None] * C[k][None, None, :]
    rows = []
    for i, j in PAIRS:
        acc = TC[:, i, 0] * T[:, j, 0]
        for m in range(1, 6):
            acc = acc + TC[:, i, m] * T[:, j, m]
        rows.append(acc)
    return np.stack(rows, -1)


def fourth_order_rotate(C, R):
    """independent_validation.py:32-42: C_ijkl rotation through 3x3x3x3."""
    tensor = np.empty((3, 3, 3, 3))
    for a, (i, j) in enumerate(VOIGT_PAIRS):
        for b, (k, l) in enumerate(VOIGT_PAIRS):
            for ii, jj in ((i, j), (j, i)):
                for kk, ll in ((k, l), (l, k)):
                    tensor[ii, jj, kk, ll] = C[a, b]
    rotated = np.einsum("ia,jb,kc,ld,abcd->ijkl", R, R, R, R, tensor)
    return np.array([[rotated[i, j, k, l] for k, l in VOIGT_PAIRS] for i, j in VOIGT_PAIRS])


def pack21(C):
    C = np.asarray(C)
    return C[..., PAIR_I, PAIR_J]


def unpack21(c21):
    c21 = np.asarray(c21)
    out = np.zeros(c21.shape[:-1] + (6, 6))
    out[..., PAIR_I, PAIR_J] = c21
    out[..., PAIR_J, PAIR_I] = c21
    return out


def full_field(phase, materials, orientation=None):
    """Per-voxel 6x6 stiffness (N,6,6), C order over (z,y,x).

    Restates solver.py:25-35: global tensors by label, local-frame voxels
    rotated with the independent fourth-order rotation.
    """
    phase = np.asarray(phase)
    labels, tensors, local = phase_table(materials)
    full = tensors[np.searchsorted(labels, phase.ravel())].copy()
    for index in np.flatnonzero(local):
        selected = np.flatnonzero(phase.ravel() == labels[index])
        if selected.size:
            rotations = orientation_values(orientation, phase.shape, selected)
            for element, rotation in zip(selected, rotations):
                full[element] = fourth_order_rotate(tensors[index], rotation)
    return full


def hybrid_pool(phase, materials, orientation=None):
    """hybrid_material.py:14-70: (pool (rows,21), row map int32 (N,)).

    Rows [0, L) are the label tensors in sorted-label order; then one Bond
    rotated row per local-frame voxel, label by label ascending, C order.
    """
    phase = np.asarray(phase)
    labels, tensors, local = phase_table(materials)
    rows_of = np.searchsorted(labels, phase.reshape(-1))
    count = int(np.count_nonzero(local[rows_of]))
    index = rows_of.astype(np.int32)
    pool = np.empty((len(labels) + count, 21))
    pool[:len(labels)] = pack21(tensors)
    offset = len(labels)
    for p in np.flatnonzero(local):
        selected = np.flatnonzero(rows_of == p)
        if not selected.size:
            continue
        index[selected] = np.arange(offset, offset + selected.size, dtype=np.int32)
        R = orientation_values(orientation, phase.shape, selected)
        pool[offset:offset + selected.size] = rotate_packed(tensors[p], R)
        offset += selected.size
    return pool, index
