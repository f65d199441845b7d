"""Host-side material law and label/orientation bookkeeping.

Mirrors the reference's material interface
(src/voxel_lmfem/_core/materials/material_field.py:14-114,
femlib/materials.py:24-49) with the same accepted inputs, validation rules and
error messages, and turns a labelled cell into the compact device form used by
libvxl: one small integer row index per voxel plus a table of 21-coefficient
rows.

Row layout (same order as the reference's hybrid pool,
src/voxel_lmfem/_core/materials/hybrid_material.py:38-61): rows [0, L) are the
phase tensors in ascending label order; local-frame voxels get Bond-rotated
rows after them, label by label, voxels in C order.  Unlike the reference,
voxels whose frames are bit-identical share one row (an exact de-duplication:
the same R gives the same rotated row bit for bit), so per-grain or
per-layer frames cost one row each instead of one row per voxel.
"""
import warnings
from collections.abc import Mapping
from numbers import Integral

import numpy as np

PAIRS = tuple((i, j) for i in range(6) for j in range(i, 6))
_PI = np.array([i for i, _ in PAIRS])
_PJ = np.array([j for _, j in PAIRS])
ORTHOTROPIC_KEYS = ("E1", "E2", "E3", "nu12", "nu13", "nu23", "G12", "G13", "G23")


def isotropic_voigt(E, nu):
    lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    mu = E / (2.0 * (1.0 + nu))
    C = np.zeros((6, 6))
    C[:3, :3] = lam
    C[np.arange(3), np.arange(3)] += 2.0 * mu
    C[np.arange(3, 6), np.arange(3, 6)] = mu
    return C


def orthotropic_voigt(E1, E2, E3, nu12, nu13, nu23, G12, G13, G23):
    S = np.zeros((6, 6))
    S[0, 0], S[1, 1], S[2, 2] = 1.0 / E1, 1.0 / E2, 1.0 / E3
    S[0, 1] = S[1, 0] = -nu12 / E1
    S[0, 2] = S[2, 0] = -nu13 / E1
    S[1, 2] = S[2, 1] = -nu23 / E2
    S[3, 3], S[4, 4], S[5, 5] = 1.0 / G23, 1.0 / G13, 1.0 / G12
    return np.linalg.inv(S)


def stiffness(spec):
    """One phase specification -> validated symmetric positive-definite 6x6."""
    kind = spec.get("kind")
    if spec.get("frame", "global") not in ("global", "local"):
        raise ValueError("Material frame must be 'global' or 'local'")
    if kind == "isotropic":
        E, nu = float(spec["E"]), float(spec["nu"])
        if not (np.isfinite([E, nu]).all() and E > 0 and -1 < nu < .5):
            raise ValueError("Isotropic elasticity requires E>0 and -1<nu<0.5")
        C = isotropic_voigt(E, nu)
    elif kind == "orthotropic":
        v = np.array([spec[key] for key in ORTHOTROPIC_KEYS], dtype=float)
        if not np.isfinite(v).all() or (v[[0, 1, 2, 6, 7, 8]] <= 0).any():
            raise ValueError("Orthotropic moduli must be finite and positive")
        try:
            C = orthotropic_voigt(*v)
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


def _stiffness_all(specs):
    """stiffness() of every spec; the anisotropic ones are checked as one
    stack (same arithmetic and result), and any spec the stack rejects is
    re-run alone for its error message."""
    out = [None] * len(specs)
    batch = [i for i, s in enumerate(specs)
             if s.get("kind") == "anisotropic" and s.get("frame", "global") in ("global", "local")]
    if len(batch) >= 8:
        try:
            C = np.stack([np.asarray(specs[i]["C"], dtype=np.float64) for i in batch])
        except (ValueError, TypeError):
            C = None
        if C is not None and C.shape[1:] == (6, 6) and np.isfinite(C).all():
            scale = np.abs(C).max(axis=(1, 2))
            sym = (np.abs(C - C.transpose(0, 2, 1)) <= (1e-12 * scale)[:, None, None]).all(axis=(1, 2))
            if sym.all():
                C = (C + C.transpose(0, 2, 1)) * .5
                try:
                    np.linalg.cholesky(C)
                    for j, i in enumerate(batch):
                        out[i] = C[j]
                except np.linalg.LinAlgError:
                    pass
    for i, s in enumerate(specs):
        if out[i] is None:
            out[i] = stiffness(s)
    return np.stack(out)


def phase_table(materials):
    """Sorted int64 labels, (L,6,6) tensors, local-frame flags."""
    if not isinstance(materials, Mapping) or not materials:
        raise ValueError("materials must be a nonempty mapping of integer labels")
    if any(not isinstance(k, Integral) or isinstance(k, bool) for k in materials):
        raise ValueError("Material labels must be integers")
    labels = np.array(sorted(materials), dtype=np.int64)
    specs = [materials[int(k)] for k in labels]
    tensors = _stiffness_all(specs)
    local = np.array([s["kind"] != "isotropic" and s.get("frame", "global") == "local" for s in specs])
    return labels, tensors, local


def host_workers():
    """Host threads for the pieced NumPy passes below (at most 16)."""
    import os
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return max(1, min(16, n))


def map_pieces(fn, length, piece=1 << 20):
    """[fn(start, stop)] over consecutive pieces of [0, length), in order, on a
    thread pool (NumPy releases the GIL inside these loops)."""
    from concurrent.futures import ThreadPoolExecutor
    bounds = [(start, min(start + piece, length)) for start in range(0, length, piece)]
    workers = host_workers()
    if workers <= 1 or len(bounds) <= 1:
        return [fn(a, b) for a, b in bounds]
    with ThreadPoolExecutor(max_workers=workers) as pool:
        return list(pool.map(lambda ab: fn(*ab), bounds))


def voxels_with_label(flat, label):
    """np.flatnonzero(flat == label), piece by piece on a thread pool."""
    parts = map_pieces(lambda a, b: np.flatnonzero(flat[a:b] == label) + a, flat.size)
    return np.concatenate(parts) if parts else np.zeros(0, dtype=np.intp)


def _as_float(values):
    """Orientation values as float64 (one conversion per array)."""
    return np.asarray(values, dtype=float)


def _pick(values, shape, tail, flat_indices, coordinates=None):
    """Entries of `values` (float64, broadcastable to shape + tail) at the
    flat voxel indices, checked finite.  A full-shape C-ordered array is read
    through its flat view, a broadcast one at the unravelled coordinates;
    both give the same values."""
    try:
        view = np.broadcast_to(values, tuple(shape) + tail)
    except ValueError as exc:
        raise ValueError("Orientation shape is not broadcastable to phase_id") from exc
    if values.shape == view.shape and values.flags.c_contiguous:
        out = values.reshape((-1,) + tail)[flat_indices]
    else:
        if coordinates is None:
            coordinates = np.unravel_index(flat_indices, shape)
        out = np.array(view[coordinates], copy=True)
    if not np.isfinite(out).all():
        raise ValueError("Active local-frame orientations must be finite")
    return out


def _axis_frames(a, b, build=True):
    """Frames from axis pairs: columns a, c x a, c with c = a x b, all unit;
    the checks (nonzero, not collinear) raise the reference's messages.

    Written per component: every norm is sqrt((x0 x0 + x1 x1) + x2 x2) and
    every cross-product entry one product minus another, the expressions and
    order of np.linalg.norm(axis=-1) and np.cross on (k, 3) arrays, so the
    values (and the checks) are the same bits, at a fraction of the cost of
    reducing along a length-3 axis."""
    a0, a1, a2 = a[:, 0], a[:, 1], a[:, 2]
    b0, b1, b2 = b[:, 0], b[:, 1], b[:, 2]
    na = np.sqrt(a0 * a0 + a1 * a1 + a2 * a2)
    nb = np.sqrt(b0 * b0 + b1 * b1 + b2 * b2)
    if (na <= 0).any() or (nb <= 0).any():
        raise ValueError("Orientation axes must be nonzero")
    a0, a1, a2 = a0 / na, a1 / na, a2 / na
    b0, b1, b2 = b0 / nb, b1 / nb, b2 / nb
    c0, c1, c2 = a1 * b2 - a2 * b1, a2 * b0 - a0 * b2, a0 * b1 - a1 * b0
    nc = np.sqrt(c0 * c0 + c1 * c1 + c2 * c2)
    if (nc <= 1e-12).any():
        raise ValueError("Orientation axes must not be collinear")
    if not build:
        return None
    c0, c1, c2 = c0 / nc, c1 / nc, c2 / nc
    d0, d1, d2 = c1 * a2 - c2 * a1, c2 * a0 - c0 * a2, c0 * a1 - c1 * a0
    return np.stack([np.stack([a0, d0, c0], -1), np.stack([a1, d1, c1], -1), np.stack([a2, d2, c2], -1)], -2)


def _matrix_frames(R):
    if not np.allclose(R.swapaxes(-1, -2) @ R, np.eye(3), rtol=0, atol=1e-10):
        raise ValueError("Orientation must be an orthonormal rotation matrix")
    if not np.allclose(np.linalg.det(R), 1, rtol=0, atol=1e-10):
        raise ValueError("Orientation must be right-handed (det=+1)")
    return R


def frames(orientation, shape, flat_indices):
    """Material frames (k,3,3) at the selected voxels; columns = local axes.

    Accepts a rotation array broadcastable to shape+(3,3) or a mapping with
    axis1/axis2 broadcastable to shape+(3,).
    """
    if isinstance(orientation, Mapping):
        a = _pick(_as_float(orientation["axis1"]), shape, (3,), flat_indices)
        b = _pick(_as_float(orientation["axis2"]), shape, (3,), flat_indices)
        return _axis_frames(a, b)
    return _matrix_frames(_pick(_as_float(orientation), shape, (3, 3), flat_indices))


def check_frames(orientation, shape, selections, chunk=65536, workers=None, keep=False):
    """Validate the frames of the selected voxels without building them.

    selections: flat voxel indices per local-frame phase, in the order they
    are checked.  Each is cut into `chunk`-voxel pieces checked in turn, as
    the reference validates (src/voxel_lmfem/inputs.py:42-47), and the error
    raised is the one the first failing piece raises there: the same checks
    in the same order (finite axis1, finite axis2, nonzero, not collinear;
    or finite, orthonormal, right-handed), on the same values.  The pieces
    run on a thread pool (NumPy releases the GIL in these loops); the
    orientation arrays are converted to float64 once up front.

    keep: also return the values read, one row per selected voxel in
    selection order, as (axis1 (m, 3), axis2 (m, 3)) or (rotation (m, 3, 3),)
    float64 arrays, for the device front end to upload instead of the whole
    grid's orientation (``build_field_device(..., compact=...)``)."""
    from concurrent.futures import ThreadPoolExecutor
    pieces = [sel[first:first + chunk] for sel in selections for first in range(0, len(sel), chunk)]
    total = sum(len(p) for p in pieces)
    mapping = isinstance(orientation, Mapping)
    tails = ((3,), (3,)) if mapping else ((3, 3),)
    kept = tuple(np.empty((total,) + tail) for tail in tails) if keep else None
    if not pieces:
        return kept
    sources = ((_as_float(orientation["axis1"]), _as_float(orientation["axis2"])) if mapping
               else (_as_float(orientation),))
    starts = np.concatenate(([0], np.cumsum([len(p) for p in pieces])))

    def one(k):
        idx = pieces[k]
        got = [_pick(src, shape, tail, idx) for src, tail in zip(sources, tails)]
        if mapping:
            _axis_frames(got[0], got[1], build=False)
        else:
            _matrix_frames(got[0])
        if keep:
            for out, values in zip(kept, got):
                out[starts[k]:starts[k + 1]] = values
    if workers is None:
        workers = host_workers()
    if workers <= 1 or len(pieces) == 1:
        for k in range(len(pieces)):
            one(k)
        return kept
    with ThreadPoolExecutor(max_workers=workers) as pool:
        for _ in pool.map(one, range(len(pieces))):  # results in piece order: the first failing piece raises
            pass
    return kept


def bond_rotate(C, R):
    """Rotated stiffness rows (k,21) = pack(T C T^T), T the Bond stress matrix
    (the reference's _rotate_packed, material_field.py:105-114).

    Every sum runs over its index in ascending order with each product rounded
    on its own (no fused multiply-add, no BLAS), the order the device
    front end (vxl_local_rows) uses, so both give the same bits."""
    R = np.asarray(R, dtype=np.float64).reshape(-1, 3, 3)
    C = np.asarray(C, dtype=np.float64)
    i, j = [1, 0, 0], [2, 2, 1]
    r1, r2 = R[:, i, :], R[:, j, :]
    T = np.concatenate((np.concatenate((R * R, 2 * R[:, :, i] * R[:, :, j]), -1),
                        np.concatenate((r1 * r2, r1[:, :, i] * r2[:, :, j] + r1[:, :, j] * r2[:, :, i]), -1)),
                       -2)
    TC = T[:, :, 0, None] * C[0][None, None, :]
    for k in range(1, 6):
        TC = TC + T[:, :, k, None] * C[k][None, None, :]
    out = []
    for a, b in PAIRS:
        acc = TC[:, a, 0] * T[:, b, 0]
        for m in range(1, 6):
            acc = acc + TC[:, a, m] * T[:, b, m]
        out.append(acc)
    return np.stack(out, -1)


def rotation_quaternions(R):
    """Unit quaternions (k,4) = (w, x, y, z) of rotations R (k,3,3), with
    R = [[1-2(y2+z2), 2(xy-wz), 2(xz+wy)], [2(xy+wz), 1-2(x2+z2), 2(yz-wx)],
    [2(xz-wy), 2(yz+wx), 1-2(x2+y2)]]: Shepperd's method (the largest of
    w, x, y, z is taken from the diagonal, the others from the off-diagonal
    sums), then normalised."""
    R = np.asarray(R, dtype=np.float64).reshape(-1, 3, 3)
    d0, d1, d2 = R[:, 0, 0], R[:, 1, 1], R[:, 2, 2]
    cand = np.stack([1 + d0 + d1 + d2, 1 + d0 - d1 - d2, 1 - d0 + d1 - d2, 1 - d0 - d1 + d2], -1)
    pick = np.argmax(cand, axis=-1)
    big = 0.5 * np.sqrt(np.maximum(cand[np.arange(len(R)), pick], 0.0))
    f = 0.25 / np.where(big > 0, big, 1.0)
    s10, s01 = R[:, 1, 0] + R[:, 0, 1], R[:, 1, 0] - R[:, 0, 1]
    s20, s02 = R[:, 2, 0] + R[:, 0, 2], R[:, 0, 2] - R[:, 2, 0]
    s21, s12 = R[:, 2, 1] + R[:, 1, 2], R[:, 2, 1] - R[:, 1, 2]
    q = np.empty((len(R), 4))
    forms = [
        (big, s12 * f, s02 * f, s01 * f),
        (s12 * f, big, s10 * f, s20 * f),
        (s02 * f, s10 * f, big, s21 * f),
        (s01 * f, s20 * f, s21 * f, big),
    ]
    for k, form in enumerate(forms):
        sel = pick == k
        for c in range(4):
            q[sel, c] = form[c][sel]
    return q / np.linalg.norm(q, axis=-1, keepdims=True)


def phase_frames(phase, materials, orientation=None):
    """Per-voxel phase slots and frames for the stencil that rotates each
    voxel's phase row itself (Cell.set_frames): (phase rows (P,21) FP64 in
    their own frames, local-frame flags (P,), slot per voxel (N,) uint8, unit
    quaternion per voxel (N,4) FP64, identity outside local-frame phases).
    The frames are the ones build_field rotates by (same validation)."""
    phase = np.asarray(phase)
    labels, tensors, local = phase_table(materials)
    if len(labels) > 256:
        raise ValueError("Per-voxel frames take at most 256 material labels")
    flat = phase.reshape(-1)
    slot = np.searchsorted(labels, flat)
    if (slot >= len(labels)).any() or not np.array_equal(labels[np.minimum(slot, len(labels) - 1)], flat):
        used = np.unique(flat)
        raise ValueError(f"Undefined material labels: {used[~np.isin(used, labels)].tolist()}")
    quat = np.zeros((flat.size, 4))
    quat[:, 0] = 1.0
    for p in np.flatnonzero(local):
        selected = np.flatnonzero(slot == p)
        if not selected.size:
            continue
        if orientation is None:
            raise ValueError("Local-frame anisotropic phases require orientation")
        quat[selected] = rotation_quaternions(frames(orientation, phase.shape, selected))
    return pack21(tensors), local, slot.astype(np.uint8), quat


def pack21(C):
    return np.asarray(C)[..., _PI, _PJ]


def unpack21(rows):
    rows = np.asarray(rows)
    out = np.zeros(rows.shape[:-1] + (6, 6))
    out[..., _PI, _PJ] = rows
    out[..., _PJ, _PI] = rows
    return out


def _unique_frames(R):
    """Exact de-duplication of (k,3,3) frames: (unique (u,3,3), inverse (k,))."""
    flat = np.ascontiguousarray(R.reshape(len(R), 9))
    keys = flat.view(np.dtype((np.void, flat.dtype.itemsize * 9))).ravel()
    _, first, inverse = np.unique(keys, return_index=True, return_inverse=True)
    return R[first], inverse.reshape(-1)


class MaterialField:
    """Device-ready material description of one cell.

    rows   : (rows, 21) FP64 coefficient table
    index  : (N,) row of every voxel (uint8/uint16/int32), C order over z,y,x
    labels, tensors, local : the validated phase table
    """

    def __init__(self, rows, index, labels, tensors, local):
        self.rows = rows
        self.index = index
        self.labels = labels
        self.tensors = tensors
        self.local = local

    @property
    def index_kind(self):
        return {np.dtype(np.uint8): 1, np.dtype(np.uint16): 2, np.dtype(np.int32): 4}[self.index.dtype]

    def voxel_rows(self, positions):
        return self.rows[self.index[positions]]

    def lattice_rows(self, shape):
        """Rows at the Zeman sampling lattice (every max(1, n//16)-th voxel per axis, C order)."""
        nz, ny, nx = shape
        z = np.arange(0, nz, max(1, nz // 16))
        y = np.arange(0, ny, max(1, ny // 16))
        x = np.arange(0, nx, max(1, nx // 16))
        pos = (z[:, None, None] * (ny * nx) + y[None, :, None] * nx + x[None, None, :]).reshape(-1)
        return self.voxel_rows(pos)

    def used_labels(self, phase):
        return np.unique(np.asarray(phase))


class DeviceField(MaterialField):
    """MaterialField whose row index lives on the GPU (built by libvxl from a
    resident label array); the host keeps the phase rows, the labels present
    in the cell and the Zeman lattice rows.  With local-frame phases the full
    row table (phase rows, then one Bond-rotated row per voxel) is `rows_dev`
    on the device."""

    def __init__(self, rows, index_dev, kind, labels, tensors, local, present, samples, rows_dev=None):
        super().__init__(rows, None, labels, tensors, local)
        self.index_dev = index_dev
        self.kind = kind
        self.present = present
        self.samples = samples
        self.rows_dev = rows_dev
        self._lattice = None

    @property
    def index_kind(self):
        return self.kind

    @property
    def row_count(self):
        return len(self.rows) if self.rows_dev is None else int(self.rows_dev.shape[0])

    def lattice_rows(self, shape):
        if self.rows_dev is None:
            return self.rows[self.samples]
        if self._lattice is None:   # gather the sampled rows from the device table
            import torch
            pick = torch.as_tensor(self.samples.astype(np.int64), device=self.rows_dev.device)
            self._lattice = self.rows_dev.index_select(0, pick).cpu().numpy()
        return self._lattice

    def used_labels(self, phase=None):
        return self.present


def index_dtype(rows):
    """Narrowest row-index type for a table of `rows` rows (libvxl index kinds)."""
    if rows <= 256:
        return np.uint8
    if rows <= 65536:
        return np.uint16
    if rows <= np.iinfo(np.int32).max:
        return np.int32
    raise ValueError("Coefficient table exceeds int32 indexing")


_FRAME_ERRORS = ("Active local-frame orientations must be finite", "Orientation axes must be nonzero",
                 "Orientation axes must not be collinear", "Orientation must be an orthonormal rotation matrix",
                 "Orientation must be right-handed (det=+1)")


def _device_frames(orientation, shape, torch, device):
    """(mode, a_dev, b_dev, stride) for vxl_local_rows: one frame for all
    voxels (stride 0) or one per voxel, from a rotation array or an axis pair
    broadcastable to the grid (the host path's accepted forms)."""
    def upload(host):
        # The host array is only read by the copy to the device; a read-only
        # one (a full-size broadcast view, a memory map) needs no host copy.
        with warnings.catch_warnings():
            warnings.filterwarnings("ignore", message="The given NumPy array is not writable")
            return torch.as_tensor(host, device=device)

    def grid_array(values, tail):
        a = np.asarray(values, dtype=np.float64)
        if a.shape == tail:
            return upload(np.ascontiguousarray(a)), 0
        try:
            full = np.broadcast_to(a, tuple(shape) + tail)
        except ValueError as exc:
            raise ValueError("Orientation shape is not broadcastable to phase_id") from exc
        return upload(np.ascontiguousarray(full.reshape(-1, *tail))), int(np.prod(tail))

    if isinstance(orientation, Mapping):
        a, sa = grid_array(orientation["axis1"], (3,))
        b, sb = grid_array(orientation["axis2"], (3,))
        if sa != sb:   # one of the two varies per voxel: expand the other too
            a, sa = grid_array(np.broadcast_to(np.asarray(orientation["axis1"], float), tuple(shape) + (3,)), (3,))
            b, sb = grid_array(np.broadcast_to(np.asarray(orientation["axis2"], float), tuple(shape) + (3,)), (3,))
        return 1, a, b, sa
    R, sr = grid_array(orientation, (3, 3))
    return 2, R, None, sr


def build_field_device(phase_dev, materials, orientation=None, deduplicate=True, compact=None):
    """Label and orientation bookkeeping of a cell whose int32/int64 labels are
    resident on the GPU.

    Global-frame phases: the same rows and per-voxel row index as
    ``build_field`` (rows = phase tensors in ascending label order), index
    computed by ``vxl_label_rows``.  Local-frame phases: every voxel of such a
    phase gets its own Bond-rotated row, built on the device by
    ``vxl_local_rows`` after the phase rows (int32 index), with the host
    path's frame checks and messages.  Raises the reference's ValueError for
    labels without a material.  deduplicate: voxels whose rows are bitwise
    equal share one row (index narrowed to u16 / u8 where it fits); False
    keeps one row per voxel (row L + voxel).  compact: (selections,
    values) with the flat voxel lists of every local-frame phase in label
    order and their frames as ``check_frames(..., keep=True)`` returns them;
    only those values are uploaded and their rows built, instead of the
    whole grid's orientation (the rows are the same bits).
    """
    import ctypes
    import torch
    from . import _native
    from ._native import check, dptr
    labels, tensors, local = phase_table(materials)
    if phase_dev.dim() != 3 or phase_dev.dtype not in (torch.int32, torch.int64) or not phase_dev.is_cuda:
        raise ValueError("phase_dev must be a 3-D int32/int64 GPU tensor")
    phase_dev = phase_dev.contiguous()
    nz, ny, nx = (int(v) for v in phase_dev.shape)
    n = nz * ny * nx
    lo = int(labels[0])
    span = int(labels[-1]) - lo + 1
    if span > (1 << 24):
        raise ValueError("label range too wide for the device lookup table; use build_field")
    L = len(labels)
    lut = np.full(span, -1, dtype=np.int32)
    lut[labels - lo] = np.arange(L, dtype=np.int32)
    lut[labels[local] - lo] = -2
    with_local = bool(local.any())
    dtype = np.int32 if with_local else index_dtype(L)
    kind = {np.uint8: 1, np.uint16: 2, np.int32: 4}[dtype]
    tdtype = {np.uint8: torch.uint8, np.uint16: torch.int16, np.int32: torch.int32}[dtype]
    index = torch.empty(n, dtype=tdtype, device=phase_dev.device)
    present = np.zeros(span, dtype=np.int32)
    undefined = np.zeros(1, dtype=np.int32)
    cz, cy, cx = (-(-v // max(1, v // 16)) for v in (nz, ny, nx))
    samples = np.zeros(cz * cy * cx, dtype=np.int32)
    ip = ctypes.POINTER(ctypes.c_int)
    lib = _native.load()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    check(lib.vxl_label_rows(nz, ny, nx, phase_dev.element_size(), dptr(phase_dev), lo, span,
                             lut.ctypes.data_as(ip), kind, dptr(index), L, present.ctypes.data_as(ip),
                             undefined.ctypes.data_as(ip), samples.ctypes.data_as(ip), stream), "vxl_label_rows")
    if undefined[0]:
        used = np.unique(phase_dev.cpu().numpy())
        raise ValueError(f"Undefined material labels: {used[~np.isin(used, labels)].tolist()}")
    used = labels[present[labels - lo] != 0]
    phase_rows = pack21(tensors)
    rows_dev = None
    if with_local:
        if orientation is None and np.isin(labels[local], used).any():
            raise ValueError("Local-frame anisotropic phases require orientation")
        if L + n > np.iinfo(np.int32).max:  # rows are addressed by an int32 index and an int row count
            raise ValueError("Too many voxels for one coefficient row each (int32 row index); use build_field, "
                             "which shares one row between voxels with equal frames")
        present_local = orientation is not None and np.isin(labels[local], used).any()
        if present_local and compact is not None:
            sels, values = compact
            vox = np.concatenate(sels) if len(sels) else np.zeros(0, dtype=np.int64)
            if len(sels) != int(local.sum()) or len(vox) != len(values[0]):
                raise ValueError("compact frames must hold every local-frame phase's voxels and their values")
            m = len(vox)
        else:
            m = n
        rows_dev = torch.zeros((L + m, 21), dtype=torch.float64, device=phase_dev.device)
        rows_dev[:L] = torch.as_tensor(phase_rows, device=phase_dev.device)
        if present_local:
            slot = np.full(span, -1, dtype=np.int32)
            slot[labels[local] - lo] = np.arange(int(local.sum()), dtype=np.int32)
            cloc = np.ascontiguousarray(phase_rows[local])
            bad = np.zeros(5, dtype=np.int32)
            if compact is None:   # every voxel's frame from the grid's orientation
                mode, a, b, stride = _device_frames(orientation, (nz, ny, nx), torch, phase_dev.device)
                grid, lab, lab_bytes = (nz, ny, nx), phase_dev, phase_dev.element_size()
            else:   # the selected voxels' frames only, as one row of m voxels
                dev = phase_dev.device
                counts = [len(sel) for sel in sels]
                lab = torch.as_tensor(np.repeat(labels[local], counts).astype(np.int32), device=dev)
                a = torch.as_tensor(np.ascontiguousarray(values[0]).reshape(m, -1), device=dev)
                b = torch.as_tensor(np.ascontiguousarray(values[1]), device=dev) if len(values) > 1 else None
                mode, stride = (1, 3) if len(values) > 1 else (2, 9)
                grid, lab_bytes = (1, 1, m), 4
            if m:
                check(lib.vxl_local_rows(grid[0], grid[1], grid[2], lab_bytes, dptr(lab), lo, span,
                                         slot.ctypes.data_as(ip), int(local.sum()),
                                         cloc.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), mode, dptr(a),
                                         dptr(b) if b is not None else None, stride, dptr(rows_dev), L,
                                         bad.ctypes.data_as(ip), stream), "vxl_local_rows")
            for flag, message in zip(bad, _FRAME_ERRORS):
                if flag:
                    raise ValueError(message)
            if compact is not None and m:
                # label_rows gave voxel v of a local phase row L + v; its row is
                # now L + (its position in the selections).
                idx = index.to(torch.int64)
                vox_dev = torch.as_tensor(vox.astype(np.int64), device=phase_dev.device)
                idx[vox_dev] = L + torch.arange(m, dtype=torch.int64, device=phase_dev.device)
                pick = samples.astype(np.int64)
                local_pick = pick >= L
                if local_pick.any():
                    pick[local_pick] = idx.index_select(
                        0, torch.as_tensor(pick[local_pick] - L, device=phase_dev.device)).cpu().numpy()
                samples = pick.astype(np.int32)
                index = idx.to(torch.int32)
    field = DeviceField(phase_rows, index, kind, labels, tensors, local, used, samples, rows_dev)
    if rows_dev is not None and deduplicate:
        _share_equal_rows(field, L, torch)
    return field


# Odd multipliers of the row key (fixed, below 2^63): the two word-mixing
# steps, then one weight per coefficient word.
_ROW_KEY_MIX = (0x3F58476D1CE4E5B9, 0x14D049BB133111EB)
_ROW_KEY_MULTIPLIERS = [(0x9E3779B97F4A7C15 * (2 * k + 1) + 0x632BE59BD9B4E019 * k) % (1 << 63) | 1 for k in range(21)]


def _share_equal_rows(field, L, torch):
    """One table row per distinct Bond-rotated row instead of one per
    local-frame voxel.  Equal frames give bitwise equal rows (the reference
    weave repeats about a thousand frames over millions of yarn voxels), and
    those then share one entry, as the host path's de-duplication does; the
    coefficients every voxel sees are unchanged bit for bit.  Rows are keyed
    by a 64-bit mix of their words; every row is then compared word for word
    with the row of its key, and a collision keeps one row per voxel.  The
    index narrows to u16 / u8 when the rows allow, and the Zeman lattice
    samples follow the new row numbers."""
    idx = field.index_dev.to(torch.int64)
    vox = torch.nonzero(idx >= L).squeeze(1)
    count = int(vox.numel())
    dev = idx.device
    new_idx = idx.clone()
    if count == 0:   # no voxel of a local-frame phase: the phase rows alone
        table = field.rows_dev[:L]
    else:
        rows = field.rows_dev.index_select(0, idx.index_select(0, vox))
        words = rows.view(torch.int64)
        # Key: every word mixed on its own (xor-shift-multiply, so that sign
        # flips of two words, as frames +theta and -theta give, do not cancel),
        # then summed with odd per-word weights, wrapping modulo 2^64.  It
        # only groups rows; the word-for-word comparison decides.
        z = words ^ (words >> 30)
        z = z * _ROW_KEY_MIX[0]
        z = z ^ (z >> 27)
        z = z * _ROW_KEY_MIX[1]
        mult = torch.tensor(_ROW_KEY_MULTIPLIERS, dtype=torch.int64, device=dev)
        key = (z * mult).sum(dim=1)
        sorted_keys, order = torch.sort(key)
        starts = torch.ones(count, dtype=torch.bool, device=dev)
        starts[1:] = sorted_keys[1:] != sorted_keys[:-1]
        group = torch.cumsum(starts, 0) - 1
        first = order[starts]                   # one voxel per group (its members are checked equal)
        inverse = torch.empty_like(order)
        inverse[order] = group
        if not torch.equal(words.index_select(0, first).index_select(0, inverse), words):
            return
        new_idx[vox] = L + inverse
        table = torch.cat([field.rows_dev[:L], rows.index_select(0, first)])
    total = int(table.shape[0])
    dtype = index_dtype(total)
    if dtype == np.uint8:
        index = new_idx.to(torch.uint8)
    elif dtype == np.uint16:   # stored as int16 with the u16 bit pattern
        index = torch.where(new_idx >= 32768, new_idx - 65536, new_idx).to(torch.int16)
    else:
        index = new_idx.to(torch.int32)
    # Lattice samples hold row numbers: map every old row to its new one.
    row_map = torch.arange(int(field.rows_dev.shape[0]), dtype=torch.int64, device=dev)
    if count:
        row_map[idx.index_select(0, vox)] = new_idx.index_select(0, vox)
    pick = torch.as_tensor(field.samples.astype(np.int64), device=dev)
    field.samples = row_map.index_select(0, pick).cpu().numpy().astype(np.int32)
    field.index_dev = index.contiguous()
    field.kind = {np.uint8: 1, np.uint16: 2, np.int32: 4}[dtype]
    field.rows_dev = table.contiguous()
    field._lattice = None


def build_field(phase, materials, orientation=None, *, deduplicate=True):
    """Label/orientation bookkeeping of one cell -> MaterialField."""
    phase = np.asarray(phase)
    labels, tensors, local = phase_table(materials)
    flat = phase.reshape(-1)
    slot = np.searchsorted(labels, flat)
    if (slot >= len(labels)).any() or not np.array_equal(labels[np.minimum(slot, len(labels) - 1)], flat):
        used = np.unique(flat)
        raise ValueError(f"Undefined material labels: {used[~np.isin(used, labels)].tolist()}")
    table = [pack21(tensors)]
    index = slot.astype(np.int64)
    offset = len(labels)
    for p in np.flatnonzero(local):
        selected = np.flatnonzero(slot == p)
        if not selected.size:
            continue
        if orientation is None:
            raise ValueError("Local-frame anisotropic phases require orientation")
        R = frames(orientation, phase.shape, selected)
        if deduplicate:
            uniq, inverse = _unique_frames(R)
            table.append(bond_rotate(tensors[p], uniq))
            index[selected] = offset + inverse
            offset += len(uniq)
        else:
            table.append(bond_rotate(tensors[p], R))
            index[selected] = offset + np.arange(selected.size)
            offset += selected.size
    rows = np.concatenate(table, axis=0)
    return MaterialField(rows, index.astype(index_dtype(len(rows))), labels, tensors, local)
