"""Input validation and the NPZ/JSON input format.

Same accepted inputs, checks and ValueError messages as the reference's
src/voxel_lmfem/inputs.py:16-64; per-voxel frames are validated when the
material field is built (materials.build_field), still before any GPU work.
"""
import json
from pathlib import Path

import numpy as np

from . import materials as mat


def used_labels(phase):
    """np.unique(phase) of an integer array: counts over the label range when
    that range is narrow (pieces counted on a thread pool), a sort otherwise.
    Same values and dtype."""
    flat = phase.reshape(-1)
    if flat.size:
        lo, hi = int(flat.min()), int(flat.max())
        if hi - lo < (1 << 22) and hi <= np.iinfo(np.int64).max:
            shift = lo != 0 or flat.dtype == np.uint64

            def count(a, b):
                piece = flat[a:b]
                return np.bincount(piece.astype(np.int64) - lo if shift else piece, minlength=hi - lo + 1)
            counts = sum(mat.map_pieces(count, flat.size))
            return (np.flatnonzero(counts) + lo).astype(phase.dtype)
    return np.unique(flat)


def validate(phase_id, materials, domain_size, orientation=None):
    """Return (phase, table, domain, orientation, labels, tensors, local)."""
    phase = np.asarray(phase_id)
    if phase.ndim != 3 or phase.dtype.kind not in "iu" or min(phase.shape) < 2:
        raise ValueError("phase_id must be an integer (Nz, Ny, Nx) array with every dimension >= 2")
    domain = np.asarray(domain_size, dtype=float)
    if domain.shape not in ((3,), (1, 3)) or not np.isfinite(domain).all() or (domain <= 0).any():
        raise ValueError("domain_size must contain three finite positive lengths in (x, y, z) order")
    domain = domain.reshape(3)
    if not isinstance(materials, dict) or not materials:
        raise ValueError("materials must be a nonempty dictionary")
    table = {}
    for key, value in materials.items():
        if isinstance(key, bool) or not isinstance(key, (str, int, np.integer)):
            raise ValueError("Material labels must be integers or integer JSON keys")
        try:
            label = int(key)
        except ValueError as exc:
            raise ValueError(f"Invalid material label: {key!r}") from exc
        if label in table:
            raise ValueError(f"Duplicate material label after conversion: {label}")
        table[label] = value
    labels, tensors, local = mat.phase_table(table)
    used = used_labels(phase)
    if not np.isin(used, labels).all():
        raise ValueError(f"Undefined material labels: {used[~np.isin(used, labels)].tolist()}")
    if orientation is None:
        for index in np.flatnonzero(local):
            if np.any(phase == labels[index]):
                raise ValueError("Local-frame anisotropic phases require orientation")
    return phase, table, domain, orientation, labels, tensors, local


def load_input(input_path, materials_path):
    """NPZ {phase_id (z,y,x), domain_size (x,y,z), orientation | axis1+axis2} + JSON."""
    with np.load(input_path, allow_pickle=False) as archive:
        if not {"phase_id", "domain_size"} <= set(archive.files):
            raise ValueError("Input NPZ must contain phase_id and domain_size")
        phase, domain = archive["phase_id"], archive["domain_size"]
        orientation = None
        if "orientation" in archive:
            orientation = archive["orientation"]
        elif "axis1" in archive or "axis2" in archive:
            if not {"axis1", "axis2"} <= set(archive.files):
                raise ValueError("Both axis1 and axis2 are required")
            orientation = dict(axis1=archive["axis1"], axis2=archive["axis2"])
    materials = json.loads(Path(materials_path).read_text(encoding="utf-8"))
    return phase, materials, domain, orientation
