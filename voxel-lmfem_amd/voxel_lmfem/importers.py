"""Import an external labelled voxel volume into the solve input format.

Contract (the reference's ``import_model``, src/voxel_lmfem/importers.py:66-157,
README.md:124-133): the output is a NEW directory holding

* ``model.npz``: ``phase_id`` in z, y, x order, ``domain_size`` in x, y, z
  order and the direction arrays (``orientation`` or ``axis1``/``axis2``);
* ``materials.json``: the validated phase table;
* ``import.json``: schema 1 -- source files with SHA-256, source and
  normalised axes and shapes, physical size, spacing, voxel count and volume
  fraction per phase.

Geometry is only permuted, never resampled; lengths and vector/tensor
components stay in global x, y, z.  Error messages are the reference's (they
are what callers and tests match on).  A rejected model leaves nothing behind.

Design: the import is a short pipeline over one ``_Source`` record --
``_read_source`` (files -> arrays), ``_normalise`` (axis permutation of the
labels and of any spatial direction field), ``_physical_size``,
``_read_materials`` and ``validate`` -- and ``_publish`` writes the three files
into a private staging directory beside the target, renames it into place
only when everything is written, and removes it on any failure.
"""
import hashlib
import json
import os
import shutil
import tempfile
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from . import materials as mat
from .inputs import validate

NORMAL_AXES = "zyx"
SCHEMA_VERSION = 1
# Properties each material kind must name (the material law reads them).
KIND_PROPERTIES = {
    "isotropic": frozenset({"E", "nu"}),
    "orthotropic": frozenset(mat.ORTHOTROPIC_KEYS),
    "anisotropic": frozenset({"C"}),
}
_INT64 = np.iinfo(np.int64)
_DIRECTION_KEYS = ("orientation", "axis1", "axis2")


@dataclass
class _Source:
    labels: np.ndarray          # as stored, source axis order
    domain: object = None       # domain_size stored beside the labels, if any
    directions: object = None   # (3,3)/(...,3,3) array, {"axis1", "axis2"} or None


def file_sha256(path, chunk=1 << 20):
    """Hex SHA-256 of a file, read in fixed-size chunks into one buffer."""
    digest = hashlib.sha256()
    buffer = bytearray(chunk)
    view = memoryview(buffer)
    with Path(path).open("rb", buffering=0) as stream:
        while True:
            got = stream.readinto(buffer)
            if not got:
                return digest.hexdigest()
            digest.update(view[:got])


def _arrays(path, wanted):
    """{name: array} of the wanted names present in an NPZ (pickles refused)."""
    with np.load(path, allow_pickle=False) as archive:
        return {name: archive[name] for name in wanted if name in archive.files}


def _directions_of(found):
    """The one direction representation among an archive's arrays, or None."""
    pair = [name for name in ("axis1", "axis2") if name in found]
    if "orientation" in found:
        if pair:
            raise ValueError("Use orientation or axis1/axis2, rather than both representations")
        return found["orientation"]
    if not pair:
        return None
    if len(pair) == 1:
        raise ValueError("Both axis1 and axis2 are required")
    return {name: found[name] for name in pair}


def _read_source(path, key):
    kind = path.suffix.lower()
    if kind == ".npy":
        return _Source(np.load(path, allow_pickle=False))
    if kind != ".npz":
        raise ValueError("Supported volume formats: .npy and .npz")
    with np.load(path, allow_pickle=False) as archive:
        if key not in archive.files:
            raise ValueError(f"NPZ has no {key!r} array; select its label array with --key")
    found = _arrays(path, (key, "domain_size") + _DIRECTION_KEYS)
    return _Source(found[key], found.get("domain_size"), _directions_of(found))


def _read_directions(path):
    kind = path.suffix.lower()
    if kind == ".npy":
        return np.load(path, allow_pickle=False)
    if kind != ".npz":
        raise ValueError("Orientation input must be .npy or .npz")
    directions = _directions_of(_arrays(path, _DIRECTION_KEYS))
    if directions is None:
        raise ValueError("Orientation NPZ needs orientation or axis1/axis2 arrays")
    return directions


def _check_labels(labels):
    if labels.ndim != 3 or labels.dtype.kind not in "biu" or min(labels.shape) < 2:
        raise ValueError("Input must be a 3D integer label array with every dimension >= 2; "
                         "segment image intensities first")
    if labels.dtype.kind == "b":
        return labels.astype(np.uint8)
    if labels.dtype.kind == "u" and int(labels.max()) > _INT64.max:
        raise ValueError("Material IDs must fit a signed 64-bit integer")
    return labels


def _spatial_permute(values, source_shape, perm, tail):
    """A direction field over the source grid, re-laid over the normalised
    grid.  Only the three leading (spatial) axes move; the trailing component
    axes `tail` are global x, y, z and stay put.  A single uniform value
    (shape == tail) applies everywhere and is returned as is."""
    field = np.asarray(values, dtype=float)
    if field.shape == tail:
        return field
    expected = tuple(source_shape) + tail
    if field.shape != expected:
        raise ValueError(f"Orientation shape must be {tail} or {expected} in source axis order")
    return np.transpose(field, perm + tuple(range(3, field.ndim)))


def _normalise(source, axes):
    """Labels and directions in z, y, x order (native byte order)."""
    perm = tuple(axes.index(axis) for axis in NORMAL_AXES)   # source axis that becomes z, y, x
    native = source.labels.dtype.newbyteorder("=")
    labels = np.ascontiguousarray(np.transpose(source.labels, perm)).astype(native, copy=False)
    directions = source.directions
    if isinstance(directions, dict):
        directions = {name: _spatial_permute(directions[name], source.labels.shape, perm, (3,))
                      for name in ("axis1", "axis2")}
    elif directions is not None:
        directions = _spatial_permute(directions, source.labels.shape, perm, (3, 3))
    return labels, directions


def _physical_size(shape_zyx, spacing, domain_size, stored):
    """domain_size (x, y, z) from an explicit spacing, an explicit size, or the
    size stored in the source (in that order of precedence)."""
    if spacing is not None:
        step = np.asarray(spacing, dtype=float)
        if step.shape != (3,) or not np.isfinite(step).all() or np.any(step <= 0):
            raise ValueError("Spacing must be three finite positive values in x,y,z order")
        return step * np.array(shape_zyx[::-1], dtype=float)
    size = stored if domain_size is None else domain_size
    if size is None:
        raise ValueError("Provide --spacing or --domain-size, or domain_size in the source NPZ")
    return size


def _read_materials(path):
    table = json.loads(Path(path).read_text(encoding="utf-8-sig"))
    if not isinstance(table, dict) or not table or any(not isinstance(spec, dict) for spec in table.values()):
        raise ValueError("Materials JSON must map phase IDs to material property objects")
    for label, spec in table.items():
        needed = KIND_PROPERTIES.get(spec.get("kind"))
        if needed is None:
            raise ValueError(f"Material {label}: kind must be isotropic, orthotropic or anisotropic")
        absent = sorted(needed.difference(spec))
        if absent:
            raise ValueError(f"Material {label}: missing {', '.join(absent)}")
        if not _INT64.min <= int(label) <= _INT64.max:
            raise ValueError(f"Material ID {label} must fit a signed 64-bit integer")
    return table


def _phase_census(phase):
    ids, counts = np.unique(phase, return_counts=True)
    total = phase.size
    return [dict(id=int(i), voxels=int(c), fraction=float(c / total)) for i, c in zip(ids, counts)]


def _publish(target, files):
    """Write {name: writer(path)} into a staging directory beside `target`
    and rename it into place; nothing is left behind on failure."""
    target.parent.mkdir(parents=True, exist_ok=True)
    staging = Path(tempfile.mkdtemp(prefix=".vxl-import-", dir=target.parent))
    try:
        for name, write in files.items():
            write(staging / name)
        if target.exists():
            raise FileExistsError(f"Choose a new output directory: {target}")
        os.rename(staging, target)
    except BaseException:
        shutil.rmtree(staging, ignore_errors=True)
        raise


def _json_writer(value):
    text = json.dumps(value, indent=2, allow_nan=False) + "\n"
    return lambda path: path.write_text(text, encoding="utf-8")


def import_model(input_path, materials_path, out, *, axes="zyx", spacing=None, domain_size=None, key="phase_id",
                 orientation_path=None):
    """Convert and validate an external labelled volume; returns the import
    report (the content of import.json).

    axes names the source array's dimensions (a permutation of "xyz");
    spacing (dx, dy, dz) or domain_size (Lx, Ly, Lz) give the physical size,
    else the source NPZ's own domain_size.  orientation_path: separate NPY
    (rotation matrices) or NPZ (orientation, or axis1 + axis2) directions,
    for a source that holds none."""
    source_path, materials_path, target = Path(input_path), Path(materials_path), Path(out).resolve()
    if target.exists():
        raise FileExistsError(f"Choose a new output directory: {target}")
    if not isinstance(axes, str) or len(axes) != 3 or set(axes) != set("xyz"):
        raise ValueError("axes must be a permutation of xyz, such as zyx or xyz")
    if spacing is not None and domain_size is not None:
        raise ValueError("Specify spacing or domain_size, rather than both")

    source = _read_source(source_path, key)
    source.labels = _check_labels(source.labels)
    if orientation_path is not None:
        if source.directions is not None:
            raise ValueError("Orientation is already present in the source NPZ")
        orientation_path = Path(orientation_path)
        source.directions = _read_directions(orientation_path)
    phase, directions = _normalise(source, axes)
    size = _physical_size(phase.shape, spacing, domain_size, source.domain)
    table = _read_materials(materials_path)
    try:
        phase, checked, domain, directions, labels, _, local = validate(phase, table, size, directions)
    except (TypeError, KeyError, OverflowError) as exc:
        raise ValueError(f"Invalid material definition: {exc}") from exc
    if local.any():   # every local frame must be a proper rotation
        flat = phase.reshape(-1)
        mat.check_frames(directions, phase.shape, [mat.voxels_with_label(flat, labels[s]) for s in np.flatnonzero(local)])

    arrays = {"phase_id": phase, "domain_size": domain}
    if isinstance(directions, dict):
        arrays.update(directions)
    elif directions is not None:
        arrays["orientation"] = directions
    provenance = {"volume": source_path, "materials": materials_path}
    if orientation_path is not None:
        provenance["orientation"] = orientation_path
    report = {
        "schema_version": SCHEMA_VERSION,
        "input_axes": axes,
        "input_shape": list(source.labels.shape),
        "axes": NORMAL_AXES,
        "shape": list(phase.shape),
        "domain_size": domain.tolist(),
        "spacing": (domain / np.array(phase.shape[::-1], dtype=float)).tolist(),
        "phases": _phase_census(phase),
        "sources": {role: {"file": str(p.resolve()), "sha256": file_sha256(p)} for role, p in provenance.items()},
    }
    _publish(target, {
        "model.npz": lambda path: np.savez_compressed(path, **arrays),
        "materials.json": _json_writer({str(label): spec for label, spec in checked.items()}),
        "import.json": _json_writer(report),
    })
    return report
