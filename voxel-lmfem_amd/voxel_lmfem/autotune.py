"""Stencil autotuner with a persistent JSON cache.

The stencil has three run-time knobs per precision (all compute the same
element sums): the workgroup tile (32 x tile_y output nodes), the schedule
(1: per-layer 8-node transform, 2: persistent face columns, 3: 2 with the
plane after next prefetched) and the fewest z planes per workgroup chunk.
``vxl_autotune_stencil`` times the 27 combinations on a cell; this module
keeps the winners in a JSON file keyed by everything the timing depends on
(grid, precision, load count, material index width, per-voxel scale, device
name), so a shape is tuned once per machine.  The reference records its own
operator-variant choice in ``_core/hybrid_operator/recommended.json:1-6``
(written by ``hybrid_operator/study.py:64-82``); this is the measured
counterpart for this build's stencil.

Cache file: $VXL_TUNE_CACHE, else ~/.cache/voxel_lmfem/stencil_tuning.json.
"""
import json
import os
import tempfile
from pathlib import Path

SCHEMA = 1


def cache_path(path=None):
    if path is not None:
        return Path(path)
    env = os.environ.get("VXL_TUNE_CACHE")
    if env:
        return Path(env)
    return Path.home() / ".cache" / "voxel_lmfem" / "stencil_tuning.json"


def load(path=None):
    """Entries of the cache ({} when absent or unreadable)."""
    p = cache_path(path)
    try:
        data = json.loads(p.read_text(encoding="utf-8"))
    except (OSError, ValueError):
        return {}
    if not isinstance(data, dict) or data.get("schema") != SCHEMA or not isinstance(data.get("entries"), dict):
        return {}
    return data["entries"]


def save(entries, path=None, *, strict=True):
    """Atomically replace the cache file with `entries`.  strict=False turns a
    failure to write (read-only home, full disk) into a warning: the tuned
    configurations still apply to the cells of this process."""
    try:
        _write(entries, path)
    except OSError as exc:
        if strict:
            raise
        import warnings
        warnings.warn(f"stencil tuning cache not written: {exc}", RuntimeWarning, stacklevel=2)


def _write(entries, path):
    p = cache_path(path)
    p.parent.mkdir(parents=True, exist_ok=True)
    text = json.dumps(dict(schema=SCHEMA, entries=entries), indent=1, sort_keys=True) + "\n"
    fd, tmp = tempfile.mkstemp(prefix=".stencil_tuning-", dir=p.parent)
    try:
        with os.fdopen(fd, "w", encoding="utf-8") as f:
            f.write(text)
        os.replace(tmp, p)
    except BaseException:
        Path(tmp).unlink(missing_ok=True)
        raise


TILES = {"fp32": (14, 8, 6), "fp64": (6, 8, 4)}   # table axis 0, in vxl_autotune_stencil's order
DEPTHS = (8, 4, 2)                                # table axis 2


def _table_ms(table, precision, tile_y, schedule, zmin):
    """Milliseconds of one configuration in the 3 x 3 x 3 autotune table
    (NaN if the configuration is not one of the tuned ones)."""
    tiles = TILES["fp64" if precision in ("float64", "fp64") else "fp32"]
    if tile_y not in tiles or zmin not in DEPTHS or not 1 <= schedule <= 3:
        return float("nan")
    return float(table[tiles.index(tile_y), schedule - 1, DEPTHS.index(zmin)])


def key(cell, precision, nrhs):
    """Cache key: device, precision, grid (z, y, x), loads, index width, scale."""
    prec = "fp64" if precision in ("float64", "fp64") else "fp32"
    nz, ny, nx = cell.shape
    width = {1: "u8", 2: "u16", 4: "i32"}.get(int(cell.index.element_size()), "?")
    device = cell.torch.cuda.get_device_name(cell.device)
    scaled = "scaled" if cell.scale is not None else "plain"
    return f"{device}|{prec}|{nz}x{ny}x{nx}|r{int(nrhs)}|{width}|{scaled}"


def tune(cell, precision="float32", nrhs=6, *, path=None, reps=3, refresh=False, entries=None):
    """Apply the cached configuration for this cell, or tune and record it.

    entries: an already loaded cache dict to consult and update in place (the
    caller saves it); otherwise the file is read and, after a new tuning,
    rewritten.  Returns the entry applied ({"tile_y", "schedule", "zmin",
    "ms", "tuned"})."""
    own = entries is None
    if own:
        entries = load(path)
    k = key(cell, precision, nrhs)
    hit = entries.get(k)
    if hit and not refresh:
        try:
            cell.set_stencil_config(precision, hit["tile_y"], hit["schedule"], hit["zmin"])
            return dict(hit, tuned=False)
        except (KeyError, TypeError, ValueError):
            pass  # stale or foreign entry: tune again
    before = cell.stencil_config(precision)   # the configuration the cell runs untuned
    (tile_y, schedule, zmin), table = cell.autotune(precision, nrhs=nrhs, reps=reps)
    entry = dict(tile_y=int(tile_y), schedule=int(schedule), zmin=int(zmin), ms=float(table.min()),
                 default_ms=_table_ms(table, precision, *before))
    entries[k] = entry
    if own:
        save(entries, path, strict=False)
    return dict(entry, tuned=True)
