"""Oracle C_eff fixtures for GPU parity tests on cells too slow to solve with
the NumPy oracle inside a GPU test (data only; regenerate on any CPU):

    python tests/golden/make_oracle_fixtures.py [workers]

Writes tests/golden/oracle_ceff.json with, per case, the generator call, the
SHA-256 of the label array (C order, its own dtype) and of the packed
material rows, and the FP64 matrix-free FFT-PCG C_eff of oracle.homogenize
(tolerance 1e-10, so the fixture's own solve error is far below the 1e-7 the
tests allow) with its per-load iterations.

Cases
  * polycrystal48: 48^3 periodic Voronoi, 200 grains, cubic C11/C12/C44 =
    168.4/121.4/75.4 GPa, one Shoemake rotation per grain (seed 1);
  * inclusion_batch7_<k>: every 32^3 cell of random_inclusion_batch(64, seed=7)
    (the 64-cell uneven batch: vf 0.1-0.4, radius 2-5, contrast 10-1000).
"""
import hashlib
import json
import sys
import time
from multiprocessing import Pool
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path[:0] = [str(ROOT), str(ROOT / "voxel-lmfem_amd")]

from oracle import homogenize as oracle  # noqa: E402
from voxel_lmfem import cells  # noqa: E402
from voxel_lmfem import materials as pm  # noqa: E402

TOL = 1e-10


def digest(array):
    a = np.ascontiguousarray(array)
    return hashlib.sha256(a.dtype.str.encode() + str(a.shape).encode() + a.tobytes()).hexdigest()


def case_list():
    out = [("polycrystal48", "voronoi_polycrystal(n=48, grains=200, seed=1)")]
    batch = cells.random_inclusion_batch(64, seed=7)
    out += [(f"inclusion_batch7_{k}", f"random_inclusion_batch(64, seed=7)[{k}]")
            for k, (_, meta) in enumerate(batch) if meta["n"] == 32]
    return out


def build(name):
    if name == "polycrystal48":
        return cells.voronoi_polycrystal(n=48, grains=200, seed=1)[0]
    k = int(name.rsplit("_", 1)[1])
    return cells.random_inclusion_batch(64, seed=7)[k][0]


def solve(item):
    name, call = item
    phase, mats, dom, ori = build(name)
    t = time.time()
    out = oracle.homogenize_matfree(phase, mats, dom, orientation=ori, tol=TOL)
    field = pm.build_field(phase, mats, ori)
    return name, dict(call=call, shape=list(phase.shape), domain_size=[float(v) for v in dom],
                      labels_sha256=digest(phase), rows_sha256=digest(field.rows),
                      CH=np.asarray(out["CH"]).tolist(), tol=TOL,
                      iterations=[c["iterations"] for c in out["calls"]],
                      reference_medium=list(out["reference"]), cpu_seconds=time.time() - t)


def main():
    workers = int(sys.argv[1]) if len(sys.argv) > 1 else 6
    items = case_list()
    with Pool(workers) as pool:
        done = dict(pool.map(solve, items, chunksize=1))
    path = Path(__file__).parent / "oracle_ceff.json"
    path.write_text(json.dumps({name: done[name] for name, _ in items}, indent=1) + "\n")
    print({k: (v["iterations"], round(v["cpu_seconds"], 1)) for k, v in done.items()})


if __name__ == "__main__":
    main()
