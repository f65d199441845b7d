"""The configured workloads against the FP64 CPU oracle.

* 200-grain Voronoi polycrystal (rotated cubic C11/C12/C44 = 168.4/121.4/75.4):
  32^3 against oracle.homogenize_matfree run here, 48^3 against the oracle's
  committed C_eff (tests/golden/oracle_ceff.json), mixed and FP64, with the
  device label bookkeeping bit-exact against the host's and the oracle's;
* the 256^3 polycrystal itself: mixed within 1e-6 of the FP64 GPU solve;
* the 64-cell uneven inclusion batch (edges 32-96, contrast 10-1000):
  every certificate, every 32^3 cell against the oracle's committed C_eff,
  one cell per other edge against its single solve;
* the fiber-angle gradient on a 64^3 woven copy at 20 seeded voxels against
  central differences (step 1e-4).
"""
import hashlib
import json

import numpy as np
import pytest

from conftest import GOLDEN
from oracle import homogenize as oracle, materials as om

pytestmark = pytest.mark.gpu


def fixture(name):
    """One case of the oracle's committed C_eff (tests/golden/make_oracle_fixtures.py)."""
    return json.loads((GOLDEN / "oracle_ceff.json").read_text())[name]


def rel(a, b):
    a, b = np.asarray(a, dtype=float), np.asarray(b, dtype=float)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def digest(array):
    a = np.ascontiguousarray(array)
    return hashlib.sha256(a.dtype.str.encode() + str(a.shape).encode() + a.tobytes()).hexdigest()


def device_bookkeeping_matches(phase, mats, pool, index):
    """Device label -> row index and row table equal the host's and the
    oracle's pool (hybrid_material.py:38-61 order), bit for bit."""
    import torch
    from voxel_lmfem import materials as pm
    host = pm.build_field(phase, mats)
    dev = pm.build_field_device(torch.as_tensor(phase.astype(np.int32), device="cuda"), mats)
    idx = dev.index_dev.cpu().numpy().view(host.index.dtype)
    assert np.array_equal(idx, host.index)
    assert np.array_equal(dev.rows, host.rows)
    assert np.array_equal(host.rows, pool) and np.array_equal(host.index.astype(np.int32), index)


@pytest.mark.parametrize("n", [32, 48])
def test_polycrystal_vs_oracle(torch_gpu, n):
    from voxel_lmfem import homogenize
    from voxel_lmfem.cells import voronoi_polycrystal
    (phase, mats, dom, _), _ = voronoi_polycrystal(n=n, grains=200, seed=1)
    print(f"case polycrystal {n}^3 200 grains vs oracle", flush=True)
    pool, index = om.hybrid_pool(phase, mats)
    device_bookkeeping_matches(phase, mats, pool, index)
    if n == 32:
        ref = oracle.homogenize_matfree(phase, mats, dom)["CH"]
    else:
        fx = fixture("polycrystal48")
        assert fx["labels_sha256"] == digest(phase)
        ref = fx["CH"]
    for precision in ("mixed", "fp64"):
        out = homogenize(phase, mats, dom, precision=precision)
        err = rel(out["CH"], ref)
        print(f"  {precision}: rel err vs oracle {err:.2e}, its {[c['iterations'] for c in out['calls']]}",
              flush=True)
        assert err < 1e-7
        assert out["max_true_residual"] <= 1e-8


def test_polycrystal_256_mixed_vs_fp64(torch_gpu):
    """The bench workload in a test: 256^3, 200 grains, mixed within 1e-6 of
    the FP64 GPU solve, both certified at 1e-8."""
    from voxel_lmfem import homogenize
    from voxel_lmfem.cells import voronoi_polycrystal
    (phase, mats, dom, _), _ = voronoi_polycrystal(n=256, grains=200, seed=3)
    print("case polycrystal 256^3 mixed vs fp64", flush=True)
    mixed = homogenize(phase, mats, dom, precision="mixed")
    fp64 = homogenize(phase, mats, dom, precision="fp64")
    err = rel(mixed["CH"], fp64["CH"])
    print(f"  mixed vs fp64 {err:.2e}; its {[c['iterations'] for c in mixed['calls']]}; "
          f"{mixed['seconds']:.3f}s vs {fp64['seconds']:.3f}s", flush=True)
    assert err < 1e-6
    assert max(mixed["max_true_residual"], fp64["max_true_residual"]) <= 1e-8
    C = np.asarray(mixed["CH"])
    assert rel(C, C.T) < 1e-8   # unsymmetrized CH of a symmetric operator


def test_inclusion_batch_64_cells(torch_gpu, tmp_path):
    from voxel_lmfem import homogenize
    from voxel_lmfem.cells import random_inclusion_batch
    from voxel_lmfem.solver import homogenize_batch
    cells = random_inclusion_batch(64, seed=7)
    edges = [m["n"] for _, m in cells]
    print(f"case inclusion batch 64 cells, {sum(n ** 3 for n in edges)} voxels", flush=True)
    inputs = [(phase, mats, dom) for (phase, mats, dom, _), _ in cells]
    out = homogenize_batch(inputs, tune_cache=tmp_path / "tuning.json")
    assert len(out) == 64
    for k, res in enumerate(out):
        assert res["max_true_residual"] <= 1e-8, k
        assert res["grid_shape"] == [edges[k]] * 3
    checked = 0
    for k, n in enumerate(edges):
        if n != 32:
            continue
        fx = fixture(f"inclusion_batch7_{k}")
        assert fx["labels_sha256"] == digest(inputs[k][0])
        err = rel(out[k]["CH"], fx["CH"])
        print(f"  cell {k} (32^3): rel err vs oracle {err:.2e}", flush=True)
        assert err < 1e-7
        checked += 1
    assert checked == edges.count(32) > 0
    for n in sorted(set(edges) - {32}):
        k = edges.index(n)
        one = homogenize(*inputs[k])
        err = rel(out[k]["CH"], one["CH"])
        print(f"  cell {k} ({n}^3): batch vs single solve {err:.2e}", flush=True)
        assert err < 1e-8


def test_angle_gradient_woven_64(torch_gpu):
    """C5 angle leg: fiber angles of the tow voxels of the woven cell at 64^3
    (the 128^3 geometry at half resolution), J = C_1111 + K_V(C_eff):
    adjoint against central differences at 20 seeded tow voxels."""
    torch = torch_gpu
    from voxel_lmfem import adjoint, materials as pm
    from voxel_lmfem.cells import woven_like
    (phase, mats, dom, ori), _ = woven_like(64, section=(2, 6), amplitude=3.0)
    print("case angle gradient woven 64^3, 20 voxels", flush=True)
    labels, tensors, local = pm.phase_table(mats)
    slot = np.searchsorted(labels, phase.reshape(-1))
    mask = local[slot]
    idx = np.flatnonzero(mask)
    frames = pm.frames(ori, phase.shape, idx)
    iso = pm.pack21(tensors[slot])
    C_loc = tensors[slot[idx[0]]]
    W = adjoint.objective_weights()
    theta = torch.zeros(phase.shape, dtype=torch.float64, requires_grad=True)
    C = adjoint.angle_ceff(theta, dom, C_loc, frames, iso, mask)
    J = (torch.as_tensor(W) * C).sum()
    J.backward()
    g = theta.grad.numpy().reshape(-1)
    assert not g[~mask].any()
    rng = np.random.default_rng(64)
    step = 1e-4
    worst = 0.0
    for e in rng.choice(idx, size=20, replace=False):
        vals = []
        for sgn in (1.0, -1.0):
            t = np.zeros(phase.size)
            t[e] = sgn * step
            Cp = adjoint.angle_ceff(torch.tensor(t.reshape(phase.shape)), dom, C_loc, frames, iso, mask)
            vals.append(float((np.asarray(Cp.numpy()) * W).sum()))
        fd = (vals[0] - vals[1]) / (2 * step)
        scale = max(abs(fd), 1e-3 * np.abs(g).max())
        worst = max(worst, abs(fd - g[e]) / scale)
    print(f"  worst FD mismatch {worst:.2e} over 20 voxels", flush=True)
    assert worst < 1e-3
