"""Batched uneven cells (concurrent solves), the woven mixed-vs-FP64 check,
and the adjoint gradients of C_eff against central finite differences and
the oracle."""
import numpy as np
import pytest

from oracle import fem
from voxel_lmfem import materials as pm

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a, dtype=float), np.asarray(b, dtype=float)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_batch_equals_individual_solves(torch_gpu, tmp_path):
    from voxel_lmfem import homogenize
    from voxel_lmfem.cells import random_inclusion_batch
    from voxel_lmfem.solver import homogenize_batch
    cells = random_inclusion_batch(count=5, seed=11, sizes=(16, 24, 32))
    print("case batch", [m["n"] for _, m in cells], flush=True)
    inputs = [(phase, mats, dom) for (phase, mats, dom, _), _ in cells]
    batch = homogenize_batch(inputs, streams=3, workers=3, autotune=False, stack=1)
    tuned = homogenize_batch(inputs, streams=3, workers=3, tune_cache=tmp_path / "tuning.json")
    assert (tmp_path / "tuning.json").is_file()
    for (phase, mats, dom), got, alt in zip(inputs, batch, tuned):
        one = homogenize(phase, mats, dom)
        assert got["max_true_residual"] <= 1e-8
        assert np.array_equal(np.asarray(got["CH"]), np.asarray(one["CH"]))  # same kernels, same bits
        assert [c["iterations"] for c in got["calls"]] == [c["iterations"] for c in one["calls"]]
        assert got["graph_blocks"] == one["graph_blocks"]
        # autotuned stencil configurations: other summation orders, same certified answer
        assert alt["max_true_residual"] <= 1e-8
        assert rel(alt["CH"], one["CH"]) < 1e-8


def test_batch_graph_replay_equals_single_solves(torch_gpu, monkeypatch):
    """Batch solves with the check blocks replayed from graphs captured before
    the worker threads start (direction ring and fused r update on): the same
    bits and iterations as single solves under the same rules, and blocks
    really replayed."""
    from voxel_lmfem import homogenize
    from voxel_lmfem.cells import random_inclusion_batch
    from voxel_lmfem.solver import homogenize_batch
    monkeypatch.setenv("VXL_GRAPH", "1")
    cells = random_inclusion_batch(count=4, seed=12, sizes=(16, 24, 32))
    print("case batch graph replay", [m["n"] for _, m in cells], flush=True)
    inputs = [(phase, mats, dom) for (phase, mats, dom, _), _ in cells]
    batch = homogenize_batch(inputs, streams=2, workers=2, autotune=False, stack=1)
    assert sum(got["graph_blocks"] for got in batch) > 0
    for (phase, mats, dom), got in zip(inputs, batch):
        one = homogenize(phase, mats, dom)
        assert got["max_true_residual"] <= 1e-8
        assert np.array_equal(np.asarray(got["CH"]), np.asarray(one["CH"]))
        assert [c["iterations"] for c in got["calls"]] == [c["iterations"] for c in one["calls"]]
        assert got["graph_blocks"] == one["graph_blocks"]


def test_woven_128_mixed_vs_fp64(torch_gpu):
    """128^3 woven-like cell, TI tows with per-voxel path tangents: mixed within 1e-6 of FP64."""
    from voxel_lmfem import homogenize
    from voxel_lmfem.cells import woven_like
    (phase, mats, dom, ori), meta = woven_like(128)
    print(f"case woven128 {meta}", flush=True)
    mixed = homogenize(phase, mats, dom, orientation=ori, precision="mixed")
    fp64 = homogenize(phase, mats, dom, orientation=ori, precision="fp64")
    err = rel(mixed["CH"], fp64["CH"])
    print(f"  mixed vs fp64 {err:.2e}, its {[c['iterations'] for c in mixed['calls']]}, "
          f"{mixed['seconds']:.3f}s vs {fp64['seconds']:.3f}s", flush=True)
    assert err < 1e-6
    assert max(mixed["max_true_residual"], fp64["max_true_residual"]) <= 1e-8


def _J(C, W):
    return float((np.asarray(C) * W).sum())


def test_gradients_when_fields_are_solved_again(torch_gpu):
    """keep_fields=False drops the fields after the forward solve and solves
    again in backward: the same gradient, bit for bit (density and angle)."""
    torch = torch_gpu
    from voxel_lmfem import adjoint, materials as pm
    from voxel_lmfem.cells import simp_density, woven_like
    print("case adjoint keep_fields False vs True", flush=True)
    rho, C1, _, _ = simp_density(12, sigma=2.0, seed=5)
    grads = []
    for keep in (True, False):
        r = torch.tensor(rho, dtype=torch.float64, requires_grad=True)
        C = adjoint.density_ceff(r, np.ones(3), C1=C1, keep_fields=keep)
        (C[0, 0] + C[:3, :3].sum() / 9.0).backward()
        grads.append(r.grad.numpy().copy())
    assert np.array_equal(grads[0], grads[1])
    (phase, mats, dom, ori), _ = woven_like(16, section=(2, 4), amplitude=1.0)
    labels, tensors, local = pm.phase_table(mats)
    slot = np.searchsorted(labels, phase.reshape(-1))
    mask = local[slot]
    idx = np.flatnonzero(mask)
    frames = pm.frames(ori, phase.shape, idx)
    iso = pm.pack21(tensors[slot])
    grads = []
    for keep in (True, False):
        theta = torch.zeros(phase.shape, dtype=torch.float64, requires_grad=True)
        C = adjoint.angle_ceff(theta, dom, tensors[slot[idx[0]]], frames, iso, mask, keep_fields=keep)
        (C[0, 0] + C[:3, :3].sum() / 9.0).backward()
        grads.append(theta.grad.numpy().copy())
    assert np.array_equal(grads[0], grads[1]) and grads[0].any()


@pytest.mark.parametrize("n,voxels", [(12, 6), (96, 20)])
def test_density_gradient_matches_central_differences(torch_gpu, n, voxels):
    """SIMP density cell, J = C_1111 + K_V(C_eff): adjoint vs central FD (step 1e-4) at seeded voxels."""
    torch = torch_gpu
    from voxel_lmfem import adjoint
    from voxel_lmfem.cells import simp_density
    rho, C1, E, dE = simp_density(n, sigma=4.0 if n >= 64 else 2.0, seed=5)
    W = adjoint.objective_weights()
    r = torch.tensor(rho, dtype=torch.float64, requires_grad=True)
    C = adjoint.density_ceff(r, np.ones(3), C1=C1)
    J = C[0, 0] + (C[:3, :3].sum()) / 9.0
    J.backward()
    g = r.grad.numpy().reshape(-1)
    assert float(J.detach()) == pytest.approx(_J(C.detach().numpy(), W), rel=1e-14)
    rng = np.random.default_rng(n)
    picks = rng.choice(rho.size, size=voxels, replace=False)
    step = 1e-4
    worst = 0.0
    for e in picks:
        vals = []
        for sgn in (1.0, -1.0):
            rr = rho.reshape(-1).copy()
            rr[e] = rr[e] + sgn * step
            Cp = adjoint.density_ceff(torch.tensor(rr.reshape(rho.shape)), np.ones(3), C1=C1)
            vals.append(_J(Cp.numpy(), W))
        fd = (vals[0] - vals[1]) / (2 * step)
        scale = max(abs(fd), 1e-3 * np.abs(g).max())
        worst = max(worst, abs(fd - g[e]) / scale)
    print(f"  case density n={n}: worst FD mismatch {worst:.2e} over {voxels} voxels", flush=True)
    assert worst < 1e-3
    if n ** 3 <= 4096:  # the oracle's sparse-direct solve and energy contraction
        Ev = 1e-3 + rho.reshape(-1) ** 3 * (70 - 1e-3)
        Chat = np.broadcast_to(C1, (n ** 3, 6, 6))
        h = np.ones(3) / n
        _, _, u, CH, _ = fem.assemble(Ev[:, None, None] * Chat, (n, n, n), h)
        gref = (3 * rho.reshape(-1) ** 2 * (70 - 1e-3)) * fem.energy_contraction(Chat, (n, n, n), h, u, np.eye(6), W)
        assert rel(C.detach().numpy(), CH) < 1e-9
        assert rel(g, gref) < 1e-7


def test_angle_gradient_matches_central_differences(torch_gpu):
    """Fiber angle of the tow voxels of a woven copy: adjoint vs central FD."""
    torch = torch_gpu
    from voxel_lmfem import adjoint
    from voxel_lmfem.cells import woven_like
    (phase, mats, dom, ori), _ = woven_like(32, section=(4, 8), amplitude=3.0)
    labels, tensors, local = pm.phase_table(mats)
    slot = np.searchsorted(labels, phase.reshape(-1))
    mask = local[slot]
    assert len(set(int(s) for s in slot[mask])) == 2   # warp and weft share one tow law
    idx = np.flatnonzero(mask)
    frames = pm.frames(ori, phase.shape, idx)
    iso = pm.pack21(tensors[slot])
    C_loc = tensors[slot[idx[0]]]
    W = adjoint.objective_weights()
    theta = torch.zeros(phase.shape, dtype=torch.float64, requires_grad=True)
    C = adjoint.angle_ceff(theta, dom, C_loc, frames, iso, mask)
    field = pm.build_field(phase, mats, ori)
    ref = pm.unpack21(field.rows[field.index.astype(np.int64)])
    rows0, _ = adjoint.angle_rows(np.zeros(phase.size), C_loc, frames, iso, mask)
    np.testing.assert_allclose(pm.unpack21(rows0), ref, rtol=1e-13, atol=1e-12)  # theta = 0 is the cell
    J = (torch.as_tensor(W) * C).sum()
    J.backward()
    g = theta.grad.numpy().reshape(-1)
    assert not g[~mask].any()
    rng = np.random.default_rng(3)
    step = 1e-4
    worst = 0.0
    for e in rng.choice(idx, size=6, replace=False):
        vals = []
        for sgn in (1.0, -1.0):
            t = np.zeros(phase.size)
            t[e] = sgn * step
            Cp = adjoint.angle_ceff(torch.tensor(t.reshape(phase.shape)), dom, C_loc, frames, iso, mask)
            vals.append(_J(Cp.numpy(), W))
        fd = (vals[0] - vals[1]) / (2 * step)
        scale = max(abs(fd), 1e-3 * np.abs(g).max())
        worst = max(worst, abs(fd - g[e]) / scale)
    print(f"  case angle: worst FD mismatch {worst:.2e}", flush=True)
    assert worst < 1e-3
