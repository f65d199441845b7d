"""The drop-in boundary on the GPU: the `solve --backend cuda` command, the
result's backend key, the uncertified (max_iter) path, and the reference's
(1, 3N) operator protocol against the oracle."""
import json

import numpy as np
import pytest

from oracle import fem, homogenize as oracle, materials as om, pcg as opcg, spectral
from test_oracle import gpu_fixture_cell

pytestmark = pytest.mark.gpu

RESULT_KEYS = {"seconds", "CH", "calls", "stages", "field_statistics", "peak_allocated_bytes",
               "peak_reserved_bytes", "backend", "storage", "operator", "scope", "max_true_residual",
               "elapsed_seconds", "voigt_order"}


def rel(a, b):
    a, b = np.asarray(a, dtype=float), np.asarray(b, dtype=float)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def two_phase(tmp_path):
    rng = np.random.default_rng(21)
    phase = (rng.random((6, 5, 4)) < 0.35).astype(np.uint8)
    table = {"0": dict(kind="isotropic", E=3000.0, nu=0.35), "1": dict(kind="isotropic", E=70000.0, nu=0.22)}
    np.savez(tmp_path / "model.npz", phase_id=phase, domain_size=np.array([1.0, 1.25, 1.5]))
    (tmp_path / "materials.json").write_text(json.dumps(table))
    return phase, {int(k): v for k, v in table.items()}, np.array([1.0, 1.25, 1.5])


def test_cli_solve_backend_cuda(torch_gpu, tmp_path, capsys):
    """`solve --backend cuda` (run in this process: the CLI's main) writes the
    reference's GPU result keys, prints {output, backend, max_true_residual},
    and refuses to overwrite its output (exit status 2)."""
    from voxel_lmfem.cli import main
    with capsys.disabled():  # stdout is the command's own output here
        print("case cli solve --backend cuda 6x5x4", flush=True)
    phase, table, domain = two_phase(tmp_path)
    out = tmp_path / "result.json"
    argv = ["solve", "--input", str(tmp_path / "model.npz"), "--materials", str(tmp_path / "materials.json"),
            "--output", str(out), "--backend", "cuda", "--warmup"]
    assert main(argv) == 0
    printed = json.loads(capsys.readouterr().out)
    assert printed == dict(output=str(out), backend="cuda", max_true_residual=printed["max_true_residual"])
    result = json.loads(out.read_text())
    assert set(result) >= RESULT_KEYS
    assert result["backend"] == "cuda" and result["voigt_order"] == ["11", "22", "33", "23", "13", "12"]
    assert len(result["calls"]) == 6 and result["max_true_residual"] <= 1e-8
    ref = oracle.homogenize_direct(phase, table, domain)
    assert rel(result["CH"], ref["CH"]) < 1e-8
    before = out.read_bytes()
    with pytest.raises(SystemExit) as stop:
        main(argv)
    assert stop.value.code == 2
    assert "error: " in capsys.readouterr().err
    assert out.read_bytes() == before


def test_backend_key_repeats_the_callers_spelling(torch_gpu):
    from voxel_lmfem import homogenize
    phase = np.zeros((2, 2, 3), dtype=np.int32)
    mats = {0: dict(kind="isotropic", E=10.0, nu=0.3)}
    assert homogenize(phase, mats, [1, 1, 1])["backend"] == "cuda"
    assert homogenize(phase, mats, [1, 1, 1], backend="hip")["backend"] == "hip"
    with pytest.raises(ValueError, match="backend must be"):
        homogenize(phase, mats, [1, 1, 1], backend="rocm")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_uncertified_solve_raises_and_keeps_the_last_iterate(torch_gpu, precision):
    """max_iter=1 forces the uncertified path (the reference's
    chain_study.py:95-102): homogenize raises RuntimeError, and the solver's
    error carries the iterate it stopped at -- in FP64 exactly one PCG step of
    the oracle from u = 0 -- not an unwritten buffer."""
    import torch
    from voxel_lmfem import homogenize, materials as pm
    from voxel_lmfem._native import ConvergenceError
    from voxel_lmfem.gpu import Cell, Solver
    print(f"case uncertified {precision}", flush=True)
    phase, table, domain, rot = gpu_fixture_cell()
    domain = np.asarray(domain, dtype=float)
    with pytest.raises(RuntimeError):
        homogenize(phase, table, domain, orientation=rot, precision=precision, max_iter=1)
    with pytest.raises(ValueError, match="max_iter"):
        homogenize(phase, table, domain, orientation=rot, max_iter=0)
    before = torch.cuda.memory_allocated()
    field = pm.build_field(phase, table, rot)
    h = domain / np.array(phase.shape[::-1], dtype=float)
    lam, mu = oracle.reference_lame(phase, table, rot)
    cell = Cell(field, phase.shape, h)
    solver = Solver(cell, lam, mu, mode=precision, nrhs=6)
    with pytest.raises(ConvergenceError) as err:
        solver.run(tol=1e-8, max_iter=1, return_fields=True)
    partial = err.value.partial
    assert partial["iterations"] == [1] * 6
    u = partial["fields"].cpu().numpy()
    assert np.isfinite(u).all()
    full = om.unpack21(field.rows[field.index.astype(np.int64)])
    op = fem.MatrixFree(full, phase.shape, h)
    for j in range(6):
        rhs = op.project_zero_mean(op.rhs(np.eye(6)[j]))
        one, _, _ = opcg.pcg(op.apply, lambda r: spectral.green_apply(r, phase.shape, h, lam, mu), rhs,
                             tol=1e-30, max_iter=1)
        got = u[j].T   # (3, N) planes -> (N, 3)
        err_j = rel(got, op.project_zero_mean(one))
        print(f"  load {j}: one-step iterate rel err {err_j:.2e}", flush=True)
        assert err_j < (1e-10 if precision == "fp64" else 1e-5)
    solver.close()
    cell.close()
    del partial, u, err
    assert torch.cuda.memory_allocated() <= before + (1 << 20)


def _small_cell():
    """8 x 6 x 10 (z, y, x) rotated-grain cell with non-cubic voxels."""
    from voxel_lmfem.cells import voronoi_polycrystal
    (phase, mats, _, _), _ = voronoi_polycrystal(n=8, grains=5, seed=3)
    phase = np.concatenate([phase[:, :6, :], phase[:, :6, :2]], axis=2)  # (8, 6, 10)
    return phase, mats, np.array([1.0, 0.7, 1.3])


def test_operator_protocol_matches_oracle(torch_gpu):
    """Cell as the reference's HybridFIOperator (hybrid_adapter.py:39-125):
    B, mesh, geometry, apply, compute_rhs, compute_avg_stress,
    project_zero_mean, _compute_reference_medium on (1, 3N) FP64 node-major
    tensors, against the oracle's Gauss-point operator."""
    import torch
    from voxel_lmfem import materials as pm
    from voxel_lmfem.gpu import Cell
    from oracle.reference_medium import zeman
    print("case operator protocol 8x6x10", flush=True)
    phase, mats, domain = _small_cell()
    shape = phase.shape
    h = domain / np.array(shape[::-1], dtype=float)
    field = pm.build_field(phase, mats)
    cell = Cell(field, shape, h)
    n = phase.size
    assert cell.B == 1 and cell.geometry() == (8, 6, 10, h[0], h[1], h[2])
    assert (cell.mesh.n_nodes, cell.mesh.n_dofs, cell.mesh.V_elem) == (n, 3 * n, float(np.prod(h)))
    pool, index = om.hybrid_pool(phase, mats)
    op = fem.MatrixFree(om.unpack21(pool[index]), shape, h)
    rng = np.random.default_rng(4)
    u = rng.normal(size=(n, 3))
    ud = torch.as_tensor(u.reshape(1, -1), device="cuda")
    Ku = cell.apply(ud).cpu().numpy().reshape(n, 3)
    ref = op.apply(u)
    assert np.abs(Ku - ref).max() <= 1e-13 * np.abs(ref).max()
    for j in range(6):
        macro = np.eye(6)[j] * (1.0 + j)
        f = cell.compute_rhs(macro).cpu().numpy().reshape(n, 3)
        fr = op.rhs(macro)
        assert np.abs(f - fr).max() <= 1e-13 * np.abs(fr).max()
        s = cell.compute_avg_stress(ud, macro)
        assert s.shape == (1, 6)
        np.testing.assert_allclose(s[0], op.avg_stress(u, macro), rtol=1e-12, atol=1e-12 * np.abs(s).max())
    centred = cell.project_zero_mean(ud).cpu().numpy().reshape(n, 3)
    np.testing.assert_allclose(centred, op.project_zero_mean(u), rtol=0, atol=1e-14 * np.abs(u).max())
    medium = cell._compute_reference_medium("zeman")
    assert medium.shape == (1, 6, 6)
    np.testing.assert_allclose(medium[0], zeman(lambda pos: pool[index[pos]], shape), rtol=1e-14, atol=0)
    for bad in (ud.float(), ud[:, :-3].contiguous(), ud.view(n, 3), ud.cpu()):
        with pytest.raises(ValueError, match="Expected contiguous FP64 displacement"):
            cell.apply(bad)
    with pytest.raises(ValueError, match="Six engineering macro strain"):
        cell.compute_rhs(np.ones(5))
    cell.close()


@pytest.mark.parametrize("precision", ["float64", "float32"])
def test_preconditioner_protocol_matches_oracle(torch_gpu, precision):
    """apply_precond (FP64 AoS in and out) and precond_packed (raw component
    planes, chain_adapter.py:17-31) against the oracle's Green operator."""
    import torch
    from voxel_lmfem import materials as pm
    from voxel_lmfem.gpu import Cell
    print(f"case preconditioner protocol {precision}", flush=True)
    phase, mats, domain = _small_cell()
    shape = phase.shape
    h = domain / np.array(shape[::-1], dtype=float)
    cell = Cell(pm.build_field(phase, mats), shape, h)
    with pytest.raises(RuntimeError, match="enable_preconditioner"):
        cell.apply_precond(torch.zeros(1, 3 * phase.size, dtype=torch.float64, device="cuda"))
    lam, mu = 90.0, 35.0
    cell.enable_preconditioner(lam, mu, precision=precision)
    rng = np.random.default_rng(8)
    r = rng.normal(size=(phase.size, 3))
    ref = spectral.green_apply(r, shape, h, lam, mu)
    got = cell.apply_precond(torch.as_tensor(r.reshape(1, -1), device="cuda")).cpu().numpy().reshape(-1, 3)
    tol = 1e-11 if precision == "float64" else 5e-6
    assert np.abs(got - ref).max() <= tol * np.abs(ref).max()
    dtype = torch.float64 if precision == "float64" else torch.float32
    planes = torch.as_tensor(np.stack([r.T, 2 * r.T]).reshape(2, 3, *shape), dtype=dtype, device="cuda")
    keep = planes.clone()
    packed = cell.precond_packed(planes)
    assert packed.shape == planes.shape and packed.dtype == dtype
    assert torch.equal(planes, keep)   # the input is not consumed
    z = packed.cpu().numpy().astype(np.float64).reshape(2, 3, -1)
    assert np.abs(z[0].T - ref).max() <= tol * np.abs(ref).max()
    assert np.abs(z[1].T - 2 * ref).max() <= 2 * tol * np.abs(ref).max()
    with pytest.raises(ValueError):
        cell.precond_packed(planes.reshape(2, 3, -1))
    cell.close()


def test_side_stream_cell_orders_against_the_current_stream(torch_gpu):
    """A cell on a stream of its own (homogenize_batch builds them so): inputs
    produced on the current stream right before each call and outputs read
    there right after give the default-stream cell's results bit for bit
    (apply, RHS, preconditioner, solver fields)."""
    import torch
    from voxel_lmfem import materials as pm
    from voxel_lmfem.gpu import Cell, Preconditioner, Solver
    print("case side-stream cell 32^3", flush=True)
    rng = np.random.default_rng(12)
    phase = (rng.random((32, 32, 32)) < 0.3).astype(np.uint8)
    mats = {0: dict(kind="isotropic", E=3000.0, nu=0.35), 1: dict(kind="isotropic", E=70000.0, nu=0.22)}
    h = np.full(3, 1.0 / 32)
    field = pm.build_field(phase, mats)
    side = torch.cuda.Stream()
    ref_cell = Cell(field, phase.shape, h)
    cell = Cell(field, phase.shape, h, stream=side)
    n = phase.size
    for dtype in (torch.float32, torch.float64):
        for _ in range(3):
            u = torch.randn(6, 3, n, dtype=dtype, device="cuda")
            for _ in range(20):  # a chain of producer kernels on the current stream
                u = u * 1.0001 + 0.5
            assert torch.equal(cell.apply_planes(u), ref_cell.apply_planes(u))
    assert torch.equal(cell.rhs_planes(np.eye(6)), ref_cell.rhs_planes(np.eye(6)))
    pc, ref_pc = Preconditioner(cell, 8000.0, 9000.0), Preconditioner(ref_cell, 8000.0, 9000.0)
    r = torch.randn(6, 3, n, dtype=torch.float32, device="cuda") * 3.0 + 1.0
    assert torch.equal(pc.apply_planes(r), ref_pc.apply_planes(r))
    pc.close()
    ref_pc.close()
    solver, ref_solver = Solver(cell, 8000.0, 9000.0), Solver(ref_cell, 8000.0, 9000.0)
    got = solver.run(return_fields=True)
    ref = ref_solver.run(return_fields=True)
    assert torch.equal(got["fields"], ref["fields"]) and np.array_equal(got["CH"], ref["CH"])
    assert max(got["true_residuals"]) <= 1e-8
    for obj in (solver, ref_solver, cell, ref_cell):
        obj.close()
