"""Stacked cells (vxl_cell_set_groups, Cell.stack): several cells of one grid
and voxel size in every launch, six loads each, each with its own material
and reference medium.  Operator, RHS and mean stress equal the single cells'
bit for bit; solves certify and match the single solves and the oracle;
homogenize_batch stacks same-shape cells; refusals."""
import numpy as np
import pytest

from oracle import homogenize as oracle
from voxel_lmfem import materials as pm

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a, dtype=float), np.asarray(b, dtype=float)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def three_cells(n=16, seed=4):
    """Same grid, different labels, materials and contrasts (the last one
    with a third label, so the row tables differ in length)."""
    from voxel_lmfem.cells import periodic_spheres
    out = []
    for k, (E1, frac) in enumerate(((30.0, .2), (300.0, .3), (5.0, .25))):
        (phase, mats, dom, _), _ = periodic_spheres(n, radius=2.5, fraction=frac, seed=seed + k, E=(1.0, E1))
        if k == 2:
            phase = phase.copy()
            phase[: n // 4] += 2
            mats = dict(mats)
            mats[2] = dict(kind="isotropic", E=2.0, nu=.3)
            mats[3] = dict(kind="isotropic", E=9.0, nu=.25)
        out.append((phase, mats, dom))
    return out


def build(cells):
    from voxel_lmfem.reference_medium import select
    fields, media = [], []
    for phase, mats, dom in cells:
        field = pm.build_field(phase, mats)
        lam, mu, _ = select(field, phase, mats)
        fields.append(field)
        media.append((lam, mu))
    return fields, media


def test_stacked_operator_rhs_stress_bitwise(torch_gpu):
    torch = torch_gpu
    from voxel_lmfem.gpu import Cell
    cells = three_cells()
    fields, _ = build(cells)
    shape, h = cells[0][0].shape, (1 / 16,) * 3
    print("case stacked operator, 3 cells of 16^3", flush=True)
    singles = [Cell(f, shape, h) for f in fields]
    stacked = Cell.stack(fields, shape, h)
    try:
        n = singles[0].n
        rng = np.random.default_rng(5)
        for dtype in (torch.float64, torch.float32):
            u = torch.tensor(rng.normal(size=(12, 3, n)), dtype=dtype, device="cuda")  # 4 fields per cell
            got = stacked.apply_planes(u).cpu()
            for g, c in enumerate(singles):
                want = c.apply_planes(u[4 * g: 4 * g + 4]).cpu()
                assert torch.equal(got[4 * g: 4 * g + 4], want), (dtype, g)
        macro = np.tile(np.eye(6), (3, 1))
        f = stacked.rhs_planes(macro).cpu()
        u64 = torch.tensor(rng.normal(size=(18, 3, n)), dtype=torch.float64, device="cuda")
        sig = stacked.avg_stress_planes(u64, macro)
        for g, c in enumerate(singles):
            assert torch.equal(f[6 * g: 6 * g + 6], c.rhs_planes(np.eye(6)).cpu())
            assert np.array_equal(sig[6 * g: 6 * g + 6], c.avg_stress_planes(u64[6 * g: 6 * g + 6], np.eye(6)))
    finally:
        for c in singles:
            c.close()
        stacked.close()


@pytest.mark.parametrize("mode", ["mixed", "fp64"])
def test_stacked_solve_matches_single_and_oracle(torch_gpu, mode):
    from voxel_lmfem.gpu import Cell, Solver
    cells = three_cells()
    fields, media = build(cells)
    shape, h = cells[0][0].shape, (1 / 16,) * 3
    print(f"case stacked solve {mode}", flush=True)
    stacked = Cell.stack(fields, shape, h)
    try:
        sv = Solver(stacked, [m[0] for m in media], [m[1] for m in media], mode=mode)
        try:
            assert sv.nrhs == 18
            out = sv.run()
        finally:
            sv.close()
    finally:
        stacked.close()
    assert max(out["true_residuals"]) <= 1e-8
    for g, ((phase, mats, dom), field, (lam, mu)) in enumerate(zip(cells, fields, media)):
        cell = Cell(field, shape, h)
        try:
            s1 = Solver(cell, lam, mu, mode=mode)
            try:
                one = s1.run()
            finally:
                s1.close()
        finally:
            cell.close()
        ch = np.asarray(out["CH"])[:, 6 * g: 6 * g + 6]
        ref = oracle.homogenize_matfree(phase, mats, dom)["CH"]
        print(f"  cell {g}: its stacked {out['iterations'][6 * g: 6 * g + 6]} single {one['iterations']}, "
              f"vs single {rel(ch, one['CH']):.1e}, vs oracle {rel(ch, ref):.1e}", flush=True)
        assert rel(ch, one["CH"]) < 1e-9
        assert rel(ch, ref) < 1e-7


def test_stacked_graph_replay_equals_eager(torch_gpu, monkeypatch):
    """Check blocks of a stacked solve replayed from a captured graph (ring and
    fused update on): the same bits and iterations as launched kernel by
    kernel, and blocks really replayed."""
    from voxel_lmfem.gpu import Cell, Solver
    cells = three_cells()
    fields, media = build(cells)
    shape, h = cells[0][0].shape, (1 / 16,) * 3
    print("case stacked graph replay", flush=True)
    outs = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("VXL_GRAPH", mode)
        stacked = Cell.stack(fields, shape, h)
        try:
            sv = Solver(stacked, [m[0] for m in media], [m[1] for m in media])
            try:
                outs[mode] = sv.run()
            finally:
                sv.close()
        finally:
            stacked.close()
    assert outs["1"]["graph_blocks"] > 0 and outs["0"]["graph_blocks"] == 0
    assert np.array_equal(outs["1"]["CH"], outs["0"]["CH"])
    assert outs["1"]["iterations"] == outs["0"]["iterations"]


def test_homogenize_batch_stacks_same_shape_cells(torch_gpu, monkeypatch):
    """Stacked against one solver per cell; and the stacked batch with its
    check blocks captured before the worker threads start and replayed
    (VXL_GRAPH=1) against the same batch launched kernel by kernel."""
    from voxel_lmfem.cells import random_inclusion_batch
    from voxel_lmfem.solver import homogenize_batch
    cells = random_inclusion_batch(count=7, seed=13, sizes=(16, 24))
    inputs = [(phase, mats, dom) for (phase, mats, dom, _), _ in cells]
    edges = [m["n"] for _, m in cells]
    print("case batch stacking", edges, flush=True)
    stacked = homogenize_batch(inputs, streams=2, workers=2, autotune=False, stack=10)
    single = homogenize_batch(inputs, streams=2, workers=2, autotune=False, stack=1)
    monkeypatch.setenv("VXL_GRAPH", "1")
    replayed = homogenize_batch(inputs, streams=2, workers=2, autotune=False, stack=10)
    assert sum(r["graph_blocks"] for r in replayed) > 0
    for k, (a, b, c) in enumerate(zip(stacked, single, replayed)):
        assert a["stacked_with"] == edges.count(edges[k]) and b["stacked_with"] == 1
        assert a["max_true_residual"] <= 1e-8
        assert a["grid_shape"] == [edges[k]] * 3
        assert rel(a["CH"], b["CH"]) < 1e-9, k
        assert np.array_equal(np.asarray(c["CH"]), np.asarray(a["CH"])), k


def test_stacked_refusals(torch_gpu):
    torch = torch_gpu
    from voxel_lmfem.gpu import Cell, Solver
    cells = three_cells(n=8)
    fields, media = build(cells)
    shape, h = cells[0][0].shape, (1 / 8,) * 3
    print("case stacked refusals", flush=True)
    stacked = Cell.stack(fields, shape, h)
    try:
        lams, mus = [m[0] for m in media], [m[1] for m in media]
        with pytest.raises(ValueError, match="not a stacked cell"):
            stacked.set_mfma(True)
        with pytest.raises(ValueError, match="multiple of the cell count"):
            stacked.apply_planes(torch.zeros(4, 3, stacked.n, dtype=torch.float64, device="cuda"))
        with pytest.raises(ValueError, match="one lam and one mu per cell"):
            Solver(stacked, lams[0], mus[0])
        with pytest.raises(ValueError, match="multiple of the cell count"):
            Solver(stacked, lams, mus, nrhs=10)
        sv = Solver(stacked, lams, mus)
        try:
            with pytest.raises(ValueError, match="per-case CG"):
                sv.set_method("block")
            with pytest.raises(ValueError, match="FFT preconditioner"):
                sv.set_preconditioner("jacobi")
        finally:
            sv.close()
        u = torch.zeros(18, 3, stacked.n, dtype=torch.float64, device="cuda")
        with pytest.raises(ValueError, match="centre fields take one cell"):
            stacked.center_fields_planes(u, np.tile(np.eye(6), (3, 1)))
        with pytest.raises(ValueError, match="not a stacked cell"):
            stacked.inverse_diagonal_planes()
    finally:
        stacked.close()
