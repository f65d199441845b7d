"""The matrix-core FP32 stencil (label-grouped dense element matrices on
v_mfma_f32_16x16x32_f16, vxl_cell_set_mfma) against the oracle: K.u on
cubic and non-cubic voxels, ragged and tiny grids, many rows per wave, a
per-voxel scale field; and C_eff of mixed solves that run it inside the CG
(the certificate is FP64 either way)."""
import numpy as np
import pytest

from oracle import fem, homogenize as oracle, materials as om
from voxel_lmfem import materials as pm

pytestmark = pytest.mark.gpu

# FP16 hi/lo products keep about 2^-21 of each element column's largest
# entry; the tolerance is relative to max |K u| as for the modal FP32 stencil.
TOL = 2e-6


def planes(a):
    return np.ascontiguousarray(a.T)


def label_cell(shape, seed, labels, *, cubic):
    rng = np.random.default_rng(seed)
    phase = rng.integers(0, labels, size=shape).astype(np.int32)
    table = {}
    for k in range(labels):
        X = rng.normal(size=(6, 6))
        table[k] = dict(kind="anisotropic", C=(X @ X.T + 6 * np.eye(6)).tolist())
    field = pm.build_field(phase, table)
    full = om.unpack21(field.rows[field.index.astype(np.int64)])
    h = (0.7, 0.7, 0.7) if cubic else tuple(rng.uniform(.3, 1.5, size=3))
    return field, full, h


@pytest.mark.parametrize("shape,labels,cubic", [((2, 2, 2), 2, True), ((4, 4, 4), 3, False),
                                                 ((5, 7, 6), 5, True), ((24, 20, 40), 3, True),
                                                 ((9, 33, 17), 70, False), ((16, 15, 70), 300, True)])
def test_mfma_apply_matches_oracle(torch_gpu, shape, labels, cubic):
    torch = torch_gpu
    from voxel_lmfem.gpu import Cell
    print(f"case mfma apply shape={shape} labels={labels} cubic={cubic}", flush=True)
    field, full, h = label_cell(shape, sum(shape) + labels, labels, cubic=cubic)
    cell = Cell(field, shape, h)
    try:
        op = fem.MatrixFree(full, shape, h)
        rng = np.random.default_rng(3)
        n = int(np.prod(shape))
        U = rng.normal(size=(3, n, 3)) * np.array([1.0, 1e-3, 1e4])[:, None, None]
        want = np.stack([planes(op.apply(u)) for u in U])
        u32 = torch.tensor(np.stack([planes(u) for u in U]), device=cell.device, dtype=torch.float32)
        modal = cell.apply_planes(u32).double().cpu().numpy()
        cell.set_mfma(True)
        got = cell.apply_planes(u32).double().cpu().numpy()
        again = cell.apply_planes(u32).double().cpu().numpy()
        cell.set_mfma(False)
        back = cell.apply_planes(u32).double().cpu().numpy()
    finally:
        cell.close()
    assert np.array_equal(got, again)      # deterministic (no atomics)
    assert np.array_equal(back, modal)     # switching back restores the modal stencil
    for r in range(3):                     # each field against its own scale
        scale = np.abs(want[r]).max()
        err = np.abs(got[r] - want[r]).max() / scale
        print(f"  rhs {r}: mfma rel err {err:.2e}, modal {np.abs(modal[r] - want[r]).max() / scale:.2e}",
              flush=True)
        assert err < TOL


def test_mfma_apply_with_scale_field(torch_gpu):
    torch = torch_gpu
    from voxel_lmfem.gpu import Cell
    shape = (12, 10, 36)
    print(f"case mfma apply with per-voxel scale {shape}", flush=True)
    field, full, h = label_cell(shape, 11, 2, cubic=True)
    rng = np.random.default_rng(5)
    scale = rng.uniform(1e-3, 1.0, size=shape)
    cell = Cell(field, shape, h, scale=scale)
    try:
        op = fem.MatrixFree(full * scale.reshape(-1)[:, None, None], shape, h)
        n = int(np.prod(shape))
        U = rng.normal(size=(2, n, 3))
        want = np.stack([planes(op.apply(u)) for u in U])
        cell.set_mfma(True)
        got = cell.apply_planes(torch.tensor(np.stack([planes(u) for u in U]), device=cell.device,
                                             dtype=torch.float32)).double().cpu().numpy()
    finally:
        cell.close()
    assert np.abs(got - want).max() / np.abs(want).max() < TOL


def test_mfma_refuses_per_voxel_rows_and_hg(torch_gpu):
    from voxel_lmfem.gpu import Cell
    rng = np.random.default_rng(2)
    shape = (17, 16, 16)  # 4352 voxels, each its own row
    phase = np.zeros(shape, dtype=np.int32)
    table = {0: dict(kind="anisotropic", C=(np.diag([9., 8, 7, 3, 2, 1])).tolist(), frame="local")}
    ori = dict(axis1=rng.normal(size=shape + (3,)), axis2=rng.normal(size=shape + (3,)))
    field = pm.build_field(phase, table, ori, deduplicate=False)
    assert len(field.rows) > 4096
    cell = Cell(field, shape, (1., 1., 1.))
    try:
        with pytest.raises(ValueError, match="4096 rows"):
            cell.set_mfma(True)
    finally:
        cell.close()
    field, _, _ = label_cell((4, 4, 4), 1, 2, cubic=True)
    cell = Cell(field, (4, 4, 4), (1., 1., 1.), element="c3d8r_hg")
    try:
        with pytest.raises(ValueError, match="FI element"):
            cell.set_mfma(True)
    finally:
        cell.close()


@pytest.mark.parametrize("n", [24, 32])
def test_mfma_mixed_solve_matches_oracle(torch_gpu, n):
    """Mixed solves with the matrix-core stencil in the CG: C_eff within 1e-8
    of the oracle and of the modal-stencil solve (both FP64-certified)."""
    from voxel_lmfem import homogenize
    from voxel_lmfem.cells import voronoi_polycrystal
    print(f"case mfma mixed solve polycrystal {n}^3", flush=True)
    (phase, mats, dom, _), _ = voronoi_polycrystal(n=n, grains=200 if n == 32 else 40, seed=4)
    ref = oracle.homogenize_matfree(phase, mats, dom)["CH"]
    modal = homogenize(phase, mats, dom, precision="mixed")
    mf = homogenize(phase, mats, dom, precision="mixed", stencil="mfma")
    err = np.linalg.norm(np.asarray(mf["CH"]) - ref) / np.linalg.norm(ref)
    print(f"  rel err vs oracle {err:.2e}; its mfma {[c['iterations'] for c in mf['calls']]} "
          f"modal {[c['iterations'] for c in modal['calls']]}", flush=True)
    assert mf["stencil"] == "mfma" and modal["stencil"] == "modal"
    assert err < 1e-7 and mf["max_true_residual"] <= 1e-8
    assert np.linalg.norm(np.asarray(mf["CH"]) - np.asarray(modal["CH"])) / np.linalg.norm(ref) < 1e-7
