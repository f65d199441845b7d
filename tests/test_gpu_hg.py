"""The C3D8R-HG control element (centroid rule + hourglass penalty,
femlib/c3d8r.py:62-78) on the GPU: operator, Jacobi diagonal, its Green
preconditioner and certified C_eff against the oracle's dense element
matrices, dense symbol inverse and sparse-direct solve."""
import numpy as np
import pytest

from oracle import fem, materials as om, spectral
from voxel_lmfem import materials as pm

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a, dtype=float), np.asarray(b, dtype=float)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def random_cell(shape, seed, labels=3):
    rng = np.random.default_rng(seed)
    phase = rng.integers(0, labels, size=shape).astype(np.int32)
    table = {}
    for k in range(labels):
        X = rng.normal(size=(6, 6))
        table[k] = dict(kind="anisotropic", C=(X @ X.T + 6 * np.eye(6)).tolist())
    h = rng.uniform(.4, 1.4, size=3)
    field = pm.build_field(phase, table, None, deduplicate=False)
    full = om.unpack21(field.rows[field.index.astype(np.int64)])
    return phase, table, field, full, h


@pytest.mark.parametrize("shape,cubic,scaled,coef", [((5, 7, 6), False, False, 0.1), ((8, 6, 10), True, False, 0.37),
                                                     ((6, 9, 7), False, True, 0.1)])
def test_hg_apply_matches_oracle(torch_gpu, shape, cubic, scaled, coef):
    torch = torch_gpu
    from voxel_lmfem.gpu import Cell
    print(f"case hg apply shape={shape} cubic={cubic} scaled={scaled} coef={coef}", flush=True)
    _, _, field, full, h = random_cell(shape, seed=sum(shape))
    if cubic:
        h = np.full(3, 0.8)
    n = int(np.prod(shape))
    scale = None
    if scaled:
        scale = np.random.default_rng(2).uniform(1e-3, 1, size=n).astype(np.float32)
        full = full * scale.astype(float)[:, None, None]
    cell = Cell(field, shape, h, scale=scale, element="c3d8r_hg", hourglass_coefficient=coef)
    op = fem.MatrixFreeHG(full, shape, h, coef)
    U = np.random.default_rng(1).normal(size=(2, n, 3))
    want = np.stack([op.apply(u).T for u in U])
    u64 = torch.tensor(np.stack([np.ascontiguousarray(u.T) for u in U]), device=cell.device)
    scl = np.abs(want).max()
    e64 = np.abs(cell.apply_planes(u64).cpu().numpy() - want).max() / scl
    e32 = np.abs(cell.apply_planes(u64.float()).double().cpu().numpy() - want).max() / scl
    print(f"  fp64 {e64:.2e} fp32 {e32:.2e}", flush=True)
    assert e64 < 1e-13 and e32 < 2e-6
    K = fem.assemble(full, shape, h, element="hg", coef=coef)[0]
    want_d = 1.0 / K.diagonal().reshape(n, 3).T
    got_d = cell.inverse_diagonal_planes().cpu().numpy()
    assert np.abs(got_d - want_d).max() / np.abs(want_d).max() < 1e-13
    cell.close()


@pytest.mark.parametrize("shape", [(8, 8, 8), (6, 5, 4), (16, 12, 32)])
def test_hg_green_matches_dense_inverse(torch_gpu, shape):
    torch = torch_gpu
    from voxel_lmfem.gpu import Cell, Preconditioner
    print(f"case hg green shape={shape}", flush=True)
    _, _, field, _, h = random_cell(shape, seed=3)
    cell = Cell(field, shape, h, element="c3d8r_hg", hourglass_coefficient=0.2)
    lam, mu = 2.5, 1.3
    n = int(np.prod(shape))
    R = np.random.default_rng(6).normal(size=(2, n, 3))
    R -= R.mean(axis=1, keepdims=True)
    want = np.stack([spectral.green_hg_dense(r, shape, h, lam, mu, 0.2).T for r in R])
    r = torch.tensor(np.stack([np.ascontiguousarray(x.T) for x in R]), device="cuda")
    for precision, tol in (("float64", 1e-11), ("float32", 5e-6)):
        pc = Preconditioner(cell, lam, mu, precision=precision, nrhs_max=2)
        z = pc.apply_planes(r.to(pc.dtype)).double().cpu().numpy()
        err = np.abs(z - want).max() / np.abs(want).max()
        print(f"  {precision} err {err:.2e}", flush=True)
        assert err < tol
        pc.close()
    cell.close()


@pytest.mark.parametrize("precision,preconditioner", [("mixed", "fft"), ("fp64", "fft"), ("mixed", "jacobi")])
def test_hg_homogenize_matches_direct_solve(torch_gpu, precision, preconditioner):
    from voxel_lmfem import homogenize
    shape = (6, 5, 7)
    print(f"case hg homogenize {precision} {preconditioner}", flush=True)
    phase, table, _, full, h = random_cell(shape, seed=21)
    domain = h * np.array(shape[::-1])
    _, _, _, CH, _ = fem.assemble(full, shape, h, element="hg", coef=0.1)
    out = homogenize(phase, table, domain, precision=precision, preconditioner=preconditioner,
                     element="c3d8r_hg", tol=1e-10)
    print(f"  rel {rel(out['CH'], CH):.2e} its {[c['iterations'] for c in out['calls']]}", flush=True)
    assert out["operator"]["element"] == "c3d8r_hg" and "C3D8R-HG" in out["scope"]
    assert rel(out["CH"], CH) < 1e-8
    fi = homogenize(phase, table, domain, tol=1e-10)
    assert rel(fi["CH"], CH) > 1e-6  # a different element, a different answer


def test_hg_patch_test_and_guards(torch_gpu):
    from voxel_lmfem import homogenize
    from voxel_lmfem.gpu import Cell, Solver
    C = np.array(om.material_stiffness(dict(kind="isotropic", E=3500., nu=.35)))
    phase = np.zeros((8, 8, 8), dtype=np.int32)
    out = homogenize(phase, {0: dict(kind="isotropic", E=3500., nu=.35)}, [1., 1., 1.], element="c3d8r_hg")
    np.testing.assert_allclose(out["CH"], C, rtol=1e-10, atol=1e-7)
    with pytest.raises(ValueError, match="hourglass"):
        homogenize(phase, {0: dict(kind="isotropic", E=3500., nu=.35)}, [1., 1., 1.], element="c3d8r_hg",
                   hourglass_coefficient=0.0)
    field = pm.build_field(phase, {0: dict(kind="isotropic", E=3500., nu=.35)})
    cell = Cell(field, phase.shape, np.full(3, 1 / 8))
    solver = Solver(cell, 1000.0, 800.0)
    cell.lib.vxl_cell_set_element(cell.handle, 1, 0.1)
    with pytest.raises(ValueError, match="element changed"):
        solver.run()
    solver.close()
    cell.close()
