"""Block CG over the load cases (vxl_solver_set_method(VXL_METHOD_BLOCK_CG)):
certified C_eff against the oracle and the per-case CG, in mixed and FP64
precision, on a polycrystal and a high-contrast inclusion cell; a zero load
case; two identical load cases (a rank-deficient block); and its refusals."""
import numpy as np
import pytest

from oracle import homogenize as oracle

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a, dtype=float), np.asarray(b, dtype=float)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_block_cg_matches_oracle_and_cg(torch_gpu, precision):
    from voxel_lmfem import homogenize
    from voxel_lmfem.cells import voronoi_polycrystal
    (phase, mats, dom, _), _ = voronoi_polycrystal(n=24, grains=12, seed=5)
    print(f"case block CG polycrystal 24^3 {precision}", flush=True)
    ref = oracle.homogenize_matfree(phase, mats, dom)["CH"]
    blk = homogenize(phase, mats, dom, precision=precision, solver="block")
    cg = homogenize(phase, mats, dom, precision=precision)
    print(f"  block its {[c['iterations'] for c in blk['calls']]}, cg its {[c['iterations'] for c in cg['calls']]}, "
          f"rel err vs oracle {rel(blk['CH'], ref):.2e}", flush=True)
    assert blk["solver"] == "block" and cg["solver"] == "cg"
    assert blk["max_true_residual"] <= 1e-8
    assert rel(blk["CH"], ref) < 1e-7
    assert rel(blk["CH"], cg["CH"]) < 1e-7


def test_block_cg_high_contrast_inclusions(torch_gpu):
    from voxel_lmfem import homogenize
    from voxel_lmfem.cells import periodic_spheres
    (phase, mats, dom, _), _ = periodic_spheres(32, radius=4.0, fraction=0.3, seed=2, E=(1.0, 1000.0))
    print("case block CG spheres 32^3 contrast 1000", flush=True)
    blk = homogenize(phase, mats, dom, solver="block")
    cg = homogenize(phase, mats, dom)
    print(f"  block its {[c['iterations'] for c in blk['calls']]}, cg its {[c['iterations'] for c in cg['calls']]}",
          flush=True)
    assert blk["max_true_residual"] <= 1e-8
    assert rel(blk["CH"], cg["CH"]) < 1e-7


def test_block_cg_zero_and_repeated_loads(torch_gpu):
    """A zero macro strain gives the zero field without entering the block;
    two equal macro strains make the block rank deficient (their directions
    coincide), and both still certify with equal stress columns."""
    from voxel_lmfem import materials as pm
    from voxel_lmfem.cells import voronoi_polycrystal
    from voxel_lmfem.gpu import Cell, Solver
    from voxel_lmfem.reference_medium import select
    (phase, mats, dom, _), _ = voronoi_polycrystal(n=16, grains=6, seed=3)
    print("case block CG zero and repeated loads 16^3", flush=True)
    field = pm.build_field(phase, mats)
    lam, mu, _ = select(field, phase, mats)
    cell = Cell(field, phase.shape, dom / 16)
    macro = np.array([[1, 0, 0, 0, 0, 0], [0, 0, 0, 0, 0, 0], [1, 0, 0, 0, 0, 0], [0, .5, 0, 0, 0, 1.]], float)
    try:
        for mode in ("mixed", "fp64"):
            blk = Solver(cell, lam, mu, mode=mode, nrhs=4, method="block")
            cg = Solver(cell, lam, mu, mode=mode, nrhs=4)
            try:
                b = blk.run(macro)
                c = cg.run(macro)
            finally:
                blk.close()
                cg.close()
            print(f"  {mode}: block its {b['iterations']} cg its {c['iterations']}", flush=True)
            assert max(b["true_residuals"]) <= 1e-8
            assert not np.asarray(b["CH"])[:, 1].any() and b["iterations"][1] == 0
            assert np.allclose(b["CH"][:, 0], b["CH"][:, 2], rtol=1e-8, atol=1e-8 * np.abs(b["CH"]).max())
            assert rel(b["CH"], c["CH"]) < 1e-7
    finally:
        cell.close()


def test_block_cg_refusals(torch_gpu):
    from voxel_lmfem import homogenize, materials as pm
    from voxel_lmfem.gpu import Cell, Solver
    phase = np.zeros((4, 4, 4), dtype=np.int32)
    phase[:2] = 1
    mats = {0: dict(kind="isotropic", E=3.0, nu=0.3), 1: dict(kind="isotropic", E=70.0, nu=0.2)}
    with pytest.raises(ValueError, match="FFT preconditioner"):
        homogenize(phase, mats, [1, 1, 1], solver="block", preconditioner="jacobi")
    with pytest.raises(ValueError, match="solver must be"):
        homogenize(phase, mats, [1, 1, 1], solver="gmres")
    cell = Cell(pm.build_field(phase, mats), phase.shape, (.25, .25, .25))
    try:
        big = Solver(cell, 30.0, 10.0, nrhs=9)
        try:
            with pytest.raises(ValueError, match="at most 8"):
                big.set_method("block")
        finally:
            big.close()
    finally:
        cell.close()
