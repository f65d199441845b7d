"""Preconditioner options of the solver: the Jacobi control (1 / diag K) and
the FFT Green preconditioner with BF16-stored spectra.  The assembled
diagonal and the solves are checked against the oracle's sparse assembly and
sparse-direct C_eff; the BF16 variant against the FP32-storage solve."""
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
    return field, full, h


@pytest.mark.parametrize("shape,scaled", [((2, 2, 2), False), ((3, 4, 5), False), ((9, 33, 17), False),
                                          ((6, 5, 7), True)])
def test_inverse_diagonal_matches_oracle(torch_gpu, shape, scaled):
    from voxel_lmfem.gpu import Cell
    print(f"case inverse diagonal shape={shape} scaled={scaled}", flush=True)
    field, full, h = random_cell(shape, seed=sum(shape))
    n = int(np.prod(shape))
    scale = None
    if scaled:
        scale = np.random.default_rng(4).uniform(1e-3, 1.0, size=n).astype(np.float32)
        full = full * scale.astype(float)[:, None, None]
    cell = Cell(field, shape, h, scale=scale)
    want = 1.0 / fem.assemble_matrix(full, shape, h).diagonal().reshape(n, 3).T
    got = cell.inverse_diagonal_planes().cpu().numpy()
    err = np.abs(got - want).max() / np.abs(want).max()
    print(f"  err {err:.2e}", flush=True)
    assert err < 1e-13
    cell.close()


def _lam_mu(full):
    """Isotropic medium from the mean tensor (any positive pair is a valid
    reference medium; the Jacobi solve does not use it)."""
    C = full.mean(axis=0)
    mu = float(np.mean([C[3, 3], C[4, 4], C[5, 5]]))
    lam = float(np.mean([C[0, 1], C[0, 2], C[1, 2]]))
    return max(lam, .1), max(mu, .1)


@pytest.mark.parametrize("mode", ["mixed", "fp64"])
def test_jacobi_solve_matches_oracle(torch_gpu, mode):
    from voxel_lmfem.gpu import Cell, Solver
    shape = (6, 5, 7)
    print(f"case jacobi solve mode={mode} shape={shape}", flush=True)
    field, full, h = random_cell(shape, seed=21)
    _, _, _, CH, _ = fem.assemble(full, shape, h)
    cell = Cell(field, shape, h)
    lam, mu = _lam_mu(full)
    jac = Solver(cell, lam, mu, mode=mode, preconditioner="jacobi")
    out_j = jac.run(np.eye(6), tol=1e-10)
    fft = Solver(cell, lam, mu, mode=mode)
    out_f = fft.run(np.eye(6), tol=1e-10)
    print(f"  iterations jacobi {out_j['iterations']} fft {out_f['iterations']}", flush=True)
    assert max(out_j["true_residuals"]) <= 1e-10
    assert rel(out_j["CH"], CH) < 1e-8
    assert rel(out_f["CH"], CH) < 1e-8
    assert out_j["trace"] != out_f["trace"]  # a different preconditioner ran
    # switching back restores the FFT solve exactly
    jac.set_preconditioner("fft")
    again = jac.run(np.eye(6), tol=1e-10)
    assert np.array_equal(again["CH"], out_f["CH"])
    assert again["iterations"] == out_f["iterations"]
    for s in (jac, fft):
        s.close()
    cell.close()


def test_jacobi_homogenize_polycrystal(torch_gpu):
    from voxel_lmfem import homogenize
    from voxel_lmfem.cells import voronoi_polycrystal
    (phase, mats, dom, _), _ = voronoi_polycrystal(n=24, grains=12, seed=2)
    print("case jacobi homogenize polycrystal 24^3", flush=True)
    a = homogenize(phase, mats, dom, preconditioner="jacobi")
    b = homogenize(phase, mats, dom)
    its_a = [c["iterations"] for c in a["calls"]]
    its_b = [c["iterations"] for c in b["calls"]]
    print(f"  iterations jacobi {its_a} fft {its_b}", flush=True)
    assert a["operator"]["preconditioner"] == "jacobi"
    assert "Jacobi" in a["scope"]
    assert a["max_true_residual"] <= 1e-8
    assert rel(a["CH"], b["CH"]) < 1e-6


@pytest.mark.parametrize("shape", [(8, 8, 8), (16, 32, 64)])
def test_bf16_spectra_apply(torch_gpu, shape):
    """BF16 storage between passes: error at BF16's rounding level (8-bit
    significand), far above FP32's, against the FP64 oracle."""
    torch = torch_gpu
    from voxel_lmfem.gpu import Cell, Preconditioner
    print(f"case bf16 precond shape={shape}", flush=True)
    field, full, h = random_cell(shape, seed=5)
    cell = Cell(field, shape, h)
    lam, mu = 2.5, 1.3
    n = int(np.prod(shape))
    R = np.random.default_rng(6).normal(size=(2, n, 3))
    R -= R.mean(axis=1, keepdims=True)
    ref = np.stack([spectral.green_apply(r, shape, h, lam, mu).T for r in R])
    r = torch.tensor(np.stack([np.ascontiguousarray(x.T) for x in R]), device="cuda", dtype=torch.float32)
    errs = {}
    for precision in ("float32", "bf16"):
        pc = Preconditioner(cell, lam, mu, precision=precision, nrhs_max=2)
        z = pc.apply_planes(r).double().cpu().numpy()
        errs[precision] = rel(z, ref)
        pc.close()
    print(f"  rel err fp32 {errs['float32']:.2e} bf16 {errs['bf16']:.2e}", flush=True)
    assert errs["float32"] < 1e-6
    assert 1e-5 < errs["bf16"] < 2e-2
    cell.close()


def test_bf16_spectra_rejects_axes_without_four_step(torch_gpu):
    from voxel_lmfem.gpu import Cell, Preconditioner
    field, _, h = random_cell((8, 8, 12), seed=1)
    cell = Cell(field, (8, 8, 12), h)
    with pytest.raises(ValueError, match="four-step factorisation"):
        Preconditioner(cell, 2.0, 1.0, precision="bf16")
    cell.close()


def test_bf16_homogenize_polycrystal(torch_gpu):
    """Certified solve with BF16-stored spectra: same C_eff as the FP32-storage
    preconditioner to well inside the 1e-8 certificate's reach."""
    from voxel_lmfem import homogenize
    from voxel_lmfem.cells import voronoi_polycrystal
    (phase, mats, dom, _), _ = voronoi_polycrystal(n=32, grains=16, seed=4)
    print("case bf16 homogenize polycrystal 32^3", flush=True)
    a = homogenize(phase, mats, dom, preconditioner="fft_bf16")
    b = homogenize(phase, mats, dom)
    its_a = [c["iterations"] for c in a["calls"]]
    its_b = [c["iterations"] for c in b["calls"]]
    print(f"  iterations bf16 {its_a} fp32 {its_b}; rel {rel(a['CH'], b['CH']):.2e}", flush=True)
    assert a["max_true_residual"] <= 1e-8
    assert rel(a["CH"], b["CH"]) < 1e-6
    with pytest.raises(ValueError, match="mixed"):
        homogenize(phase, mats, dom, preconditioner="fft_bf16", precision="fp64")
