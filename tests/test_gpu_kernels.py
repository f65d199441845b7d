"""GPU parity of the libvxl kernels against the CPU oracle (FP64 bitwise-close,
FP32 to its rounding level): stiffness action, RHS, mean stress, FFT Green
preconditioner."""
import numpy as np
import pytest

from oracle import fem, materials as om, spectral
from voxel_lmfem import materials as pm

pytestmark = pytest.mark.gpu


def make_cell(shape, seed, *, labels=3, local=True, big_table=False):
    rng = np.random.default_rng(seed)
    phase = rng.integers(0, labels, size=shape).astype(np.int32)
    table = {}
    for k in range(labels):
        X = rng.normal(size=(6, 6))
        table[k] = dict(kind="anisotropic", C=(X @ X.T + 6 * np.eye(6)).tolist())
    orientation = None
    if local:
        table[labels - 1]["frame"] = "local"
        a, b = rng.normal(size=shape + (3,)), rng.normal(size=shape + (3,))
        orientation = dict(axis1=a, axis2=b)
    if big_table:  # more than 65536 rows -> int32 index
        pass
    h = rng.uniform(.3, 1.5, size=3)
    field = pm.build_field(phase, table, orientation, deduplicate=False)
    full = om.unpack21(field.rows[field.index.astype(np.int64)])
    return field, full, h


def planes(a):
    return np.ascontiguousarray(a.T)  # (N,3) -> (3,N)


SHAPES = [(2, 2, 2), (4, 4, 4), (3, 4, 5), (5, 7, 6), (24, 20, 40), (9, 33, 17)]


@pytest.mark.parametrize("shape", SHAPES)
def test_apply_matches_oracle(torch_gpu, shape):
    torch = torch_gpu
    from voxel_lmfem.gpu import Cell
    print(f"case apply shape={shape}", flush=True)
    field, full, h = make_cell(shape, seed=sum(shape))
    cell = Cell(field, shape, h)
    op = fem.MatrixFree(full, shape, h)
    rng = np.random.default_rng(1)
    n = int(np.prod(shape))
    U = rng.normal(size=(2, n, 3))
    ref = np.stack([op.apply(u) for u in U])
    u64 = torch.tensor(np.stack([planes(u) for u in U]), device=cell.device)
    want = np.stack([planes(r) for r in ref])
    scale = np.abs(ref).max()
    for variant in (1, 2, 3):  # every stencil schedule
        cell.set_variant("float64", variant)
        cell.set_variant("float32", variant)
        out64 = cell.apply_planes(u64).cpu().numpy()
        err64 = np.abs(out64 - want).max() / scale
        out32 = cell.apply_planes(u64.float()).double().cpu().numpy()
        err32 = np.abs(out32 - want).max() / scale
        print(f"  schedule {variant}: fp64 err {err64:.2e} fp32 err {err32:.2e}", flush=True)
        assert err64 < 1e-13
        assert err32 < 2e-6
    cell.close()


def test_tune_stencil_keeps_fastest_schedule(torch_gpu):
    torch = torch_gpu
    from voxel_lmfem.gpu import Cell
    shape = (16, 24, 32)
    print(f"case tune shape={shape}", flush=True)
    field, full, h = make_cell(shape, seed=5)
    cell = Cell(field, shape, h)
    for precision in ("float32", "float64"):
        best, ms = cell.tune(precision, nrhs=6 if precision == "float32" else 1, reps=3)
        print(f"  {precision}: best {best} ms {ms}", flush=True)
        assert best in (1, 2, 3) and all(m > 0 for m in ms)
        assert ms[best - 1] == min(ms)
        code = 0 if precision == "float32" else 1
        assert cell.lib.vxl_get_stencil_variant(cell.handle, code) == best
    with pytest.raises(ValueError):
        cell.set_variant("float32", 4)
    cell.close()


def test_apply_symmetry_rigid_and_index_kinds(torch_gpu):
    torch = torch_gpu
    from voxel_lmfem.gpu import Cell
    shape = (6, 5, 7)
    print(f"case symmetry shape={shape}", flush=True)
    field, full, h = make_cell(shape, seed=11)
    n = int(np.prod(shape))
    rng = np.random.default_rng(2)
    u = torch.tensor(rng.normal(size=(1, 3, n)), device="cuda")
    v = torch.tensor(rng.normal(size=(1, 3, n)), device="cuda")
    outs = []
    for dtype in (np.uint8, np.uint16, np.int32):
        f = pm.MaterialField(field.rows, field.index.astype(dtype), field.labels, field.tensors, field.local)
        cell = Cell(f, shape, h)
        Ku, Kv = cell.apply_planes(u), cell.apply_planes(v)
        outs.append(Ku.cpu().numpy())
        a, b = float((v * Ku).sum()), float((u * Kv).sum())
        assert abs(a - b) < 1e-12 * max(abs(a), 1)
        rigid = torch.zeros(1, 3, n, dtype=torch.float64, device="cuda")
        rigid[0, 0], rigid[0, 1], rigid[0, 2] = 1.0, -2.0, 0.5
        assert float(cell.apply_planes(rigid).abs().max()) < 1e-12 * float(Ku.abs().max())
        cell.close()
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])


def test_apply_scale_field(torch_gpu):
    torch = torch_gpu
    from voxel_lmfem.gpu import Cell
    shape = (5, 6, 7)
    print(f"case scale shape={shape}", flush=True)
    field, full, h = make_cell(shape, seed=5, local=False)
    rng = np.random.default_rng(3)
    s = rng.uniform(.1, 2., size=shape).astype(np.float32)
    cell = Cell(field, shape, h, scale=s)
    op = fem.MatrixFree(full * s.reshape(-1).astype(np.float64)[:, None, None], shape, h)
    u = rng.normal(size=(int(np.prod(shape)), 3))
    ref = planes(op.apply(u))
    out = cell.apply_planes(torch.tensor(planes(u)[None], device="cuda"))[0].cpu().numpy()
    assert np.abs(out - ref).max() < 1e-13 * np.abs(ref).max()
    cell.close()


@pytest.mark.parametrize("shape", [(4, 4, 4), (3, 5, 7), (24, 20, 40)])
def test_rhs_and_stress_match_oracle(torch_gpu, shape):
    torch = torch_gpu
    from voxel_lmfem.gpu import Cell
    print(f"case rhs/stress shape={shape}", flush=True)
    field, full, h = make_cell(shape, seed=7)
    cell = Cell(field, shape, h)
    op = fem.MatrixFree(full, shape, h)
    macro = np.eye(6) + 0.1 * np.random.default_rng(0).normal(size=(6, 6))
    f = cell.rhs_planes(macro).cpu().numpy()
    for j in range(6):
        ref = planes(op.rhs(macro[j]))
        assert np.abs(f[j] - ref).max() < 1e-13 * np.abs(ref).max()
    n = int(np.prod(shape))
    U = np.random.default_rng(9).normal(size=(6, n, 3))
    u = torch.tensor(np.stack([planes(x) for x in U]), device="cuda")
    s = cell.avg_stress_planes(u, macro)
    for j in range(6):
        ref = op.avg_stress(U[j], macro[j])
        assert np.abs(s[j] - ref).max() < 1e-12 * np.abs(ref).max()
    cell.close()


@pytest.mark.parametrize("shape", [(4, 4, 4), (5, 3, 7)])
def test_center_fields_match_oracle(torch_gpu, shape):
    torch = torch_gpu
    from voxel_lmfem.gpu import Cell
    print(f"case centre fields shape={shape}", flush=True)
    field, full, h = make_cell(shape, seed=11)
    cell = Cell(field, shape, h)
    op = fem.MatrixFree(full, shape, h)
    rng = np.random.default_rng(12)
    n = int(np.prod(shape))
    macro = rng.normal(size=(2, 6))
    U = rng.normal(size=(2, n, 3))
    u = torch.tensor(np.stack([planes(x) for x in U]), device="cuda")
    eps, sig, w = (t.cpu().numpy() for t in cell.center_fields_planes(u, macro))
    for j in range(2):
        e_ref, s_ref, w_ref = op.center_fields(U[j], macro[j])
        assert np.abs(eps[j] - e_ref.T).max() < 1e-13 * np.abs(e_ref).max()
        assert np.abs(sig[j] - s_ref.T).max() < 1e-13 * np.abs(s_ref).max()
        assert np.abs(w[j] - w_ref).max() < 1e-13 * np.abs(w_ref).max()
    # reference protocol: (1, N, 6) per load, node-major (1, 3N) input
    u_ref = torch.tensor(U[0].reshape(1, -1), device="cuda")
    strain, stress = cell.recover_center_fields(u_ref, macro[0])
    assert tuple(strain.shape) == (1, n, 6) and tuple(stress.shape) == (1, n, 6)
    np.testing.assert_allclose(stress[0].cpu().numpy(), op.center_fields(U[0], macro[0])[1], rtol=1e-12,
                               atol=1e-13 * np.abs(sig).max())
    cell.close()


PC_SHAPES = [(4, 4, 4), (8, 8, 8), (6, 5, 4), (3, 7, 5), (16, 12, 10), (17, 2, 9), (32, 48, 20),
             (5, 3, 16), (3, 6, 64), (2, 5, 128), (3, 3, 256), (2, 3, 512),
             # four-step passes of 2^a times 3 or 5 on every axis
             (24, 40, 48), (48, 24, 80), (96, 5, 40), (160, 3, 24), (8, 80, 96), (192, 2, 8), (2, 192, 8),
             (3, 4, 160), (2, 3, 192)]
FOUR_STEP = {8, 16, 32, 64, 128, 256, 512, 24, 40, 48, 80, 96, 160, 192}


@pytest.mark.parametrize("shape", PC_SHAPES)
def test_precond_matches_oracle(torch_gpu, shape, monkeypatch):
    """nx with a four-step factorisation (2^a >= 8, or 2^a times 3 or 5 from 24
    to 192) runs the four-step x passes; they are also checked against the
    Stockham x passes (VXL_FFT_XFOURSTEP=0) on the same input.  y and z axes
    with a factorisation run the four-step y and z passes."""
    torch = torch_gpu
    from voxel_lmfem.gpu import Cell, Preconditioner
    print(f"case precond shape={shape}", flush=True)
    field, full, h = make_cell(shape, seed=3, local=False)
    cell = Cell(field, shape, h)
    lam, mu = 2.5, 1.3
    n = int(np.prod(shape))
    R = np.random.default_rng(6).normal(size=(3, n, 3))
    R -= R.mean(axis=1, keepdims=True)
    ref = np.stack([planes(spectral.green_apply(r, shape, h, lam, mu)) for r in R])
    r = torch.tensor(np.stack([planes(x) for x in R]), device="cuda")
    for xfour in (("1", "0") if shape[2] in FOUR_STEP else ("1",)):
        monkeypatch.setenv("VXL_FFT_XFOURSTEP", xfour)
        for precision, tol in (("float64", 1e-11), ("float32", 5e-6)):
            pc = Preconditioner(cell, lam, mu, precision=precision, nrhs_max=3)
            z = pc.apply_planes(r.to(pc.dtype)).double().cpu().numpy()
            err = np.abs(z - ref).max() / np.abs(ref).max()
            print(f"  x four-step {xfour} {precision} err {err:.2e}", flush=True)
            assert err < tol
            pc.close()
    cell.close()


TILES = {"float32": (14, 8, 6), "float64": (6, 8, 4)}


@pytest.mark.parametrize("shape,cubic", [((9, 20, 40), True), ((7, 17, 33), False)])
def test_every_stencil_config_matches_oracle(torch_gpu, shape, cubic):
    """All 27 (tile, schedule, chunk depth) configurations per precision; cubic
    voxels take the folded-factor element form."""
    torch = torch_gpu
    from voxel_lmfem.gpu import Cell
    print(f"case stencil configs shape={shape} cubic={cubic}", flush=True)
    field, full, h = make_cell(shape, seed=sum(shape) + 1)
    if cubic:
        h = np.full(3, 0.7)
    cell = Cell(field, shape, h)
    op = fem.MatrixFree(full, shape, h)
    n = int(np.prod(shape))
    U = np.random.default_rng(2).normal(size=(2, n, 3))
    want = np.stack([planes(op.apply(u)) for u in U])
    scale = np.abs(want).max()
    u64 = torch.tensor(np.stack([planes(u) for u in U]), device=cell.device)
    worst = {"float32": 0.0, "float64": 0.0}
    for precision, tol in (("float64", 1e-13), ("float32", 2e-6)):
        uu = u64 if precision == "float64" else u64.float()
        for ty in TILES[precision]:
            for v in (1, 2, 3):
                for z in (8, 4, 2):
                    cell.set_stencil_config(precision, ty, v, z)
                    assert cell.stencil_config(precision) == (ty, v, z)
                    got = cell.apply_planes(uu).double().cpu().numpy()
                    err = np.abs(got - want).max() / scale
                    worst[precision] = max(worst[precision], err)
                    assert err < tol, (precision, ty, v, z, err)
    print(f"  worst fp64 {worst['float64']:.2e} fp32 {worst['float32']:.2e}", flush=True)
    with pytest.raises(ValueError):
        cell.set_stencil_config("float32", 16, 2, 8)   # not a supported tile
    with pytest.raises(ValueError):
        cell.set_stencil_config("float64", 8, 2, 1)    # chunk depth below 2
    cell.close()


def test_autotune_and_persistent_cache(torch_gpu, tmp_path):
    from voxel_lmfem import autotune
    from voxel_lmfem.gpu import Cell
    shape = (16, 24, 32)
    print(f"case autotune shape={shape}", flush=True)
    field, full, h = make_cell(shape, seed=5)
    cell = Cell(field, shape, h)
    (ty, v, z), table = cell.autotune("float32", nrhs=6, reps=2)
    assert (table > 0).all()
    i, j, k = np.unravel_index(np.argmin(table), table.shape)
    assert (TILES["float32"][i], j + 1, (8, 4, 2)[k]) == (ty, v, z)
    assert cell.stencil_config("float32") == (ty, v, z)
    path = tmp_path / "tuning.json"
    first = autotune.tune(cell, "float32", 6, path=path, reps=2)
    assert first["tuned"] and path.is_file()
    cell.set_stencil_config("float32", 14, 1, 8)
    second = autotune.tune(cell, "float32", 6, path=path)
    assert not second["tuned"]
    assert cell.stencil_config("float32") == (first["tile_y"], first["schedule"], first["zmin"])
    print(f"  tuned {first}", flush=True)
    cell.close()
