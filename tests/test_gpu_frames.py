"""Per-voxel material frames rebuilt in the FP32 stencil (vxl_cell_set_frames,
homogenize(stencil="frames")): K.u against the oracle's per-voxel tensors
(its own fourth-order rotation) on cells whose every local-frame voxel has a
random frame -- cubic phases take the cubic closed form, orthotropic phases
the orthotropic closed form, general anisotropic phases the Bond rotation,
global-frame phases none; cells with cubic phases only, with cubic and
orthotropic phases, and with all kinds (the three kernel forms) -- given as
rotation matrices and as axis pairs; the same cell's row-based FP32 stencil;
C_eff of mixed solves against the oracle; refusals."""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from oracle import fem, homogenize as oracle, materials as om
from voxel_lmfem import materials as pm

pytestmark = pytest.mark.gpu

# FP32 stencil tolerance relative to max |K u|, as for the row-based stencil;
# the frame rebuilt from an FP32 quaternion adds a few ulp to each coefficient.
TOL = 2e-6

CUBIC = [[168.4, 121.4, 121.4, 0, 0, 0], [121.4, 168.4, 121.4, 0, 0, 0], [121.4, 121.4, 168.4, 0, 0, 0],
         [0, 0, 0, 75.4, 0, 0], [0, 0, 0, 0, 75.4, 0], [0, 0, 0, 0, 0, 75.4]]


def planes(a):
    return np.ascontiguousarray(a.T)


def frame_cell(shape, seed, *, axis_pairs=False, kinds=5):
    """Labels 0 (isotropic, global), 1 (cubic, local), 2 (orthotropic tow,
    local), 3 (general anisotropic, local), 4 (anisotropic, global) with a
    random frame at every voxel; the first `kinds` labels only: 2 runs the
    cubic-only kernels and 3 the cubic-and-orthotropic ones (both on the
    cell's own tile), 5 the kernels with the Bond rotation."""
    rng = np.random.default_rng(seed)
    phase = rng.integers(0, kinds, size=shape).astype(np.int32)
    X = rng.normal(size=(6, 6))
    Y = rng.normal(size=(6, 6))
    mats = {0: dict(kind="isotropic", E=3.5, nu=.35),
            1: dict(kind="anisotropic", frame="local", C=CUBIC),
            2: dict(kind="orthotropic", frame="local", E1=230., E2=15., E3=15., nu12=.2, nu13=.2, nu23=.3,
                    G12=27., G13=27., G23=15. / 2.6),
            3: dict(kind="anisotropic", frame="local", C=(X @ X.T + 6 * np.eye(6)).tolist()),
            4: dict(kind="anisotropic", C=(Y @ Y.T + 6 * np.eye(6)).tolist())}
    n = int(np.prod(shape))
    if axis_pairs:
        a = rng.normal(size=shape + (3,))
        b = rng.normal(size=shape + (3,))
        orientation = dict(axis1=a, axis2=b)
    else:
        orientation = Rotation.random(n, random_state=seed).as_matrix().reshape(shape + (3, 3))
    return phase, mats, orientation


def apply_pair(torch, shape, h, phase, mats, orientation):
    """(oracle K u, frames FP32 K u, row-based FP32 K u) for three fields."""
    from voxel_lmfem.gpu import Cell
    op = fem.MatrixFree(om.full_field(phase, mats, orientation), shape, h)
    rng = np.random.default_rng(3)
    n = int(np.prod(shape))
    U = rng.normal(size=(3, n, 3)) * np.array([1.0, 1e-3, 1e4])[:, None, None]
    want = np.stack([planes(op.apply(u)) for u in U])
    field = pm.build_field(phase, mats, orientation)
    cell = Cell(field, shape, h)
    try:
        u32 = torch.tensor(np.stack([planes(u) for u in U]), device=cell.device, dtype=torch.float32)
        rows = cell.apply_planes(u32).double().cpu().numpy()
        cell.set_frames(*pm.phase_frames(phase, mats, orientation))
        got = cell.apply_planes(u32).double().cpu().numpy()
        again = cell.apply_planes(u32).double().cpu().numpy()
        cell.set_frames(None, None, None, None)
        back = cell.apply_planes(u32).double().cpu().numpy()
    finally:
        cell.close()
    assert np.array_equal(got, again)   # deterministic
    assert np.array_equal(back, rows)   # frames off restores the row-based stencil
    return want, got, rows


@pytest.mark.parametrize("shape,cubic,axis_pairs,kinds",
                         [((2, 2, 2), True, False, 5), ((6, 5, 7), False, False, 5),
                          ((10, 12, 9), True, True, 5), ((17, 9, 40), True, False, 5),
                          ((8, 20, 33), False, True, 5), ((9, 30, 35), True, False, 2),
                          ((5, 16, 20), False, True, 2), ((7, 31, 18), True, False, 3),
                          ((12, 15, 29), False, True, 3)])
def test_frames_apply_matches_oracle(torch_gpu, shape, cubic, axis_pairs, kinds):
    print(f"case frames apply shape={shape} cubic={cubic} axis_pairs={axis_pairs} kinds={kinds}", flush=True)
    phase, mats, orientation = frame_cell(shape, sum(shape), axis_pairs=axis_pairs, kinds=kinds)
    rng = np.random.default_rng(7)
    h = (0.5, 0.5, 0.5) if cubic else tuple(rng.uniform(.3, 1.5, size=3))
    want, got, rows = apply_pair(torch_gpu, shape, h, phase, mats, orientation)
    for r in range(3):
        scale = np.abs(want[r]).max()
        err = np.abs(got[r] - want[r]).max() / scale
        print(f"  rhs {r}: frames rel err {err:.2e}, rows {np.abs(rows[r] - want[r]).max() / scale:.2e}",
              flush=True)
        assert err < TOL


def test_frames_global_only_cell_equals_rows(torch_gpu):
    """No local-frame phase: every element is used as given, so the frames
    stencil reads the same FP32 coefficients as the row-based one."""
    shape = (9, 11, 32)
    print(f"case frames, global phases only {shape}", flush=True)
    rng = np.random.default_rng(2)
    phase = rng.integers(0, 3, size=shape).astype(np.int32)
    mats = {k: dict(kind="isotropic", E=float(3 + 30 * k), nu=.3) for k in range(3)}
    want, got, rows = apply_pair(torch_gpu, shape, (1.0, 1.0, 1.0), phase, mats, None)
    assert np.array_equal(got, rows)


@pytest.mark.parametrize("axis_pairs,kinds", [(False, 5), (True, 5), (False, 2), (True, 3)])
def test_frames_homogenize_matches_oracle(torch_gpu, axis_pairs, kinds):
    from voxel_lmfem import homogenize
    shape = (12, 12, 12)
    print(f"case frames homogenize {shape} axis_pairs={axis_pairs} kinds={kinds}", flush=True)
    phase, mats, orientation = frame_cell(shape, 21, axis_pairs=axis_pairs, kinds=kinds)
    dom = [1.0, 1.0, 1.0]
    ref = oracle.homogenize_matfree(phase, mats, dom, orientation=orientation)["CH"]
    out = homogenize(phase, mats, dom, orientation=orientation, stencil="frames")
    rows = homogenize(phase, mats, dom, orientation=orientation)
    err = np.linalg.norm(np.asarray(out["CH"]) - ref) / np.linalg.norm(ref)
    print(f"  iterations frames {[c['iterations'] for c in out['calls']]}, rows "
          f"{[c['iterations'] for c in rows['calls']]}; rel err vs oracle {err:.2e}", flush=True)
    assert out["stencil"] == "frames" and out["operator"]["variant"] == "modal_fi_lds_stencil_frames"
    assert out["max_true_residual"] <= 1e-8
    assert err < 1e-7


def _set_hg(cell):
    from voxel_lmfem import _native
    _native.check(cell.lib.vxl_cell_set_element(cell.handle, _native.VXL_ELEMENT_C3D8R_HG, 0.1),
                  "vxl_cell_set_element")


def test_frames_refusals(torch_gpu):
    torch = torch_gpu
    from voxel_lmfem import homogenize
    from voxel_lmfem.gpu import Cell
    shape = (4, 4, 4)
    print("case frames refusals", flush=True)
    phase, mats, orientation = frame_cell(shape, 5)
    rows_p, local_p, slot_p, quat_p = pm.phase_frames(phase, mats, orientation)
    field = pm.build_field(phase, mats, orientation)
    cell = Cell(field, shape, (.25, .25, .25))
    try:
        with pytest.raises(ValueError, match="out of range"):
            cell.set_frames(rows_p[:2], local_p[:2], slot_p, quat_p)
        with pytest.raises(ValueError, match="unit quaternions"):
            cell.set_frames(rows_p, local_p, slot_p, 2 * quat_p)
        cell.set_frames(rows_p, local_p, slot_p, quat_p)
        with pytest.raises(ValueError, match="per-voxel frames"):
            cell.set_mfma(True)
        with pytest.raises(ValueError, match="FI element only"):
            _set_hg(cell)
        cell.set_frames(None, None, None, None)
    finally:
        cell.close()
    hg = Cell(field, shape, (.25, .25, .25), element="c3d8r_hg")
    try:
        with pytest.raises(ValueError, match="FI element"):
            hg.set_frames(rows_p, local_p, slot_p, quat_p)
    finally:
        hg.close()
    with pytest.raises(ValueError, match="FI element only"):
        homogenize(phase, mats, [1, 1, 1], orientation=orientation, stencil="frames", element="c3d8r_hg")
    many = {k: dict(kind="isotropic", E=1.0 + k, nu=.3) for k in range(300)}
    with pytest.raises(ValueError, match="at most 256"):
        pm.phase_frames(np.arange(300).reshape(3, 10, 10), many)
