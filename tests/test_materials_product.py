"""Product host bookkeeping (labels, frames, rotated rows, reference medium)
against the oracle: integer maps and frames bit-exact."""
import numpy as np
import pytest

from oracle import materials as om, homogenize as oh
from voxel_lmfem import materials as pm, inputs, reference_medium
from voxel_lmfem.cells import reference_weave, periodic_spheres
from test_oracle import gpu_fixture_cell


def test_phase_table_bit_exact():
    phase, table, domain, rot = gpu_fixture_cell()
    a = pm.phase_table(table)
    b = om.phase_table(table)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_frames_bit_exact_both_representations():
    (phase, mats, dom, ori), _ = reference_weave(8)
    sel = np.flatnonzero(phase.ravel() == 1)
    assert np.array_equal(pm.frames(ori, phase.shape, sel), om.orientation_values(ori, phase.shape, sel))
    _, _, _, rot = gpu_fixture_cell()
    sel = np.arange(64)
    assert np.array_equal(pm.frames(rot, (4, 4, 4), sel), om.orientation_values(rot, (4, 4, 4), sel))


@pytest.mark.parametrize("dedup", [True, False])
def test_field_rows_equal_reference_pool(dedup):
    (phase, mats, dom, ori), _ = reference_weave(8)
    f = pm.build_field(phase, mats, ori, deduplicate=dedup)
    pool, index = om.hybrid_pool(phase, mats, ori)
    # every voxel sees bit-identical coefficients
    assert np.array_equal(f.rows[f.index.astype(np.int64)], pool[index])
    if not dedup:
        assert np.array_equal(f.index.astype(np.int64), index)
    else:
        assert len(f.rows) < len(pool)


def test_index_dtype_follows_row_count():
    phase, mats, dom, rot = gpu_fixture_cell()
    f = pm.build_field(phase, mats, rot)
    assert f.index.dtype == np.uint8 and f.index_kind == 1
    big = np.arange(300 * 8).reshape(300, 2, 4) % 300
    table = {k: dict(kind="isotropic", E=1. + k, nu=.3) for k in range(300)}
    f = pm.build_field(big, table)
    assert f.index.dtype == np.uint16 and f.index_kind == 2


def test_reference_medium_matches_oracle():
    (phase, mats, dom, ori), _ = reference_weave(8)
    f = pm.build_field(phase, mats, ori)
    lam, mu, meta = reference_medium.select(f, phase, mats)
    assert (lam, mu) == oh.reference_lame(phase, mats, ori)
    (phase, mats, dom, ori), _ = periodic_spheres(16, seed=4)
    f = pm.build_field(phase, mats)
    lam, mu, _ = reference_medium.select(f, phase, mats)
    assert (lam, mu) == oh.reference_lame(phase, mats, None)


@pytest.mark.parametrize("phase,materials,domain,orientation,message", [
    (np.zeros((2, 2, 2)), {0: dict(kind="isotropic", E=1., nu=.2)}, [1, 1, 1], None, "integer"),
    (np.ones((2, 2, 2), dtype=int), {0: dict(kind="isotropic", E=1., nu=.2)}, [1, 1, 1], None, "Undefined"),
    (np.zeros((2, 2, 2), dtype=int), {0: dict(kind="isotropic", E=1., nu=.2)}, [1, 0, 1], None, "positive"),
    (np.zeros((2, 2, 2), dtype=int), {0: dict(kind="anisotropic", C=-np.eye(6))}, [1, 1, 1], None, "positive definite"),
    (np.zeros((2, 2, 2), dtype=int), {0: dict(kind="anisotropic", C=np.eye(6), frame="local")}, [1, 1, 1], None,
     "require orientation"),
    (np.zeros((2, 2, 2), dtype=int), {0: dict(kind="anisotropic", C=np.eye(6), frame="local")}, [1, 1, 1],
     np.diag([1, 1, -1]), "right-handed"),
])
def test_homogenize_rejects_invalid_inputs_before_gpu(phase, materials, domain, orientation, message):
    from voxel_lmfem import homogenize
    with pytest.raises(ValueError, match=message):
        homogenize(phase, materials, domain, orientation=orientation, backend="gpu")


def test_backend_names():
    from voxel_lmfem import homogenize
    with pytest.raises(ValueError, match="backend"):
        homogenize(np.zeros((2, 2, 2), dtype=int), {0: dict(kind="isotropic", E=1., nu=.2)}, [1, 1, 1],
                   backend="tpu")


def test_angle_rows_and_their_derivative():
    from voxel_lmfem import adjoint
    from voxel_lmfem.cells import random_rotations
    rng = np.random.default_rng(8)
    R = random_rotations(5, rng)
    C = pm.orthotropic_voigt(135., 9., 9., .3, .3, .35, 5., 5., 3.5)
    np.testing.assert_allclose(pm.pack21(adjoint._bond(R) @ C @ adjoint._bond(R).transpose(0, 2, 1)),
                               pm.bond_rotate(C, R), rtol=1e-13, atol=1e-12)
    n = 7
    local = np.array([True, False, True, True, False, True, True])
    iso = np.tile(pm.pack21(pm.isotropic_voigt(3.5, .35)), (n, 1))
    theta = rng.uniform(-.5, .5, n)
    rows, drows = adjoint.angle_rows(theta, C, R, iso, local)
    assert np.array_equal(rows[~local], iso[~local]) and not drows[~local].any()
    step = 1e-6
    plus, _ = adjoint.angle_rows(theta + step, C, R, iso, local)
    minus, _ = adjoint.angle_rows(theta - step, C, R, iso, local)
    np.testing.assert_allclose(drows, (plus - minus) / (2 * step), rtol=1e-6, atol=1e-6)
    # theta = 0 reproduces the frames' Bond-rotated rows
    rows0, _ = adjoint.angle_rows(np.zeros(n), C, R, iso, local)
    np.testing.assert_allclose(rows0[local], pm.bond_rotate(C, R), rtol=1e-13, atol=1e-12)


def test_objective_weights():
    from voxel_lmfem import adjoint
    C = np.arange(36.).reshape(6, 6)
    C = C + C.T
    KV = (C[0, 0] + C[1, 1] + C[2, 2] + 2 * (C[0, 1] + C[0, 2] + C[1, 2])) / 9
    assert float((adjoint.objective_weights() * C).sum()) == pytest.approx(C[0, 0] + KV)


def test_rotation_quaternions_match_scipy():
    """Quaternions of the per-voxel frames stencil: scipy's rotation built from
    them (its (x, y, z, w) order) gives back R, including the half-turns where
    w = 0 and Shepperd's method takes another component from the diagonal."""
    from scipy.spatial.transform import Rotation
    R = Rotation.random(500, random_state=4).as_matrix()
    R = np.concatenate([R, np.eye(3)[None], np.diag([1., -1., -1.])[None], np.diag([-1., 1., -1.])[None],
                        np.diag([-1., -1., 1.])[None]])
    q = pm.rotation_quaternions(R)
    assert np.allclose(np.linalg.norm(q, axis=1), 1.0, rtol=0, atol=1e-15)
    back = Rotation.from_quat(q[:, [1, 2, 3, 0]]).as_matrix()
    assert np.abs(back - R).max() < 1e-14


def test_phase_frames_bookkeeping():
    from scipy.spatial.transform import Rotation
    rng = np.random.default_rng(1)
    phase = rng.choice([3, 7, 9], size=(4, 5, 6)).astype(np.int32)
    mats = {3: dict(kind="isotropic", E=2.0, nu=.3),
            7: dict(kind="orthotropic", frame="local", E1=230., E2=15., E3=15., nu12=.2, nu13=.2, nu23=.3,
                    G12=27., G13=27., G23=5.769),
            9: dict(kind="isotropic", E=70.0, nu=.2)}
    R = Rotation.random(phase.size, random_state=2).as_matrix().reshape(phase.shape + (3, 3))
    rows, local, slot, quat = pm.phase_frames(phase, mats, R)
    labels, tensors, flags = pm.phase_table(mats)
    assert np.array_equal(rows, pm.pack21(tensors)) and np.array_equal(local, flags)
    assert slot.dtype == np.uint8 and np.array_equal(labels[slot], phase.reshape(-1))
    loc = phase.reshape(-1) == 7
    assert np.array_equal(quat[~loc], np.tile([1.0, 0, 0, 0], ((~loc).sum(), 1)))
    back = Rotation.from_quat(quat[loc][:, [1, 2, 3, 0]]).as_matrix()
    assert np.abs(back - R.reshape(-1, 3, 3)[loc]).max() < 1e-14
    with pytest.raises(ValueError, match="require orientation"):
        pm.phase_frames(phase, mats)


def _first_error(fn):
    try:
        fn()
    except Exception as exc:  # noqa: BLE001 - the type and text are what is compared
        return type(exc), str(exc)
    return None


def _reference_loop(orientation, shape, selections, chunk):
    """The reference's frame validation loop (src/voxel_lmfem/inputs.py:42-47)
    over the oracle's restatement of its frame routine."""
    for sel in selections:
        for first in range(0, len(sel), chunk):
            om.orientation_values(orientation, shape, sel[first:first + chunk])


def _axes(shape, rng):
    a = rng.normal(size=shape + (3,))
    b = rng.normal(size=shape + (3,))
    return a, b


def _case(name, rng):
    shape = (18, 20, 24)
    phase = rng.integers(0, 3, size=shape)
    sels = [np.flatnonzero(phase.ravel() == k) for k in (1, 2)]
    a, b = _axes(shape, rng)
    flat_a, flat_b = a.reshape(-1, 3), b.reshape(-1, 3)
    s1, s2 = sels
    if name == "valid":
        pass
    elif name == "nan_late_collinear_early":
        flat_b[s1[400]] = np.nan
        flat_b[s1[60]] = 3.0 * flat_a[s1[60]]
    elif name == "collinear_late_zero_early":
        flat_b[s1[300]] = -2.0 * flat_a[s1[300]]
        flat_a[s1[120]] = 0.0
    elif name == "nan_and_collinear_same_chunk":
        flat_b[s1[5]] = flat_a[s1[5]]
        flat_a[s1[7]] = np.inf
    elif name == "second_label_only":
        flat_b[s2[200]] = 0.0
    elif name == "broadcast_collinear":
        return shape, dict(axis1=np.array([1.0, 0.0, 0.0]), axis2=np.array([2.0, 0.0, 0.0])), sels
    elif name == "broadcast_valid":
        return shape, dict(axis1=np.array([1.0, 0.0, 0.0]), axis2=np.array([0.0, 1.0, 0.0])), sels
    elif name == "not_broadcastable":
        return shape, dict(axis1=a[:, :, :5], axis2=b), sels
    elif name == "float32_axes":
        flat_b[s1[350]] = flat_a[s1[350]]
        return shape, dict(axis1=a.astype(np.float32), axis2=b.astype(np.float32)), sels
    elif name in ("rot_valid", "rot_det_early_skew_late", "rot_nan"):
        q = rng.normal(size=shape + (4,))
        q /= np.linalg.norm(q, axis=-1, keepdims=True)
        w, x, y, z = np.moveaxis(q, -1, 0)
        R = np.stack([np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], -1),
                      np.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], -1),
                      np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], -1)], -2)
        flat = R.reshape(-1, 3, 3)
        if name == "rot_det_early_skew_late":
            flat[s1[420]] *= 1.01
            flat[s1[90]] = -flat[s1[90]]
        if name == "rot_nan":
            flat[s2[10], 1, 2] = np.nan
        return shape, R, sels
    return shape, dict(axis1=a, axis2=b), sels


@pytest.mark.parametrize("name", ["valid", "nan_late_collinear_early", "collinear_late_zero_early",
                                  "nan_and_collinear_same_chunk", "second_label_only", "broadcast_collinear",
                                  "broadcast_valid", "not_broadcastable", "float32_axes", "rot_valid",
                                  "rot_det_early_skew_late", "rot_nan"])
@pytest.mark.parametrize("workers", [1, 4])
def test_check_frames_raises_what_the_chunked_reference_loop_raises(name, workers):
    """The pooled frame check (pieces on threads, no frames built) raises the
    exception the reference's piece-by-piece loop raises first, or none."""
    shape, orientation, sels = _case(name, np.random.default_rng(17))
    chunk = 64  # dozens of pieces per label at this size
    expect = _first_error(lambda: _reference_loop(orientation, shape, sels, chunk))
    got = _first_error(lambda: pm.check_frames(orientation, shape, sels, chunk=chunk, workers=workers))
    assert got == expect
    if name in ("valid", "broadcast_valid", "rot_valid"):
        assert got is None
    elif name != "not_broadcastable":
        assert got is not None and got[0] is ValueError


@pytest.mark.parametrize("dtype,lo,hi", [(np.uint8, 0, 255), (np.uint8, 3, 9), (np.int8, -128, 127),
                                         (np.int16, -300, 2000), (np.int32, -7, 40), (np.int64, -(2 ** 40), 2 ** 40),
                                         (np.uint64, 2 ** 63 + 5, 2 ** 63 + 90), (np.uint64, 0, 90),
                                         (np.uint16, 0, 65535)])
def test_used_labels_equal_numpy_unique(dtype, lo, hi):
    rng = np.random.default_rng(3)
    for shape in ((7, 6, 5), (40, 180, 300)):   # one piece; several pieces on the thread pool
        phase = rng.integers(lo, hi, size=shape, dtype=dtype, endpoint=True)
        got = inputs.used_labels(phase)
        ref = np.unique(phase)
        assert got.dtype == ref.dtype and np.array_equal(got, ref)


def test_voxels_with_label_equal_flatnonzero():
    rng = np.random.default_rng(4)
    flat = rng.integers(0, 5, size=3_000_001).astype(np.uint8)
    for label in (0, 3, 7):
        assert np.array_equal(pm.voxels_with_label(flat, label), np.flatnonzero(flat == label))
    assert pm.voxels_with_label(flat[:0], 1).size == 0


def _fake_device_field(torch, L=3, n=4000, distinct=40, seed=5, compact=False):
    """A DeviceField-shaped object on CPU tensors: L phase rows, local voxels
    (every third voxel) with one of `distinct` rows each, at row L + voxel
    (or, compact, at row L + their rank among the local voxels)."""
    from types import SimpleNamespace
    rng = np.random.default_rng(seed)
    index = rng.integers(0, L, size=n).astype(np.int32)
    local = np.arange(n) % 3 == 0
    vox = np.flatnonzero(local)
    rows_of_local = L + (np.arange(len(vox)) if compact else vox)
    table = np.zeros((L + (len(vox) if compact else n), 21))
    table[:L] = rng.normal(size=(L, 21))
    choices = rng.normal(size=(distinct, 21))
    choices[1::2] = choices[0::2]
    choices[1::2, [4, 9]] *= -1   # rows of frames +theta / -theta: two words differ in the sign bit
    table[rows_of_local] = choices[rng.integers(0, distinct, size=len(vox))]
    index[vox] = rows_of_local
    samples = index[np.arange(0, n, 7)].astype(np.int32)
    return SimpleNamespace(index_dev=torch.as_tensor(index), rows_dev=torch.as_tensor(table), samples=samples,
                           kind=4, _lattice=None)


def _effective(field):
    idx = field.index_dev.numpy()
    if idx.dtype == np.int16:
        idx = idx.view(np.uint16)
    return field.rows_dev.numpy()[idx.astype(np.int64)]


@pytest.mark.parametrize("compact", [False, True])
def test_equal_rows_share_one_table_row(compact):
    """Device row sharing (run here on CPU tensors): every voxel keeps its
    coefficients bit for bit, the table shrinks to the distinct rows, the
    index narrows, and the lattice samples point at rows with the same
    coefficients, for a table with one row per voxel and for a compact one
    (one row per local voxel, numbered by rank)."""
    import torch
    field = _fake_device_field(torch, compact=compact)
    before = _effective(field).copy()
    sampled = field.rows_dev.numpy()[field.samples.astype(np.int64)].copy()
    pm._share_equal_rows(field, 3, torch)
    assert field.rows_dev.shape[0] <= 3 + 40 and field.kind == 1
    assert np.array_equal(_effective(field).view(np.int64), before.view(np.int64))
    assert np.array_equal(field.rows_dev.numpy()[field.samples.astype(np.int64)], sampled)


def test_row_key_collisions_keep_one_row_per_voxel(monkeypatch):
    """With every key equal (a forced collision) the word-for-word check
    fails and the one-row-per-voxel table stays as built."""
    import torch
    field = _fake_device_field(torch)
    rows_before = field.rows_dev.clone()
    monkeypatch.setattr(pm, "_ROW_KEY_MIX", (0, 0))
    pm._share_equal_rows(field, 3, torch)
    assert torch.equal(field.rows_dev, rows_before) and field.kind == 4


def test_no_local_voxels_keep_the_phase_rows_only():
    import torch
    field = _fake_device_field(torch)
    field.index_dev = torch.as_tensor(np.arange(4000, dtype=np.int32) % 3)
    field.samples = np.array([0, 1, 2], dtype=np.int32)
    pm._share_equal_rows(field, 3, torch)
    assert field.rows_dev.shape[0] == 3 and field.kind == 1
    assert np.array_equal(field.index_dev.numpy(), np.arange(4000) % 3)
