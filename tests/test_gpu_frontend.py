"""Device material front end (labels -> rows on the GPU), element-centre
fields in the solve result, and the device allocation cache."""
import ctypes

import numpy as np
import pytest

from voxel_lmfem import materials as pm

pytestmark = pytest.mark.gpu


def labelled_cell(shape, seed, labels):
    rng = np.random.default_rng(seed)
    phase = rng.choice(labels, size=shape).astype(np.int64)
    table = {}
    for k in labels:
        X = rng.normal(size=(6, 6))
        table[int(k)] = dict(kind="anisotropic", C=(X @ X.T + 6 * np.eye(6)).tolist())
    return phase, table


@pytest.mark.parametrize("shape,labels,dtype", [
    ((4, 5, 6), [3, 7, 11], "int32"),
    ((17, 9, 33), [-5, 0, 2, 900], "int64"),
    ((8, 8, 8), list(range(300)), "int32"),      # uint16 index
])
def test_device_label_rows_equal_host_bookkeeping(torch_gpu, shape, labels, dtype):
    torch = torch_gpu
    print(f"case label rows shape={shape} labels={len(labels)} {dtype}", flush=True)
    phase, table = labelled_cell(shape, seed=len(labels), labels=labels)
    phase[0, 0, :] = labels[0]  # a long x-run of one label
    host = pm.build_field(phase, table)
    dev = pm.build_field_device(torch.tensor(phase.astype(dtype), device="cuda"), table)
    assert dev.index_kind == host.index_kind
    idx = dev.index_dev.cpu().numpy()
    if host.index.dtype == np.uint16:
        idx = idx.view(np.uint16)
    np.testing.assert_array_equal(idx, host.index)                     # bit-exact bookkeeping
    np.testing.assert_array_equal(dev.rows, host.rows)
    np.testing.assert_array_equal(dev.used_labels(), np.unique(phase))
    np.testing.assert_array_equal(dev.lattice_rows(shape), host.lattice_rows(shape))


def test_device_label_rows_reports_undefined_labels(torch_gpu):
    torch = torch_gpu
    phase, table = labelled_cell((4, 4, 4), seed=1, labels=[1, 2])
    phase[2, 3, 1] = 9
    with pytest.raises(ValueError, match=r"Undefined material labels: \[9\]"):
        pm.build_field_device(torch.tensor(phase, device="cuda"), table)


def test_solve_from_device_field_matches_host_field(torch_gpu):
    torch = torch_gpu
    from voxel_lmfem.cells import periodic_spheres
    from voxel_lmfem.gpu import Cell, Solver
    from voxel_lmfem.reference_medium import select
    (phase, mats, dom, _), _ = periodic_spheres(24, seed=2)
    out = []
    for field in (pm.build_field(phase, mats),
                  pm.build_field_device(torch.tensor(phase.astype(np.int32), device="cuda"), mats)):
        lam, mu, _ = select(field, phase, mats)
        h = dom / np.array(phase.shape[::-1], dtype=float)
        cell = Cell(field, phase.shape, h)
        solver = Solver(cell, lam, mu, mode="mixed", nrhs=6)
        out.append((lam, mu, solver.run()["CH"]))
        solver.close()
        cell.close()
    assert out[0][:2] == out[1][:2]
    np.testing.assert_array_equal(out[0][2], out[1][2])   # same device work, same bits


def test_allocation_cache_reuses_and_releases(torch_gpu):
    torch = torch_gpu
    from voxel_lmfem import _native
    from voxel_lmfem.cells import periodic_spheres
    from voxel_lmfem.gpu import Cell, Solver
    lib = _native.load()
    (phase, mats, dom, _), _ = periodic_spheres(16, seed=3)
    field = pm.build_field(phase, mats)
    h = dom / np.array(phase.shape[::-1], dtype=float)
    cached = ctypes.c_uint64()
    results = []
    for _ in range(2):
        cell = Cell(field, phase.shape, h)
        solver = Solver(cell, 1.0, 1.0, mode="mixed", nrhs=6)
        results.append(solver.run()["CH"])
        solver.close()
        cell.close()
        assert lib.vxl_cached_bytes(ctypes.byref(cached)) == 0
        assert cached.value > 0
    np.testing.assert_array_equal(results[0], results[1])
    assert lib.vxl_release_cached_memory() == 0
    lib.vxl_cached_bytes(ctypes.byref(cached))
    assert cached.value == 0


def test_homogenize_field_statistics(torch_gpu):
    from voxel_lmfem import homogenize
    from voxel_lmfem.cells import periodic_spheres
    (phase, mats, dom, _), _ = periodic_spheres(16, seed=4)
    out = homogenize(phase, mats, dom, fields=True)
    assert out["storage"]["front_end"] == "device"
    stats = out["field_statistics"]
    assert set(stats) == {"0", "5"}
    frac = {str(k): float(np.mean(phase == k)) for k in np.unique(phase)}
    for j in ("0", "5"):
        s = stats[j]
        assert s["stress_l2"] > 0 and s["mean_energy"] > 0
        mix = sum(frac[k] * v for k, v in s["phase_mean_energy"].items())
        assert mix == pytest.approx(s["mean_energy"], rel=1e-12)
    assert out["slices"]["energy_load0"].shape == phase.shape[1:]


def _voxel_tensors_from_device(field):
    rows = field.rows_dev.cpu().numpy()
    idx = field.index_dev.cpu().numpy()
    if idx.dtype == np.int16:   # u16 row numbers stored as int16
        idx = idx.view(np.uint16)
    return pm.unpack21(rows[idx.astype(np.int64)])


@pytest.mark.parametrize("kind", ["woven_axis_pairs", "fixture_matrices", "single_matrix"])
def test_device_local_frames_equal_host_rows(torch_gpu, kind):
    """Frames and Bond-rotated rows of local-frame voxels built on the device
    equal the host's bit for bit."""
    torch = torch_gpu
    from test_oracle import gpu_fixture_cell
    from voxel_lmfem.cells import woven_like
    print(f"case local frames {kind}", flush=True)
    if kind == "woven_axis_pairs":
        (phase, mats, dom, ori), _ = woven_like(32, section=(4, 8), amplitude=3.0)
    elif kind == "fixture_matrices":
        phase, mats, dom, ori = gpu_fixture_cell()
    else:
        phase, mats, dom, _ = gpu_fixture_cell()
        ori = np.array([[0., -1., 0.], [1., 0., 0.], [0., 0., 1.]])
    host = pm.build_field(phase, mats, ori)
    want = pm.unpack21(host.rows[host.index.astype(np.int64)])
    dev = pm.build_field_device(torch.tensor(np.asarray(phase).astype(np.int32), device="cuda"), mats, ori)
    got = _voxel_tensors_from_device(dev)
    # same expressions in the same order, no fused multiply-add: the same bits
    np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(dev.used_labels(), np.unique(phase))


@pytest.mark.parametrize("orientation,message", [
    (dict(axis1=[1., 0., 0.], axis2=[2., 0., 0.]), "collinear"),
    (dict(axis1=[0., 0., 0.], axis2=[0., 1., 0.]), "nonzero"),
    (np.ones((3, 3)), "orthonormal"),
    (np.diag([1., 1., -1.]), "right-handed"),
])
def test_device_local_frames_reject_bad_frames(torch_gpu, orientation, message):
    torch = torch_gpu
    phase = np.zeros((2, 2, 3), dtype=np.int32)
    mats = {0: dict(kind="anisotropic", C=np.diag([10., 20., 30., 4., 5., 6.]).tolist(), frame="local")}
    with pytest.raises(ValueError, match=message):
        pm.build_field_device(torch.tensor(phase, device="cuda"), mats, orientation)


def test_woven_solve_device_front_end_matches_host_field(torch_gpu):
    torch = torch_gpu
    from voxel_lmfem import homogenize
    from voxel_lmfem.cells import woven_like
    from voxel_lmfem.gpu import Cell, Solver
    from voxel_lmfem.reference_medium import select
    (phase, mats, dom, ori), _ = woven_like(32, section=(4, 8), amplitude=3.0)
    out = homogenize(phase, mats, dom, orientation=ori, precision="fp64")
    assert out["storage"]["front_end"] == "device"
    field = pm.build_field(phase, mats, ori)
    lam, mu, _ = select(field, phase, mats)
    h = dom / np.array(phase.shape[::-1], dtype=float)
    cell = Cell(field, phase.shape, h)
    solver = Solver(cell, lam, mu, mode="fp64", nrhs=6)
    CH = solver.run()["CH"]
    solver.close()
    cell.close()
    assert np.linalg.norm(np.asarray(out["CH"]) - CH) <= 1e-11 * np.linalg.norm(CH)


def test_device_rows_shared_between_equal_frames(torch_gpu):
    """Voxels whose Bond-rotated rows are bitwise equal share one table row
    (the reference weave at 64^3: a few hundred rows instead of one per yarn
    voxel).  Every voxel sees the same coefficients bit for bit, the Zeman
    lattice samples follow the new row numbers, and the solve matches the
    one-row-per-voxel table's."""
    torch = torch_gpu
    from voxel_lmfem.cells import reference_weave
    from voxel_lmfem.gpu import Cell, Solver
    from voxel_lmfem.reference_medium import select
    print("case shared device rows, reference weave 64^3", flush=True)
    (phase, mats, dom, ori), _ = reference_weave(64)
    labels = torch.tensor(phase.astype(np.int32), device="cuda")
    full = pm.build_field_device(labels, mats, ori, deduplicate=False)
    shared = pm.build_field_device(labels, mats, ori)
    assert full.row_count == 2 + phase.size and full.index_kind == 4
    assert shared.row_count < 2000 and shared.index_kind in (1, 2)
    np.testing.assert_array_equal(_voxel_tensors_from_device(shared), _voxel_tensors_from_device(full))
    np.testing.assert_array_equal(shared.lattice_rows(phase.shape), full.lattice_rows(phase.shape))
    h = dom / np.array(phase.shape[::-1], dtype=float)
    out = []
    for field in (full, shared):
        lam, mu, _ = select(field, phase, mats)
        cell = Cell(field, phase.shape, h)
        solver = Solver(cell, lam, mu, mode="mixed", nrhs=6)
        got = solver.run()
        out.append((lam, mu, np.asarray(got["CH"]), got["iterations"], max(got["true_residuals"])))
        solver.close()
        cell.close()
    assert out[0][:2] == out[1][:2]
    assert np.linalg.norm(out[1][2] - out[0][2]) <= 1e-12 * np.linalg.norm(out[0][2])
    assert all(abs(a - b) <= 1 for a, b in zip(out[0][3], out[1][3]))
    assert out[1][4] <= 1e-8


@pytest.mark.parametrize("kind", ["woven_axis_pairs", "fixture_matrices", "two_local_phases"])
def test_compact_frames_equal_grid_frames(torch_gpu, kind):
    """The device front end fed only the local voxels' frames, as homogenize
    feeds it after its host check (check_frames(..., keep=True)), gives every
    voxel the same coefficients, bit for bit, and the same Zeman lattice rows
    as the one reading the whole grid's orientation."""
    torch = torch_gpu
    from test_oracle import gpu_fixture_cell
    from voxel_lmfem.cells import woven_like
    print(f"case compact frames {kind}", flush=True)
    if kind == "fixture_matrices":
        phase, mats, dom, ori = gpu_fixture_cell()
    else:
        (phase, mats, dom, ori), _ = woven_like(32, section=(4, 8), amplitude=3.0)
        if kind == "two_local_phases":
            phase = phase.copy()
            phase[:, :, :16][phase[:, :, :16] == 1] = 2
            mats = dict(mats)
            mats[2] = dict(kind="anisotropic", frame="local",
                           C=(np.diag([90000., 8000., 7000., 3000., 4000., 4500.]) + 500.).tolist())
    labels_dev = torch.tensor(np.asarray(phase).astype(np.int32), device="cuda")
    labels, _, local = pm.phase_table(mats)
    flat = np.asarray(phase).reshape(-1)
    sels = [pm.voxels_with_label(flat, labels[k]) for k in np.flatnonzero(local)]
    compact = (sels, pm.check_frames(ori, np.asarray(phase).shape, sels, keep=True))
    for dedup in (False, True):
        full = pm.build_field_device(labels_dev, mats, ori, deduplicate=dedup)
        comp = pm.build_field_device(labels_dev, mats, ori, deduplicate=dedup, compact=compact)
        np.testing.assert_array_equal(_voxel_tensors_from_device(comp).view(np.int64),
                                      _voxel_tensors_from_device(full).view(np.int64))
        np.testing.assert_array_equal(comp.lattice_rows(phase.shape), full.lattice_rows(phase.shape))
        np.testing.assert_array_equal(comp.used_labels(), full.used_labels())
        if not dedup:
            assert comp.row_count == len(labels) + sum(len(s) for s in sels)
