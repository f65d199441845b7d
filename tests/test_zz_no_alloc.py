"""The allocation-free zero-mean projection (vxl_project_zero_mean_noalloc)
against vxl_project_zero_mean, through the C ABI: bit-identical fields on two
calls in a row, the second on a smaller cell with fewer fields, and the
16-field limit of its static scratch refused with VXL_ERR_ARG.  Both entry
points are also checked against the NumPy mean."""
import ctypes

import numpy as np
import pytest

from voxel_lmfem import _native, materials as pm

pytestmark = pytest.mark.gpu


def cell_for(shape, seed):
    from voxel_lmfem.gpu import Cell
    rng = np.random.default_rng(seed)
    phase = rng.integers(0, 2, size=shape).astype(np.int32)
    table = {0: dict(kind="isotropic", E=3000.0, nu=0.35), 1: dict(kind="isotropic", E=70000.0, nu=0.22)}
    field = pm.build_field(phase, table)
    return Cell(field, shape, (1.0, 1.0, 1.0))


def project_both(torch, cell, nrhs, seed):
    lib = _native.load()
    rng = np.random.default_rng(seed)
    host = rng.normal(loc=3.0, size=(nrhs, 3, cell.n))
    with torch.cuda.stream(cell.stream):
        a = torch.tensor(host, device=cell.device)
        b = a.clone()
    cell.stream.synchronize()
    _native.check(lib.vxl_project_zero_mean(cell.handle, nrhs, _native.dptr(a)), "vxl_project_zero_mean")
    _native.check(lib.vxl_project_zero_mean_noalloc(cell.handle, nrhs, _native.dptr(b)),
                  "vxl_project_zero_mean_noalloc")
    cell.stream.synchronize()  # the no-alloc call returns before its work ends
    a, b = a.cpu().numpy(), b.cpu().numpy()
    want = host - host.mean(axis=2, keepdims=True)
    return a, b, want


def test_noalloc_projection_matches_the_allocating_entry_point(torch_gpu):
    torch = torch_gpu
    for shape, nrhs, seed in (((24, 20, 40), 6, 1), ((10, 12, 14), 2, 2)):
        print(f"case noalloc zero mean shape={shape} nrhs={nrhs}", flush=True)
        cell = cell_for(shape, seed)
        try:
            a, b, want = project_both(torch, cell, nrhs, seed)
        finally:
            cell.close()
        assert np.array_equal(a, b)
        assert np.abs(b - want).max() < 1e-12
        assert np.abs(b.mean(axis=2)).max() < 1e-13


def test_noalloc_projection_refuses_more_than_16_fields(torch_gpu):
    torch = torch_gpu
    lib = _native.load()
    cell = cell_for((4, 4, 4), 3)
    try:
        u = torch.zeros((17, 3, cell.n), dtype=torch.float64, device=cell.device)
        rc = lib.vxl_project_zero_mean_noalloc(cell.handle, 17, _native.dptr(u))
        assert rc == _native.VXL_ERR_ARG and b"at most 16" in lib.vxl_last_error()
        rc = lib.vxl_project_zero_mean_noalloc(cell.handle, 0, _native.dptr(u))
        assert rc == _native.VXL_ERR_ARG
        rc = lib.vxl_project_zero_mean_noalloc(cell.handle, 16, _native.dptr(u[:16]))
        assert rc == _native.VXL_OK
        cell.stream.synchronize()
        assert not u.any()
    finally:
        cell.close()
