"""The C-ABI library loads and exports every symbol include/vxl.h declares
(no compute calls: these run without a GPU)."""
import ctypes
import re

from conftest import ROOT


def declared_symbols():
    text = (ROOT / "include" / "vxl.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+(vxl_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_full_api():
    names = declared_symbols()
    for required in ("vxl_cell_create", "vxl_apply", "vxl_rhs", "vxl_avg_stress", "vxl_precond_create",
                     "vxl_precond_apply", "vxl_solver_create", "vxl_solver_run", "vxl_last_error"):
        assert required in names


def test_library_exports_every_declared_symbol():
    from voxel_lmfem import _native
    lib = ctypes.CDLL(str(_native.library_path()))
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing
    assert set(declared_symbols()) == set(_native.SIGNATURES)


def test_version_and_error_text_without_gpu():
    from voxel_lmfem import _native
    lib = _native.load()
    assert lib.vxl_version() >= 100
    assert isinstance(lib.vxl_last_error(), bytes)
    # argument validation happens before any HIP call
    h = ctypes.c_void_p()
    rc = lib.vxl_cell_create(ctypes.byref(h), 1, 4, 4, 1., 1., 1., 1, None, 1, None, None, None)
    assert rc == _native.VXL_ERR_ARG and b"grid" in lib.vxl_last_error()
