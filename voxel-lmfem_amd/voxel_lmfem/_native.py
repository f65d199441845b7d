"""ctypes binding of libvxl.so, the gfx950 HIP library of this package.

The library is built in-tree (``make -C voxel-lmfem_amd``) next to this file.
There is no fallback: if the library is missing or the HIP runtime cannot be
initialised, every GPU entry point raises ``RuntimeError``.
"""
import ctypes
import os
from pathlib import Path

HERE = Path(__file__).resolve().parent
LIBRARY = HERE / "libvxl.so"

VXL_OK = 0
VXL_ERR_ARG = -1
VXL_ERR_HIP = -2
VXL_ERR_CONVERGENCE = -3
VXL_INDEX_U8, VXL_INDEX_U16, VXL_INDEX_I32 = 1, 2, 4
VXL_F32, VXL_F64, VXL_F32_BF16 = 0, 1, 2
VXL_PRECOND_FFT, VXL_PRECOND_JACOBI, VXL_PRECOND_FFT_BF16 = 0, 1, 2
VXL_ELEMENT_FI, VXL_ELEMENT_C3D8R_HG = 0, 1
VXL_SOLVE_FP64, VXL_SOLVE_MIXED = 0, 1
VXL_METHOD_CG, VXL_METHOD_BLOCK_CG = 0, 1

_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_PP = ctypes.POINTER(ctypes.c_void_p)
_DP = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int)

# name -> (restype, argtypes); mirrors include/vxl.h
SIGNATURES = {
    "vxl_last_error": (ctypes.c_char_p, []),
    "vxl_version": (_I, []),
    "vxl_cached_bytes": (_I, [ctypes.POINTER(ctypes.c_uint64)]),
    "vxl_release_cached_memory": (_I, []),
    "vxl_label_rows": (_I, [_I, _I, _I, _I, _P, ctypes.c_longlong, _I, _IP, _I, _P, ctypes.c_longlong, _IP, _IP,
                            _IP, _P]),
    "vxl_local_rows": (_I, [_I, _I, _I, _I, _P, ctypes.c_longlong, _I, _IP, _I, _DP, _I, _P, _P, ctypes.c_longlong,
                            _P, ctypes.c_longlong, _IP, _P]),
    "vxl_cell_create_device_rows": (_I, [_PP, _I, _I, _I, _D, _D, _D, _I, _P, _I, _P, _P, _P]),
    "vxl_cell_create": (_I, [_PP, _I, _I, _I, _D, _D, _D, _I, _P, _I, _DP, _P, _P]),
    "vxl_cell_destroy": (_I, [_P]),
    "vxl_apply": (_I, [_P, _I, _I, _P, _P]),
    "vxl_set_stencil_variant": (_I, [_P, _I, _I]),
    "vxl_get_stencil_variant": (_I, [_P, _I]),
    "vxl_tune_stencil": (_I, [_P, _I, _I, _P, _P, _I, _IP, _DP]),
    "vxl_set_stencil_config": (_I, [_P, _I, _I, _I, _I]),
    "vxl_cell_set_element": (_I, [_P, _I, _D]),
    "vxl_get_stencil_config": (_I, [_P, _I, _IP, _IP, _IP]),
    "vxl_cell_set_mfma": (_I, [_P, _I]),
    "vxl_cell_set_frames": (_I, [_P, _I, _DP, _IP, _P, _P]),
    "vxl_cell_set_groups": (_I, [_P, _I]),
    "vxl_solver_set_media": (_I, [_P, _DP, _DP]),
    "vxl_autotune_stencil": (_I, [_P, _I, _I, _P, _P, _I, _IP, _IP, _IP, _DP]),
    "vxl_rhs": (_I, [_P, _I, _DP, _P]),
    "vxl_avg_stress": (_I, [_P, _I, _P, _DP, _DP]),
    "vxl_center_fields": (_I, [_P, _I, _P, _DP, _P, _P, _P]),
    "vxl_energy_gradient": (_I, [_P, _I, _P, _DP, _DP, _P, _P]),
    "vxl_project_zero_mean": (_I, [_P, _I, _P]),
    "vxl_project_zero_mean_noalloc": (_I, [_P, _I, _P]),
    "vxl_precond_create": (_I, [_PP, _P, _I, _D, _D, _I]),
    "vxl_precond_apply": (_I, [_P, _I, _P, _P]),
    "vxl_precond_destroy": (_I, [_P]),
    "vxl_solver_create": (_I, [_PP, _P, _I, _I, _D, _D]),
    "vxl_solver_run": (_I, [_P, _DP, _D, _I, _DP, _IP, _IP, _DP, _P]),
    "vxl_solver_run_batch": (_I, [_PP, _I, _DP, _D, _I, _DP, _IP, _IP, _DP, _I]),
    "vxl_cell_inverse_diagonal": (_I, [_P, _P]),
    "vxl_solver_set_preconditioner": (_I, [_P, _I]),
    "vxl_solver_set_method": (_I, [_P, _I]),
    "vxl_solver_graph_blocks": (_I, [_P]),
    "vxl_solver_timings": (_I, [_P, _DP]),
    "vxl_solver_trace": (_I, [_P, _I, _IP, _DP, _I]),
    "vxl_solver_destroy": (_I, [_P]),
}

_lib = None


class ConvergenceError(RuntimeError):
    """The FP64 true-residual certificate was not reached."""


def library_path():
    return Path(os.environ.get("VXL_LIBRARY", str(LIBRARY)))


def load():
    """Load libvxl.so once; raise RuntimeError if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    path = library_path()
    if not path.is_file():
        raise RuntimeError(f"libvxl.so not built at {path}; run `make -C voxel-lmfem_amd` "
                           "(or __graft_entry__.build())")
    lib = ctypes.CDLL(str(path))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc, what=""):
    """Map a libvxl status to the reference's exception types."""
    if rc == VXL_OK:
        return
    text = load().vxl_last_error().decode(errors="replace")
    message = f"{what}: {text}" if what else text
    if rc == VXL_ERR_ARG:
        raise ValueError(message)
    if rc == VXL_ERR_CONVERGENCE:
        raise ConvergenceError(message)
    raise RuntimeError(message)


def dptr(tensor):
    """Device pointer of a contiguous torch tensor as a ctypes void pointer."""
    if not tensor.is_contiguous():
        raise ValueError("device buffers must be contiguous")
    return ctypes.c_void_p(tensor.data_ptr())


def darray(values):
    import numpy as np
    arr = np.ascontiguousarray(values, dtype=np.float64)
    return arr, arr.ctypes.data_as(_DP)
