"""Device objects over libvxl: cell (operator), FFT Green preconditioner, solver.

PyTorch supplies device memory and the current stream; all arithmetic runs in
the HIP kernels of libvxl.so.  The operator object exposes the reference's
operator protocol (src/voxel_lmfem/_core/materials/hybrid_adapter.py:39-125:
``B``, ``mesh``, ``apply``, ``compute_rhs``, ``compute_avg_stress``,
``project_zero_mean``, ``apply_precond``) on (1, 3N) FP64 node-major tensors,
and component-plane batched entry points used by the solver.
"""
import contextlib
import ctypes
from types import SimpleNamespace

import numpy as np

from . import _native
from ._native import check, darray, dptr


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("No HIP device is available to PyTorch; the GPU backend needs an MI355X")
    return torch


class Cell:
    """Grid + material field resident on the current device."""

    ELEMENTS = {"fi": _native.VXL_ELEMENT_FI, "c3d8r_hg": _native.VXL_ELEMENT_C3D8R_HG}

    def __init__(self, field, shape, h, *, scale=None, device=None, stream=None, element="fi",
                 hourglass_coefficient=0.1):
        """stream: torch.cuda.Stream the cell (and everything built on it) runs
        on; default the current stream.  element: "fi" (fully integrated HEX8)
        or "c3d8r_hg" (the reference's reduced-integration control with an
        hourglass penalty of `hourglass_coefficient`)."""
        if element not in self.ELEMENTS:
            raise ValueError(f"element must be one of {sorted(self.ELEMENTS)}")
        torch = _torch()
        self.stream = torch.cuda.current_stream() if stream is None else stream
        self.lib = _native.load()
        self.torch = torch
        self.device = torch.device("cuda", torch.cuda.current_device()) if device is None else device
        self.shape = tuple(int(s) for s in shape)
        self.h = tuple(float(v) for v in h)
        self.n = int(np.prod(self.shape))
        self.field = field
        if getattr(field, "index_dev", None) is not None:  # built on the device
            self.index = field.index_dev
        else:
            index = np.ascontiguousarray(field.index)
            if index.dtype == np.uint16:
                index = index.view(np.int16)
            self.index = torch.from_numpy(index).to(self.device)
        self.scale = None
        if scale is not None:
            self.scale = torch.as_tensor(np.ascontiguousarray(scale, dtype=np.float32).reshape(-1),
                                         device=self.device)
            if self.scale.numel() != self.n:
                raise ValueError("scale must have one value per voxel")
        if stream is not None:  # uploads above ran on the current stream
            torch.cuda.current_stream().synchronize()
        handle = ctypes.c_void_p()
        nz, ny, nx = self.shape
        scale_p = dptr(self.scale) if self.scale is not None else None
        rows_dev = getattr(field, "rows_dev", None)
        if rows_dev is not None:   # row table built on the device
            check(self.lib.vxl_cell_create_device_rows(
                ctypes.byref(handle), nz, ny, nx, self.h[0], self.h[1], self.h[2], field.index_kind,
                dptr(self.index), int(rows_dev.shape[0]), dptr(rows_dev), scale_p,
                ctypes.c_void_p(self.stream.cuda_stream)), "vxl_cell_create_device_rows")
        else:
            rows, rows_p = darray(field.rows.reshape(-1, 21))
            check(self.lib.vxl_cell_create(ctypes.byref(handle), nz, ny, nx, self.h[0], self.h[1], self.h[2],
                                           field.index_kind, dptr(self.index), len(rows), rows_p, scale_p,
                                           ctypes.c_void_p(self.stream.cuda_stream)), "vxl_cell_create")
        self.handle = handle
        self.element = element
        if element != "fi":
            check(self.lib.vxl_cell_set_element(handle, self.ELEMENTS[element], float(hourglass_coefficient)),
                  "vxl_cell_set_element")
        self.mesh = SimpleNamespace(Nz=nz, Ny=ny, Nx=nx, dx=self.h[0], dy=self.h[1], dz=self.h[2],
                                    n_elem=self.n, n_nodes=self.n, n_dofs=3 * self.n,
                                    V_elem=float(np.prod(self.h)))
        self.B = 1
        self.groups = 1

    @classmethod
    def stack(cls, fields, shape, h, *, stream=None):
        """One cell that runs several cells of the same grid and voxel size in
        every launch (vxl_cell_set_groups): their row tables concatenated,
        their index blocks one after the other (shifted to the concatenated
        rows).  Label-only fields (no per-voxel local-frame rows), FI element.
        Solvers on it take 6 right-hand sides per cell and one reference
        medium per cell (Solver(cell, lams, mus))."""
        from .materials import index_dtype
        torch = _torch()
        fields = list(fields)
        if not 1 <= len(fields) <= 64:
            raise ValueError("a stacked cell takes 1 to 64 cells")
        if any(getattr(f, "rows_dev", None) is not None for f in fields):
            raise ValueError("stacked cells take label-only fields (no per-voxel local-frame rows)")
        n = int(np.prod(shape))
        rows = [np.asarray(f.rows, dtype=np.float64).reshape(-1, 21) for f in fields]
        total = sum(len(r) for r in rows)
        dtype = index_dtype(total)
        kind = {np.uint8: _native.VXL_INDEX_U8, np.uint16: _native.VXL_INDEX_U16, np.int32: _native.VXL_INDEX_I32}[dtype]
        device = torch.device("cuda", torch.cuda.current_device())
        parts, offset = [], 0
        for f, r in zip(fields, rows):
            idx = getattr(f, "index_dev", None)
            if idx is None:
                idx = torch.from_numpy(np.ascontiguousarray(np.asarray(f.index).astype(np.int64))).to(device)
            idx = idx.reshape(-1).to(torch.int64)
            if idx.numel() != n:
                raise ValueError("every stacked field needs one index per voxel of the shared grid")
            parts.append(idx + offset)
            offset += len(r)
        index = torch.cat(parts).to(torch.int32)
        index = {np.uint8: lambda t: t.to(torch.uint8), np.uint16: lambda t: t.to(torch.int16),
                 np.int32: lambda t: t}[dtype](index)
        torch.cuda.current_stream().synchronize()
        stacked = SimpleNamespace(index_dev=index, rows=np.concatenate(rows), index_kind=kind, rows_dev=None,
                                  fields=fields)
        cell = cls(stacked, shape, h, stream=stream)
        check(cell.lib.vxl_cell_set_groups(cell.handle, len(fields)), "vxl_cell_set_groups")
        cell.groups = len(fields)
        return cell

    def close(self):
        pc = getattr(self, "_preconditioner", None)
        if pc is not None:
            pc.close()
            self._preconditioner = None
        if getattr(self, "handle", None):
            self.lib.vxl_cell_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @contextlib.contextmanager
    def ordered(self):
        """libvxl launches on the cell's stream; tensors handed to it are made
        and read on the caller's current stream.  Inside this block the cell's
        stream waits for the work queued so far on the current stream (inputs
        written, freed blocks reused), and on leaving the current stream waits
        for the cell's (outputs read there).  Device-side waits only; nothing
        to do when the cell runs on the current stream."""
        cur = self.torch.cuda.current_stream(self.device)
        same = cur.cuda_stream == self.stream.cuda_stream
        if not same:
            self.stream.wait_stream(cur)
        try:
            yield
        finally:
            if not same:
                cur.wait_stream(self.stream)

    # -- component-plane batched entry points ---------------------------------
    def apply_planes(self, u, out=None):
        """u: (R, 3, N) float32/float64 device tensor -> K u (same layout)."""
        torch = self.torch
        if u.dim() != 3 or u.shape[1:] != (3, self.n) or u.device != self.device:
            raise ValueError("expected a (R, 3, N) tensor on the cell's device")
        precision = {torch.float32: _native.VXL_F32, torch.float64: _native.VXL_F64}.get(u.dtype)
        if precision is None:
            raise ValueError("apply supports float32 and float64")
        u = u.contiguous()
        out = torch.empty_like(u) if out is None else out
        with self.ordered():
            check(self.lib.vxl_apply(self.handle, precision, u.shape[0], dptr(u), dptr(out)), "vxl_apply")
        return out

    def _timing_buffers(self, nrhs, dtype):
        """Random input and output fields for the stencil timers, filled on the
        cell's own stream: libvxl launches the timed kernels there, so they
        are ordered after the fill whatever stream is current."""
        torch = self.torch
        with torch.cuda.stream(self.stream):
            u = torch.randn(nrhs, 3, self.n, dtype=dtype, device=self.device)
            out = torch.empty_like(u)
        return u, out

    def tune(self, precision="float32", nrhs=6, reps=5):
        """Time every stencil schedule on this cell and keep the fastest.

        Returns (best schedule, [ms per launch of schedules 1, 2, 3])."""
        torch = self.torch
        dtype = torch.float32 if precision in ("float32", "fp32", "mixed") else torch.float64
        code = _native.VXL_F32 if dtype == torch.float32 else _native.VXL_F64
        u, out = self._timing_buffers(nrhs, dtype)
        best = ctypes.c_int()
        ms = (ctypes.c_double * 3)()
        check(self.lib.vxl_tune_stencil(self.handle, code, nrhs, dptr(u), dptr(out), int(reps),
                                        ctypes.byref(best), ms), "vxl_tune_stencil")
        return best.value, [ms[0], ms[1], ms[2]]

    def set_variant(self, precision, variant):
        code = _native.VXL_F64 if precision in ("float64", "fp64") else _native.VXL_F32
        check(self.lib.vxl_set_stencil_variant(self.handle, code, int(variant)), "vxl_set_stencil_variant")

    @staticmethod
    def _code(precision):
        return _native.VXL_F64 if precision in ("float64", "fp64") else _native.VXL_F32

    def stencil_config(self, precision="float32"):
        """(tile_y, schedule, fewest z planes per chunk) of the stencil."""
        t, v, z = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(self.lib.vxl_get_stencil_config(self.handle, self._code(precision), ctypes.byref(t), ctypes.byref(v),
                                              ctypes.byref(z)), "vxl_get_stencil_config")
        return t.value, v.value, z.value

    def set_mfma(self, enable=True):
        """Run the cell's FP32 applies (plain and CG form) on the matrix cores:
        label-grouped dense element matrices on FP16 hi/lo MFMA products
        (FI element, at most 4096 coefficient rows).  False restores the modal
        VALU stencil.  Switch before building solvers that should use it."""
        check(self.lib.vxl_cell_set_mfma(self.handle, 1 if enable else 0), "vxl_cell_set_mfma")

    def set_frames(self, phase_rows, phase_local, phase_index, quaternions):
        """Rebuild element stiffness in the FP32 stencil from each voxel's
        phase and frame instead of reading its coefficient row.

        phase_rows: (P, 21) FP64 phase rows in their own frames; phase_local:
        (P,) flags of the local-frame phases; phase_index: (N,) uint8 phase
        slot per voxel (z, y, x order; host or device); quaternions: (N, 4)
        unit quaternions (w, x, y, z) of the voxel rotations.  FP64 applies,
        RHS, certificate and stress keep the cell's rows.  None as
        phase_index turns the frames off."""
        torch = self.torch
        if phase_index is None:
            check(self.lib.vxl_cell_set_frames(self.handle, 0, None, None, None, None), "vxl_cell_set_frames")
            return
        rows = np.ascontiguousarray(phase_rows, dtype=np.float64)
        flags = np.ascontiguousarray(np.asarray(phase_local).astype(np.int32))
        if rows.ndim != 2 or rows.shape[1] != 21 or flags.shape != (rows.shape[0],):
            raise ValueError("phase_rows must be (P, 21) with one frame flag per phase")
        index = torch.as_tensor(phase_index).reshape(-1)
        quat = torch.as_tensor(quaternions).reshape(-1, 4)
        if index.numel() != self.n or quat.shape[0] != self.n:
            raise ValueError("phase_index and quaternions need one entry per voxel")
        index = index.to(device=self.device, dtype=torch.uint8).contiguous()
        quat = quat.to(device=self.device, dtype=torch.float32).contiguous()
        self.torch.cuda.current_stream(self.device).synchronize()
        check(self.lib.vxl_cell_set_frames(self.handle, rows.shape[0], rows.ctypes.data_as(_native._DP),
                                           flags.ctypes.data_as(_native._IP), dptr(index), dptr(quat)),
              "vxl_cell_set_frames")

    def set_stencil_config(self, precision, tile_y, variant, zmin):
        check(self.lib.vxl_set_stencil_config(self.handle, self._code(precision), int(tile_y), int(variant),
                                              int(zmin)), "vxl_set_stencil_config")

    def autotune(self, precision="float32", nrhs=6, reps=3):
        """Time all 27 (tile, schedule, chunk depth) stencil configurations on
        this cell and keep the fastest.  Returns (tile_y, schedule, zmin) and
        the 3 x 3 x 3 table of milliseconds per launch."""
        torch = self.torch
        dtype = torch.float64 if precision in ("float64", "fp64") else torch.float32
        u, out = self._timing_buffers(nrhs, dtype)
        t, v, z = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        ms = (ctypes.c_double * 27)()
        check(self.lib.vxl_autotune_stencil(self.handle, self._code(precision), int(nrhs), dptr(u), dptr(out),
                                            int(reps), ctypes.byref(t), ctypes.byref(v), ctypes.byref(z), ms),
              "vxl_autotune_stencil")
        return (t.value, v.value, z.value), np.array(ms[:]).reshape(3, 3, 3)

    def rhs_planes(self, macro):
        macro, p = darray(np.atleast_2d(macro))
        out = self.torch.empty((len(macro), 3, self.n), dtype=self.torch.float64, device=self.device)
        with self.ordered():
            check(self.lib.vxl_rhs(self.handle, len(macro), p, dptr(out)), "vxl_rhs")
        return out

    def avg_stress_planes(self, u, macro):
        macro, p = darray(np.atleast_2d(macro))
        u = u.contiguous()
        if u.dtype != self.torch.float64 or u.shape != (len(macro), 3, self.n):
            raise ValueError("expected FP64 (R, 3, N) displacements, one per macro strain")
        out = np.empty((len(macro), 6))
        with self.ordered():
            check(self.lib.vxl_avg_stress(self.handle, len(macro), dptr(u), p,
                                          out.ctypes.data_as(_native._DP)), "vxl_avg_stress")
        return out

    def center_fields_planes(self, u, macro, *, strain=True, stress=True, energy=True):
        """Element-centre strain/stress planes (R, 6, N) and energy (R, N), FP64."""
        torch = self.torch
        macro, p = darray(np.atleast_2d(macro))
        u = u.contiguous()
        if u.dtype != torch.float64 or u.shape != (len(macro), 3, self.n):
            raise ValueError("expected FP64 (R, 3, N) displacements, one per macro strain")
        R = len(macro)
        eps = torch.empty(R, 6, self.n, dtype=torch.float64, device=self.device) if strain else None
        sig = torch.empty(R, 6, self.n, dtype=torch.float64, device=self.device) if stress else None
        w = torch.empty(R, self.n, dtype=torch.float64, device=self.device) if energy else None
        null = ctypes.c_void_p()
        with self.ordered():
            check(self.lib.vxl_center_fields(self.handle, R, dptr(u), p, dptr(eps) if strain else null,
                                             dptr(sig) if stress else null, dptr(w) if energy else null),
                  "vxl_center_fields")
        return eps, sig, w

    def inverse_diagonal_planes(self):
        """(3, N) FP64: 1 / diag K, assembled from the 8 elements around each node."""
        out = self.torch.empty((3, self.n), dtype=self.torch.float64, device=self.device)
        with self.ordered():
            check(self.lib.vxl_cell_inverse_diagonal(self.handle, dptr(out)), "vxl_cell_inverse_diagonal")
        return out

    def project_planes(self, u):
        """Subtract each (R, 3) component plane's mean in place (FP64)."""
        if u.dtype != self.torch.float64 or u.dim() != 3 or u.shape[1:] != (3, self.n) or not u.is_contiguous():
            raise ValueError("expected contiguous FP64 (R, 3, N) planes")
        with self.ordered():
            check(self.lib.vxl_project_zero_mean(self.handle, u.shape[0], dptr(u)), "vxl_project_zero_mean")
        return u

    # -- reference operator protocol on (1, 3N) node-major FP64 -----------------
    def geometry(self):
        """(Nz, Ny, Nx, dx, dy, dz), as HybridFIOperator.geometry (hybrid_adapter.py:70-72)."""
        m = self.mesh
        return (m.Nz, m.Ny, m.Nx, m.dx, m.dy, m.dz)

    def _to_planes(self, v):
        if (v.shape != (1, 3 * self.n) or v.dtype != self.torch.float64 or v.device != self.device
                or not v.is_contiguous()):
            raise ValueError("Expected contiguous FP64 displacement (1,3N) on the material device")
        return v.view(self.n, 3).t().contiguous().view(1, 3, self.n)

    def _from_planes(self, v):
        return v.view(3, self.n).t().contiguous().view(1, 3 * self.n)

    def apply(self, u):
        return self._from_planes(self.apply_planes(self._to_planes(u)))

    def compute_rhs(self, macro):
        if np.shape(macro) != (6,):
            raise ValueError("Six engineering macro strain components required")
        return self._from_planes(self.rhs_planes(np.asarray(macro, dtype=float)[None]))

    def compute_avg_stress(self, u, macro):
        if np.shape(macro) != (6,):
            raise ValueError("Six engineering macro strain components required")
        return self.avg_stress_planes(self._to_planes(u), np.asarray(macro, dtype=float)[None])

    def project_zero_mean(self, value):
        return self._from_planes(self.project_planes(self._to_planes(value)))

    def _compute_reference_medium(self, strategy="zeman", *, verbose=False):
        """(1, 6, 6) isotropic Zeman medium from the sampling lattice
        (HybridFIOperator._compute_reference_medium, hybrid_adapter.py:101-125)."""
        if strategy != "zeman":
            raise ValueError("This controlled prototype uses the frozen zeman reference only")
        from .reference_medium import zeman
        return zeman(self.field, self.shape)[None]

    def enable_preconditioner(self, lam=None, mu=None, *, precision="float32"):
        """Attach the FFT Green preconditioner of the isotropic medium (lam, mu)
        (default: the Zeman medium) and expose the reference's preconditioner
        protocol on the operator: ``apply_precond`` (FP64 (1,3N) node-major in
        and out) and ``precond_packed`` (the raw FP32 component-plane apply of
        chain_adapter.attach_raw, chain_adapter.py:17-31).  Returns it."""
        if lam is None or mu is None:
            ref = self._compute_reference_medium()[0]
            lam, mu = float(ref[0, 1]), float(ref[3, 3])
        old = getattr(self, "_preconditioner", None)
        if old is not None:
            old.close()
        self._preconditioner = Preconditioner(self, lam, mu, precision=precision, nrhs_max=1)
        return self._preconditioner

    def _attached(self):
        pc = getattr(self, "_preconditioner", None)
        if pc is None:
            raise RuntimeError("no preconditioner attached; call enable_preconditioner first")
        return pc

    def apply_precond(self, residual):
        """z = G0 r on (1, 3N) FP64 node-major (the operator's apply_precond)."""
        return self._attached().apply_precond(residual)

    def precond_packed(self, packed):
        """Raw preconditioner apply on FP32 component planes (B, 3, Nz, Ny, Nx)
        -> new FP32 tensor of the same shape (chain_adapter.py:17-31; the input
        is not modified)."""
        return self._attached().precond_packed(packed)

    def recover_center_fields(self, u, macro):
        """Reference protocol: element-centre (strain, stress), each (1, N, 6) FP64."""
        if np.shape(macro) != (6,):
            raise ValueError("Six engineering macro strain components required")
        eps, sig, _ = self.center_fields_planes(self._to_planes(u), np.asarray(macro, dtype=float)[None],
                                                energy=False)
        return eps.transpose(1, 2), sig.transpose(1, 2)


class Preconditioner:
    """FFT Green preconditioner of the isotropic reference medium (lam, mu)."""

    def __init__(self, cell, lam, mu, *, precision="float32", nrhs_max=6):
        self.cell = cell
        torch = cell.torch
        if precision not in ("float32", "fp32", "mixed", "bf16", "float64", "fp64"):
            raise ValueError("precision must be float32, bf16 (FP32 arithmetic, BF16 spectra) or float64")
        self.dtype = torch.float64 if precision in ("float64", "fp64") else torch.float32
        prec = {"bf16": _native.VXL_F32_BF16}.get(
            precision, _native.VXL_F32 if self.dtype == torch.float32 else _native.VXL_F64)
        handle = ctypes.c_void_p()
        check(cell.lib.vxl_precond_create(ctypes.byref(handle), cell.handle, prec, float(lam), float(mu),
                                          int(nrhs_max)), "vxl_precond_create")
        self.handle = handle
        self.nrhs_max = nrhs_max

    def apply_planes(self, r, out=None):
        torch = self.cell.torch
        if r.dtype != self.dtype or r.dim() != 3 or r.shape[1:] != (3, self.cell.n):
            raise ValueError("expected (R, 3, N) in the preconditioner's precision")
        r = r.contiguous()
        out = torch.empty_like(r) if out is None else out
        with self.cell.ordered():
            check(self.cell.lib.vxl_precond_apply(self.handle, r.shape[0], dptr(r), dptr(out)), "vxl_precond_apply")
        return out

    def apply_precond(self, residual):
        """Reference protocol: (1,3N) FP64 node-major in and out."""
        planes = self.cell._to_planes(residual).to(self.dtype)
        return self.cell._from_planes(self.apply_planes(planes).to(self.cell.torch.float64))

    def precond_packed(self, packed):
        """(B, 3, Nz, Ny, Nx) component planes in the preconditioner's
        precision -> new tensor of the same shape and type."""
        shape = (3,) + self.cell.shape
        if packed.dim() != 5 or tuple(packed.shape[1:]) != shape or packed.dtype != self.dtype:
            raise ValueError(f"expected (B, {', '.join(map(str, shape))}) planes in {self.dtype}")
        if packed.shape[0] > self.nrhs_max:
            out = [self.precond_packed(packed[b:b + self.nrhs_max]) for b in range(0, packed.shape[0], self.nrhs_max)]
            return self.cell.torch.cat(out)
        flat = packed.reshape(packed.shape[0], 3, self.cell.n)
        return self.apply_planes(flat).view_as(packed)

    def close(self):
        if getattr(self, "handle", None):
            self.cell.lib.vxl_precond_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Solver:
    """Block multi-RHS C_eff solver (mode 'mixed' or 'fp64')."""

    PRECONDITIONERS = {"fft": _native.VXL_PRECOND_FFT, "jacobi": _native.VXL_PRECOND_JACOBI,
                       "fft_bf16": _native.VXL_PRECOND_FFT_BF16}

    METHODS = {"cg": _native.VXL_METHOD_CG, "block": _native.VXL_METHOD_BLOCK_CG}

    def __init__(self, cell, lam, mu, *, mode="mixed", nrhs=None, preconditioner="fft", method="cg"):
        """lam, mu: the isotropic reference medium; for a stacked cell
        (Cell.stack) one value per cell.  nrhs: default 6 per cell."""
        self.cell = cell
        self.mode = mode
        code = {"mixed": _native.VXL_SOLVE_MIXED, "fp64": _native.VXL_SOLVE_FP64}.get(mode)
        if code is None:
            raise ValueError("precision must be 'mixed' or 'fp64'")
        groups = getattr(cell, "groups", 1)
        lams, mus = np.atleast_1d(np.asarray(lam, dtype=np.float64)), np.atleast_1d(np.asarray(mu, dtype=np.float64))
        if groups > 1 and (lams.shape != (groups,) or mus.shape != (groups,)):
            raise ValueError("a stacked cell needs one lam and one mu per cell")
        nrhs = 6 * groups if nrhs is None else int(nrhs)
        handle = ctypes.c_void_p()
        check(cell.lib.vxl_solver_create(ctypes.byref(handle), cell.handle, code, nrhs, float(lams[0]),
                                         float(mus[0])), "vxl_solver_create")
        self.handle = handle
        self.nrhs = nrhs
        self.groups = groups
        if groups > 1:
            lp, mp = np.ascontiguousarray(lams), np.ascontiguousarray(mus)
            check(cell.lib.vxl_solver_set_media(handle, lp.ctypes.data_as(_native._DP), mp.ctypes.data_as(_native._DP)),
                  "vxl_solver_set_media")
        self.preconditioner = "fft"
        if preconditioner != "fft":
            self.set_preconditioner(preconditioner)
        self.method = "cg"
        if method != "cg":
            self.set_method(method)

    def default_macro(self):
        """The six unit engineering strains, once per cell of a stacked cell
        (the first nrhs of them for one cell)."""
        if self.groups > 1:
            return np.tile(np.eye(6), (self.groups, 1))[: self.nrhs]
        return np.eye(6)[: self.nrhs]

    def set_method(self, method):
        """'cg' (one CG recurrence per load case, sharing every launch) or
        'block' (block CG over all load cases, FFT preconditioner, <= 8 cases)."""
        code = self.METHODS.get(method)
        if code is None:
            raise ValueError(f"method must be one of {sorted(self.METHODS)}")
        check(self.cell.lib.vxl_solver_set_method(self.handle, code), "vxl_solver_set_method")
        self.method = method

    def set_preconditioner(self, kind):
        """'fft' (Green operator of the reference medium), 'fft_bf16' (same,
        spectra stored in BF16; mixed mode) or 'jacobi' (1 / diag K control)."""
        code = self.PRECONDITIONERS.get(kind)
        if code is None:
            raise ValueError(f"preconditioner must be one of {sorted(self.PRECONDITIONERS)}")
        check(self.cell.lib.vxl_solver_set_preconditioner(self.handle, code), "vxl_solver_set_preconditioner")
        self.preconditioner = kind

    def run(self, macro=None, *, tol=1e-8, max_iter=10000, return_fields=False):
        torch = self.cell.torch
        macro = self.default_macro() if macro is None else np.atleast_2d(macro)
        if macro.shape != (self.nrhs, 6):
            raise ValueError("macro strains must be (nrhs, 6)")
        m, mp = darray(macro)
        ch = np.empty((6, self.nrhs))
        its = (ctypes.c_int * self.nrhs)()
        rst = (ctypes.c_int * self.nrhs)()
        res = (ctypes.c_double * self.nrhs)()
        fields = None
        if return_fields:
            fields = torch.empty((self.nrhs, 3, self.cell.n), dtype=torch.float64, device=self.cell.device)
        with self.cell.ordered():
            rc = self.cell.lib.vxl_solver_run(self.handle, mp, float(tol), int(max_iter),
                                              ch.ctypes.data_as(_native._DP), its, rst, res,
                                              dptr(fields) if fields is not None else None)
        if rc not in (_native.VXL_OK, _native.VXL_ERR_CONVERGENCE):
            check(rc, "vxl_solver_run")  # before any other call replaces the error text
        error = None if rc == _native.VXL_OK else self.cell.lib.vxl_last_error()
        ms = (ctypes.c_double * 3)()
        self.cell.lib.vxl_solver_timings(self.handle, ms)
        out = dict(CH=ch, iterations=list(its), restarts=list(rst), true_residuals=list(res),
                   device_ms=dict(setup_rhs=ms[0], iterations=ms[1], certify_stress=ms[2]),
                   graph_blocks=int(self.cell.lib.vxl_solver_graph_blocks(self.handle)),
                   trace=[self.trace(k) for k in range(self.nrhs)])
        if fields is not None:
            out["fields"] = fields
        if error is not None:
            exc = _native.ConvergenceError(f"vxl_solver_run: {error.decode(errors='replace')}")
            exc.partial = out   # the last iterate (fields, CH of it, residual history)
            raise exc
        return out

    def trace(self, k):
        """[(iteration, relative residual)] of system k; negative iteration =
        FP64 true residual."""
        lib = self.cell.lib
        count = lib.vxl_solver_trace(self.handle, k, None, None, 0)
        if count < 0:
            check(count, "vxl_solver_trace")
        its = (ctypes.c_int * max(count, 1))()
        res = (ctypes.c_double * max(count, 1))()
        lib.vxl_solver_trace(self.handle, k, its, res, count)
        return [(int(its[i]), float(res[i])) for i in range(count)]

    def close(self):
        if getattr(self, "handle", None):
            self.cell.lib.vxl_solver_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_batch(solvers, macro=None, *, tol=1e-8, max_iter=10000, workers=4):
    """Run independent solvers (each on its own cell and stream) concurrently
    from `workers` host threads, largest cells first; returns one dict per
    solver with CH (6 x R), iterations, restarts and true_residuals."""
    if not solvers:
        return []
    lib = solvers[0].cell.lib
    Rs = [s.nrhs for s in solvers]
    total = sum(Rs)
    if macro is None:
        macro = np.concatenate([s.default_macro() for s in solvers])
    macro, mp = darray(macro)
    if macro.shape != (total, 6):
        raise ValueError("macro strains must stack (R_i, 6) blocks of every solver")
    handles = (ctypes.c_void_p * len(solvers))(*[s.handle.value for s in solvers])
    ch = np.empty(6 * total)
    its = (ctypes.c_int * total)()
    rst = (ctypes.c_int * total)()
    res = (ctypes.c_double * total)()
    for s in solvers:
        s.cell.stream.synchronize()
    check(lib.vxl_solver_run_batch(handles, len(solvers), mp, float(tol), int(max_iter),
                                   ch.ctypes.data_as(_native._DP), its, rst, res, int(workers)),
          "vxl_solver_run_batch")
    out, o = [], 0
    for s, R in zip(solvers, Rs):
        out.append(dict(CH=ch[6 * o:6 * (o + R)].reshape(6, R).copy(), iterations=list(its[o:o + R]),
                        restarts=list(rst[o:o + R]), true_residuals=list(res[o:o + R]),
                        graph_blocks=int(lib.vxl_solver_graph_blocks(s.handle))))
        o += R
    return out
