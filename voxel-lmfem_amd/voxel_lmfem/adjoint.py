"""Differentiable C_eff: adjoint gradients behind torch.autograd.Function.

The reference has no gradient.  Here C_eff(theta) of a cell whose element
stiffness depends on per-voxel parameters is differentiated without a second
solve: the periodic cell problem is self-adjoint, so with the converged
load-case fluctuations u_j of the macro strains E_j and any symmetric weights W,

    d(sum_ij W_ij C_ij) / d t_e = (1/|Omega|) int_e (E_i + B u_i)^T (dC_e/dt_e) (E_j + B u_j) dV,

evaluated by ``vxl_energy_gradient`` (the operator's 2x2x2 Gauss rule) from
the fields the solve already holds.  Two parameterisations:

* ``density_ceff(rho, ...)``: SIMP density, C_e = E(rho_e) C1 with
  E(rho) = E_min + rho^p (E_max - E_min) and C1 the unit-modulus isotropic
  tensor; dC_e/drho_e = E'(rho_e) C1.
* ``angle_ceff(theta, ...)``: fiber angle of local-frame voxels, the frame R_e
  turned by theta_e about its third axis, C_e = Bond(R_e Rz(theta_e)) C_loc;
  dC_e/dtheta_e is formed on the host and contracted on the device.

Every element gets its own FP64 coefficient row (int32 index), so C_e is exact
in FP64 and finite differences of C_eff see the same model.

keep_fields: the six FP64 fields (144 B per voxel) and the cell stay on the
device between forward and backward (True), or only the host rows are kept and
backward solves again (False; the solve is deterministic, so the gradient is
bit-identical).  None (default) keeps them when they take at most a quarter of
the device memory free after the forward solve.
"""
import ctypes

import numpy as np

from . import _native
from . import materials as mat
from ._native import check, darray, dptr
from .reference_medium import zeman


def objective_weights(kind="c1111_plus_bulk"):
    """W of J = sum_ij W_ij C_ij.  'c1111_plus_bulk': J = C_1111 + K_V(C) with
    K_V = (sum_{i<3} C_ii + 2 sum_{i<j<3} C_ij) / 9."""
    if kind != "c1111_plus_bulk":
        raise ValueError("unknown objective")
    W = np.zeros((6, 6))
    W[0, 0] = 1.0
    W[:3, :3] += 1.0 / 9.0
    return W


class _RowField(mat.MaterialField):
    """One coefficient row per element (C order), int32 index."""

    def __init__(self, rows):
        super().__init__(rows, np.arange(len(rows), dtype=np.int32), None, None, None)


def _solve(rows, shape, h, precision, tol):
    from .gpu import Cell, Solver
    field = _RowField(rows)
    ref = zeman(field, shape)
    cell = Cell(field, shape, h)
    solver = None
    try:
        solver = Solver(cell, float(ref[0, 1]), float(ref[3, 3]), mode=precision, nrhs=6)
        out = solver.run(np.eye(6), tol=tol, return_fields=True)
    except BaseException:
        cell.close()  # the caller only receives (and later closes) the cell of a converged solve
        raise
    finally:
        if solver is not None:
            solver.close()
    return cell, out


def _keep(cell, keep_fields):
    """Whether to hold the fields until backward (see the module note)."""
    if keep_fields is not None:
        return bool(keep_fields)
    free, _ = cell.torch.cuda.mem_get_info(cell.device)
    return 6 * 3 * 8 * cell.n <= free // 4


def _contract(cell, u, W, d21=None):
    """out[e] = sum_ij W_ij int_e eps_i^T D_e eps_j (D_e: cell row, or d21[e])."""
    torch = cell.torch
    W = np.ascontiguousarray(0.5 * (np.asarray(W, dtype=float) + np.asarray(W, dtype=float).T))
    out = torch.empty(cell.n, dtype=torch.float64, device=cell.device)
    macro, mp = darray(np.eye(6))
    w, wp = darray(W)
    d = None
    if d21 is not None:
        d = torch.as_tensor(np.ascontiguousarray(d21, dtype=np.float64), device=cell.device)
    check(cell.lib.vxl_energy_gradient(cell.handle, 6, dptr(u), mp, wp, dptr(d) if d is not None else None,
                                       dptr(out)), "vxl_energy_gradient")
    return out


def _torch_function():
    import torch

    class DensityCeff(torch.autograd.Function):
        @staticmethod
        def forward(ctx, rho, shape, h, C1, E_min, E_max, p, precision, tol, keep_fields):
            r = rho.detach().cpu().numpy().astype(np.float64).reshape(-1)
            E = E_min + r ** p * (E_max - E_min)
            rows = E[:, None] * mat.pack21(C1)[None, :]
            cell, out = _solve(rows, shape, h, precision, tol)
            ctx.solve_args = (rows, shape, h, precision, tol)
            if _keep(cell, keep_fields):
                ctx.cell, ctx.u = cell, out["fields"]
            else:  # backward solves again
                cell.close()
                ctx.cell = ctx.u = None
            ctx.dE = p * r ** (p - 1) * (E_max - E_min)
            ctx.E = E
            ctx.volume = float(np.prod(h)) * cell.n
            ctx.rho_meta = (rho.shape, rho.dtype, rho.device)
            return torch.as_tensor(out["CH"], dtype=torch.float64)

        @staticmethod
        def backward(ctx, grad):
            # rows = E_e C1, so the row contraction is E_e int eps^T C1 eps.
            if ctx.cell is None:
                ctx.cell, out = _solve(*ctx.solve_args)
                ctx.u = out["fields"]
            g = _contract(ctx.cell, ctx.u, grad.detach().cpu().numpy()).cpu().numpy()
            d = ctx.dE * g / (ctx.E * ctx.volume)
            ctx.cell.close()
            shape, dtype, device = ctx.rho_meta
            return (torch.as_tensor(d.reshape(shape), dtype=dtype, device=device),) + (None,) * 9

    class AngleCeff(torch.autograd.Function):
        @staticmethod
        def forward(ctx, theta, shape, h, C_loc, frames, iso_rows, local_mask, precision, tol, keep_fields):
            t = theta.detach().cpu().numpy().astype(np.float64).reshape(-1)
            rows, drows = angle_rows(t, C_loc, frames, iso_rows, local_mask)
            cell, out = _solve(rows, shape, h, precision, tol)
            ctx.drows = drows
            ctx.solve_args = (rows, shape, h, precision, tol)
            if _keep(cell, keep_fields):
                ctx.cell, ctx.u = cell, out["fields"]
            else:  # backward solves again
                cell.close()
                ctx.cell = ctx.u = None
            ctx.volume = float(np.prod(h)) * cell.n
            ctx.meta = (theta.shape, theta.dtype, theta.device)
            return torch.as_tensor(out["CH"], dtype=torch.float64)

        @staticmethod
        def backward(ctx, grad):
            if ctx.cell is None:
                ctx.cell, out = _solve(*ctx.solve_args)
                ctx.u = out["fields"]
            g = _contract(ctx.cell, ctx.u, grad.detach().cpu().numpy(), d21=ctx.drows).cpu().numpy()
            ctx.cell.close()
            shape, dtype, device = ctx.meta
            return (torch.as_tensor((g / ctx.volume).reshape(shape), dtype=dtype, device=device),) + (None,) * 9

    return DensityCeff, AngleCeff


def density_ceff(rho, domain, *, C1=None, nu=0.3, E_min=1e-3, E_max=70.0, p=3.0, precision="fp64", tol=1e-8,
                 keep_fields=None):
    """C_eff (6x6 torch FP64) of a SIMP density cell rho (nz, ny, nx) torch
    tensor, differentiable with respect to rho through the adjoint.
    keep_fields: hold the fields for backward (True), solve again in backward
    (False), or decide from free device memory (None)."""
    DensityCeff, _ = _torch_function()
    shape = tuple(int(v) for v in rho.shape)
    h = np.asarray(domain, dtype=float) / np.array(shape[::-1], dtype=float)
    if C1 is None:
        C1 = mat.isotropic_voigt(1.0, nu)
    return DensityCeff.apply(rho, shape, h, np.asarray(C1, dtype=float), float(E_min), float(E_max), float(p),
                             precision, tol, keep_fields)


def _rz(t):
    c, s = np.cos(t), np.sin(t)
    R = np.zeros(t.shape + (3, 3))
    R[..., 0, 0], R[..., 0, 1], R[..., 1, 0], R[..., 1, 1], R[..., 2, 2] = c, -s, s, c, 1.0
    return R


def _drz(t):
    c, s = np.cos(t), np.sin(t)
    R = np.zeros(t.shape + (3, 3))
    R[..., 0, 0], R[..., 0, 1], R[..., 1, 0], R[..., 1, 1] = -s, -c, c, -s
    return R


def _bond(R):
    """Bond stress matrices T (k,6,6) of rotations R (k,3,3): C' = T C T^T."""
    i, j = [1, 0, 0], [2, 2, 1]
    r1, r2 = R[:, i, :], R[:, j, :]
    return np.concatenate((np.concatenate((R * R, 2 * R[:, :, i] * R[:, :, j]), -1),
                           np.concatenate((r1 * r2, r1[:, :, i] * r2[:, :, j] + r1[:, :, j] * r2[:, :, i]), -1)),
                          -2)


def _dbond(R, dR):
    """Derivative of _bond along dR (T is quadratic in R)."""
    i, j = [1, 0, 0], [2, 2, 1]
    r1, r2 = R[:, i, :], R[:, j, :]
    d1, d2 = dR[:, i, :], dR[:, j, :]
    top = np.concatenate((2 * R * dR, 2 * (dR[:, :, i] * R[:, :, j] + R[:, :, i] * dR[:, :, j])), -1)
    bot = np.concatenate((d1 * r2 + r1 * d2,
                          d1[:, :, i] * r2[:, :, j] + r1[:, :, i] * d2[:, :, j]
                          + d1[:, :, j] * r2[:, :, i] + r1[:, :, j] * d2[:, :, i]), -1)
    return np.concatenate((top, bot), -2)


def angle_rows(theta, C_loc, frames, iso_rows, local_mask):
    """Per-element rows of C_e(theta) and of dC_e/dtheta (zero outside local voxels).

    frames (k,3,3) are the frames of the local voxels (local_mask True, C
    order); iso_rows (N,21) hold the rows of every other voxel."""
    rows = np.array(iso_rows, dtype=np.float64, copy=True)
    drows = np.zeros_like(rows)
    idx = np.flatnonzero(local_mask)
    t = theta[idx]
    R = frames @ _rz(t)
    dR = frames @ _drz(t)
    T = _bond(R)
    dT = _dbond(R, dR)
    C = np.asarray(C_loc, dtype=float)
    rows[idx] = mat.pack21(T @ C @ T.transpose(0, 2, 1))
    dC = dT @ C @ T.transpose(0, 2, 1)
    drows[idx] = mat.pack21(dC + dC.transpose(0, 2, 1))
    return rows, drows


def angle_ceff(theta, domain, C_loc, frames, iso_rows, local_mask, *, precision="fp64", tol=1e-8,
               keep_fields=None):
    """C_eff (6x6 torch FP64) as a function of per-voxel fiber angles theta
    (nz, ny, nx) torch tensor (radians, 0 = the given frames); voxels outside
    local_mask keep iso_rows.  keep_fields as in density_ceff."""
    _, AngleCeff = _torch_function()
    shape = tuple(int(v) for v in theta.shape)
    h = np.asarray(domain, dtype=float) / np.array(shape[::-1], dtype=float)
    return AngleCeff.apply(theta, shape, h, np.asarray(C_loc, dtype=float), np.asarray(frames, dtype=float),
                           np.asarray(iso_rows, dtype=float), np.asarray(local_mask, dtype=bool).reshape(-1),
                           precision, tol, keep_fields)
