"""Oracle C_eff solvers: the reference's CPU backend and an FP64 matrix-free
FFT-PCG solver for larger cells.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates
  * ``src/voxel_lmfem/inputs.py:16-48`` (input validation and messages);
  * ``src/voxel_lmfem/solver.py:20-51`` (CPU backend: 4096-voxel guard,
    fourth-order rotation of local-frame voxels, exact zero-fluctuation
    shortcut for a uniform tensor, sparse direct solve, true residuals);
  * ``src/voxel_lmfem/_core/runtime/benchmark.py:92-112`` (six unit loads,
    zero-mean RHS, certified solve, unsymmetrized CH columns) over the
    Gauss-B operator of ``oracle.fem`` and the Green preconditioner of
    ``oracle.spectral`` with the reference medium of ``oracle.reference_medium``.

Pinning: tests/test_solver.py:17-77 KATs (homogeneous, laminate Schur
formula, rotated anisotropic, invalid inputs, size guard) and the
tests/test_import.py:35 closed form; matrix-free agrees with sparse direct.
"""
import numpy as np

from . import fem, materials as mat, pcg, reference_medium, spectral


def validate(phase_id, materials, domain_size, orientation=None):
    """inputs.py:16-48."""
    phase = np.asarray(phase_id)
    if phase.ndim != 3 or phase.dtype.kind not in "iu" or min(phase.shape) < 2:
        raise ValueError("phase_id must be an integer (Nz, Ny, Nx) array with every dimension >= 2")
    domain = np.asarray(domain_size, dtype=float)
    if domain.shape not in ((3,), (1, 3)) or not np.isfinite(domain).all() or (domain <= 0).any():
        raise ValueError("domain_size must contain three finite positive lengths in (x, y, z) order")
    domain = domain.reshape(3)
    if not isinstance(materials, dict) or not materials:
        raise ValueError("materials must be a nonempty dictionary")
    table = {}
    for key, value in materials.items():
        if isinstance(key, bool) or not isinstance(key, (str, int, np.integer)):
            raise ValueError("Material labels must be integers or integer JSON keys")
        try:
            label = int(key)
        except ValueError as exc:
            raise ValueError(f"Invalid material label: {key!r}") from exc
        if label in table:
            raise ValueError(f"Duplicate material label after conversion: {label}")
        table[label] = value
    labels, tensors, local = mat.phase_table(table)
    used = np.unique(phase)
    if not np.isin(used, labels).all():
        raise ValueError(f"Undefined material labels: {used[~np.isin(used, labels)].tolist()}")
    for index in np.flatnonzero(local):
        selected = np.flatnonzero(phase.ravel() == labels[index])
        if selected.size and orientation is None:
            raise ValueError("Local-frame anisotropic phases require orientation")
        for start in range(0, len(selected), 65536):
            mat.orientation_values(orientation, phase.shape, selected[start:start + 65536])
    return phase, table, domain, orientation


def homogenize_direct(phase_id, materials, domain_size, *, orientation=None):
    """solver.py:20-51 (the CPU backend)."""
    phase, table, domain, orientation = validate(phase_id, materials, domain_size, orientation)
    if phase.size > 4096:
        raise ValueError("The CPU reference is limited to 4096 voxels; use the CUDA backend for larger cells")
    full = mat.full_field(phase, table, orientation)
    voxel = domain / np.array(phase.shape[::-1])
    if np.array_equal(full, np.broadcast_to(full[0], full.shape)):
        CH = full[0].copy()
        residuals = [0.] * 6
        kind = "homogeneous-exact-zero-fluctuation"
    else:
        K, rhs, u, CH, _ = fem.assemble(full, phase.shape, voxel)
        residuals = [float(np.linalg.norm(rhs[:, j] - K @ u[..., j].ravel()) /
                           max(np.linalg.norm(rhs[:, j]), 1e-300)) for j in range(6)]
        kind = "independent-sparse-direct"
    return dict(CH=CH, true_residuals=residuals, solution_kind=kind)


def reference_lame(phase, table, orientation, pool=None, index=None):
    """Zeman sample + closed-form design (hybrid_adapter.py:147-163)."""
    if pool is None:
        pool, index = mat.hybrid_pool(phase, table, orientation)
    base = reference_medium.zeman(lambda pos: pool[index[pos]], phase.shape)
    labels = np.unique(phase)
    phases = [mat.material_stiffness(table[int(k)]) for k in labels]
    ref, _ = reference_medium.design_closed(phases, base)
    return float(ref[0, 1]), float(ref[3, 3])


def homogenize_matfree(phase_id, materials, domain_size, *, orientation=None, tol=1e-8,
                       return_fields=False):
    """FP64 matrix-free FFT-PCG C_eff (benchmark.py:92-112 semantics)."""
    phase, table, domain, orientation = validate(phase_id, materials, domain_size, orientation)
    pool, index = mat.hybrid_pool(phase, table, orientation)
    full = mat.unpack21(pool[index])
    shape = phase.shape
    h = domain / np.array(shape[::-1])
    op = fem.MatrixFree(full, shape, h)
    lam, mu = reference_lame(phase, table, orientation, pool, index)
    precond = lambda r: spectral.green_apply(r, shape, h, lam, mu)
    CH = np.empty((6, 6))
    calls, fields = [], []
    for j in range(6):
        e = np.eye(6)[j]
        rhs = op.project_zero_mean(op.rhs(e))
        u, its, restarts, rr = pcg.solve_certified(op.apply, precond, rhs, op.project_zero_mean, tol=tol)
        CH[:, j] = op.avg_stress(u, e)
        calls.append(dict(iterations=its, restarts=restarts, true_residuals=[rr]))
        if return_fields:
            fields.append(u)
    out = dict(CH=CH, calls=calls, reference=(lam, mu))
    if return_fields:
        out["fields"] = fields
    return out
