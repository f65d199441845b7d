"""Oracle preconditioned CG with the reference's safeguards, FP64 NumPy.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates ``src/voxel_lmfem/_core/runtime/repro/source/femlib/batched/pcg.py:20-252``
(eager mode): alpha = rho/pAp with the update frozen when pAp is non-finite
or <= 0, stop when ||r||/||f|| < tol, best-iterate bookkeeping; and the outer
true-residual restart loop of ``src/voxel_lmfem/_core/runtime/benchmark.py:77-90``
(zero RHS -> zero solution, zero-mean projection, restart from x0 = u,
at most 10000 iterations in total).

Pinning: CG on a small SPD system reaches the scipy direct solution; the
restart loop certifies ||f - K u||/||f|| <= tol by recomputation.
"""
import numpy as np


def pcg(apply, precond, rhs, *, tol=1e-8, max_iter=2000, x0=None):
    """Return (best_u, iterations, best_residual)."""
    rhs = np.asarray(rhs, dtype=np.float64)
    norm = max(float(np.linalg.norm(rhs)), 1e-30)
    if x0 is None:
        u = np.zeros_like(rhs)
        r = rhs.copy()
    else:
        u = np.array(x0, dtype=np.float64)
        r = rhs - apply(u)
    z = precond(r)
    p = z.copy()
    rz = float(np.vdot(r, z))
    best_u, best_res = u.copy(), np.inf
    for it in range(max_iter):
        Ap = apply(p)
        pAp = float(np.vdot(p, Ap))
        alpha = 0.0 if (not np.isfinite(pAp) or pAp <= 0) else rz / max(pAp, 1e-30)
        u = u + alpha * p
        r = r - alpha * Ap
        res = float(np.linalg.norm(r)) / norm
        if res < best_res:
            best_u, best_res = u.copy(), res
        if res < tol:
            return best_u, it + 1, best_res
        z = precond(r)
        rz_new = float(np.vdot(r, z))
        beta = rz_new / max(rz, 1e-30)
        p = z + beta * p
        rz = rz_new
    return best_u, max_iter, best_res


def solve_certified(apply, precond, rhs, project, *, tol=1e-8, total_max=10000):
    """benchmark.py:77-90 outer loop; returns (u, iterations, restarts, residual)."""
    rhs = np.asarray(rhs, dtype=np.float64)
    norm = float(np.linalg.norm(rhs))
    if norm == 0:
        return np.zeros_like(rhs), 0, 0, 0.0
    u, total, restarts = None, 0, 0
    while total < total_max:
        u, count, _ = pcg(apply, precond, rhs, tol=tol, max_iter=total_max - total, x0=u)
        total += int(count)
        u = project(u)
        rr = float(np.linalg.norm(rhs - apply(u))) / norm
        if not np.isfinite(rr):
            raise RuntimeError("Nonfinite true residual")
        if rr <= tol:
            return u, total, restarts, rr
        if not count:
            break
        restarts += 1
    raise RuntimeError(f"Unconverged at {total}: {rr}")
