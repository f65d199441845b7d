"""Oracle isotropic reference-medium selection, FP64 NumPy/SciPy.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates
  * ``src/voxel_lmfem/_core/materials/hybrid_adapter.py:101-125`` (Zeman
    sampling: every max(1, n//16) voxel per axis; mu_s = (C44+C55+C66)/4,
    K_s = (sum Cii + 2 sum Cij)/9, midpoint of extremes, lam = max(0, K-2mu/3));
  * ``src/voxel_lmfem/_core/solver_chain/reference_closed_form.py:45-171`` with
    ``reference_design.py:25-88`` (ordered isotropic soft phase, scaled to the
    baseline shear modulus, kept only if it strictly lowers the energy bound).

Pinning
  * the closed form's optimality statement: for C0 = s*Cm the bound equals
    max_e lambda_max(Ce, Cm) (checked numerically on random SPD phases);
  * the reference's own self-check fixture
    (independent_validation.py:191-197): three unordered isotropic phases
    give status "unsupported" and return the baseline unchanged.
"""
import numpy as np
from scipy.linalg import eigvalsh

from .materials import PAIRS


def isotropic_stiffness(bulk, mu):
    lam = bulk - 2.0 * mu / 3.0
    C = np.zeros((6, 6))
    C[:3, :3] = lam
    C[np.arange(3), np.arange(3)] += 2.0 * mu
    C[np.arange(3, 6), np.arange(3, 6)] = mu
    return C


def zeman(c21_rows_of_voxel, shape):
    """Zeman reference from the per-voxel C21 rows (callable index -> rows).

    ``c21_rows_of_voxel(flat_positions)`` returns (k, 21) packed tensors.
    Returns the isotropic 6x6 reference.
    """
    nz, ny, nx = shape
    z = np.arange(0, nz, max(1, nz // 16))
    y = np.arange(0, ny, max(1, ny // 16))
    x = np.arange(0, nx, max(1, nx // 16))
    positions = (z[:, None, None] * (ny * nx) + y[None, :, None] * nx + x[None, None, :]).reshape(-1)
    packed = np.asarray(c21_rows_of_voxel(positions))
    sub = np.zeros((len(packed), 6, 6))
    for k, (i, j) in enumerate(PAIRS):
        sub[:, i, j] = sub[:, j, i] = packed[:, k]
    mus = .5 * (sub[:, 3, 3] + sub[:, 4, 4] + sub[:, 5, 5]) / 2.
    Ks = (sub[:, 0, 0] + sub[:, 1, 1] + sub[:, 2, 2]
          + 2 * (sub[:, 0, 1] + sub[:, 0, 2] + sub[:, 1, 2])) / 9.
    mu = .5 * (float(mus.max()) + float(mus.min()))
    K = .5 * (float(Ks.max()) + float(Ks.min()))
    lam = max(0., K - 2 * mu / 3.)
    out = np.zeros((6, 6))
    out[:3, :3] = lam
    out[np.arange(3), np.arange(3)] += 2 * mu
    out[np.arange(3, 6), np.arange(3, 6)] = mu
    return out


def bound(phases, reference):
    spectra = [eigvalsh(c, reference, check_finite=False) for c in phases]
    a = float(min(s[0] for s in spectra))
    b = float(max(s[-1] for s in spectra))
    return dict(a=a, b=b, condition_bound=b / a)


def _isotropic(matrix, tolerance):
    bulk = float(np.sum(matrix[:3, :3]) / 9.0)
    mu = float(np.trace(matrix[3:, 3:]) / 3.0)
    candidate = isotropic_stiffness(bulk, mu)
    error = float(np.max(np.abs(matrix - candidate)) / np.max(np.abs(matrix)))
    return error <= tolerance


def design_closed(phases, baseline, *, isotropy_rtol=1e-10, order_rtol=1e-10, tie_rtol=1e-10):
    """reference_closed_form.py:45-171 (kind 'fi', fallback 'baseline').

    Returns (reference 6x6, status) with status 'closed_form', 'baseline'
    (closed form exists but does not improve), or 'unsupported'.
    """
    phases = [0.5 * (np.asarray(c, float) + np.asarray(c, float).T) for c in phases]
    baseline = 0.5 * (np.asarray(baseline, float) + np.asarray(baseline, float).T)
    baseline_bound = bound(phases, baseline)
    mu = float(baseline[3, 3])
    candidates = []
    for soft in phases:
        if not _isotropic(soft, isotropy_rtol):
            continue
        spectra = [eigvalsh(c, soft, check_finite=False) for c in phases]
        if min(v[0] for v in spectra) >= 1.0 - order_rtol:
            reference = (mu / float(soft[3, 3])) * soft
            candidates.append((reference, bound(phases, reference)))
    if candidates:
        reference, b = min(candidates, key=lambda item: item[1]["condition_bound"])
        if baseline_bound["condition_bound"] <= b["condition_bound"] * (1.0 + tie_rtol):
            return baseline.copy(), "baseline"
        return reference, "closed_form"
    return baseline.copy(), "unsupported"
