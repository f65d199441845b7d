"""Isotropic reference medium (lam0, mu0) of the FFT Green preconditioner.

Same selection as the reference chain:
  * Zeman sampling (src/voxel_lmfem/_core/materials/hybrid_adapter.py:101-125):
    every max(1, n//16)-th voxel per axis; per sample mu_s = (C44+C55+C66)/4
    and K_s = (sum Cii + 2 sum_{i<j<=3} Cij)/9; mu0 and K0 are the midpoints of
    the sampled extremes; lam0 = max(0, K0 - 2 mu0/3);
  * closed-form soft-phase design
    (src/voxel_lmfem/_core/solver_chain/reference_closed_form.py:45-171):
    an isotropic phase that is Loewner-below every used phase, scaled to the
    Zeman shear modulus, replaces the Zeman medium when it strictly lowers the
    material energy bound max lambda_max / min lambda_min.
The choice changes iteration counts, never the converged C_eff.
"""
import numpy as np


def _iso(lam, mu):
    C = np.zeros((6, 6))
    C[:3, :3] = lam
    C[np.arange(3), np.arange(3)] += 2 * mu
    C[np.arange(3, 6), np.arange(3, 6)] = mu
    return C


def zeman(field, shape):
    c = field.lattice_rows(shape)  # (k, 21) packed: indices of C44,C55,C66 = 15, 18, 20
    mus = .5 * (c[:, 15] + c[:, 18] + c[:, 20]) / 2.
    Ks = (c[:, 0] + c[:, 6] + c[:, 11] + 2 * (c[:, 1] + c[:, 2] + c[:, 7])) / 9.
    mu = .5 * (float(mus.max()) + float(mus.min()))
    K = .5 * (float(Ks.max()) + float(Ks.min()))
    return _iso(max(0., K - 2 * mu / 3.), mu)


def _spectra(phases, reference):
    """(k, 6) ascending generalized eigenvalues of each phase against the SPD
    reference (Cholesky reduction L^-1 C L^-T, as LAPACK's sygv)."""
    L = np.linalg.cholesky(reference)
    Li = np.linalg.inv(L)
    M = Li @ np.asarray(phases) @ Li.T
    return np.linalg.eigvalsh(0.5 * (M + M.transpose(0, 2, 1)))


def _bound(phases, reference):
    spectra = _spectra(phases, reference)
    return float(spectra[:, -1].max()) / float(spectra[:, 0].min())


def _is_isotropic(C, rtol=1e-10):
    """Per tensor of a (k, 6, 6) stack (or one 6x6): equal to the isotropic
    tensor with its own bulk and shear moduli to rtol."""
    C = np.asarray(C, dtype=float)
    one = C.ndim == 2
    C = C[None] if one else C
    bulk = C[:, :3, :3].sum(axis=(1, 2)) / 9.0
    mu = np.trace(C[:, 3:, 3:], axis1=1, axis2=2) / 3.0
    lam = bulk - 2.0 * mu / 3.0
    cand = np.zeros_like(C)
    cand[:, :3, :3] = lam[:, None, None]
    cand[:, np.arange(3), np.arange(3)] += 2 * mu[:, None]
    cand[:, np.arange(3, 6), np.arange(3, 6)] = mu[:, None]
    ok = np.abs(C - cand).max(axis=(1, 2)) / np.abs(C).max(axis=(1, 2)) <= rtol
    return bool(ok[0]) if one else ok


def design(phases, baseline, *, order_rtol=1e-10, tie_rtol=1e-10):
    """Return (reference 6x6, status in {'closed_form','baseline','unsupported'})."""
    phases = np.asarray(phases, dtype=float)
    phases = 0.5 * (phases + phases.transpose(0, 2, 1))
    base_bound = _bound(phases, baseline)
    mu = float(baseline[3, 3])
    best = None
    for soft in phases[_is_isotropic(phases)]:
        spectra = _spectra(phases, soft)
        if float(spectra[:, 0].min()) < 1.0 - order_rtol:
            continue
        ref = (mu / float(soft[3, 3])) * soft
        b = _bound(phases, ref)
        if best is None or b < best[1]:
            best = (ref, b)
    if best is None:
        return baseline.copy(), "unsupported"
    if base_bound <= best[1] * (1.0 + tie_rtol):
        return baseline.copy(), "baseline"
    return best[0], "closed_form"


def select(field, phase, materials):
    """(lam0, mu0, metadata) for a cell."""
    base = zeman(field, phase.shape)
    used = np.asarray(field.used_labels(phase))
    phases = field.tensors[np.searchsorted(field.labels, used)]  # validated when the field was built
    ref, status = design(phases, base)
    lam, mu = float(ref[0, 1]), float(ref[3, 3])
    return lam, mu, dict(status=status, lam=lam, mu=mu, zeman_lam=float(base[0, 1]),
                         zeman_mu=float(base[3, 3]))
