"""Seeded synthetic cells: the reference's analytic weave and the benchmark cells.

Every generator returns ``(phase_id (Nz,Ny,Nx) int, materials dict,
domain_size (x,y,z), orientation or None)`` -- the arguments of
``homogenize`` -- plus a small metadata dict.  Moduli are in the units given
(GPa for the benchmark cells, MPa for the reference weave).
"""
import hashlib

import numpy as np

from .materials import bond_rotate, unpack21


def reference_weave(N):
    """The reference's fixed analytic weave (voxel-centre classification).

    Restates the geometry used for every woven record of the reference
    (src/voxel_lmfem/_core/runtime/repro/extension_20261006/run_resolution.py:14-32,
    materials material_control.py:5-6 and the legacy adapter
    material_field.py:168-187): four vertical elliptical yarns (two along x,
    two along y) following z = +-A cos(2 pi s), yarn label 1 with an
    orthotropic local frame (axis1 = yarn tangent, axis2 = in-plane normal),
    matrix label 0.  Returns the homogenize arguments and the SHA-256 of
    (yarn_id, orient1, orient2) used by the reference's records.
    """
    width, thickness, amplitude, height = .18, .035, .09, .4
    x = (np.arange(N) + .5) / N
    y = x.copy()
    z = (x - .5) * height
    X, Y = np.meshgrid(x, y, indexing="xy")
    ids = np.full((N, N, N), -1, dtype=np.int8)
    o1 = np.zeros((N, N, N, 3), dtype=np.float64)
    o2 = np.zeros_like(o1)
    for k, zz in enumerate(z):
        for i, (axis, center, sign) in enumerate([(0, 0., 1), (0, .5, -1), (1, 0., -1), (1, .5, 1)]):
            s = X if axis == 0 else Y
            t = Y if axis == 0 else X
            d = (t - center + .5) % 1 - .5
            zc = sign * amplitude * np.cos(2 * np.pi * s)
            dz = (zz - zc + height / 2) % height - height / 2
            inside = (d / width) ** 2 + (dz / thickness) ** 2 < 1
            if np.any(inside & (ids[k] >= 0)):
                raise ValueError("Overlapping yarn voxels")
            ids[k][inside] = i
            slope = -sign * amplitude * 2 * np.pi * np.sin(2 * np.pi * s[inside])
            norm = np.sqrt(1 + slope * slope)
            v = np.zeros((len(slope), 3))
            v[:, axis] = 1 / norm
            v[:, 2] = slope / norm
            w = np.zeros_like(v)
            w[:, 1 - axis] = 1 if axis == 0 else -1
            o1[k][inside] = v
            o2[k][inside] = w
    digest = hashlib.sha256()
    for arr in (ids[None], o1[None], o2[None]):
        digest.update(np.ascontiguousarray(arr).tobytes())
    phase = (ids >= 0).astype(np.uint8)
    materials = {0: dict(kind="isotropic", E=3500., nu=.35),
                 1: dict(kind="orthotropic", frame="local", E1=135000., E2=9000., E3=9000.,
                         nu12=.30, nu13=.30, nu23=.35, G12=5000., G13=5000., G23=3500.)}
    orientation = dict(axis1=o1, axis2=o2)
    domain = np.array([1., 1., height])
    return (phase, materials, domain, orientation), dict(input_hash=digest.hexdigest(),
                                                          yarn_fraction=float(phase.mean()))


def periodic_spheres(n=32, radius=3.0, fraction=0.30, seed=1, E=(3.0, 70.0), nu=(0.35, 0.22)):
    """Two-phase cell: spheres at seeded uniform centres until vf >= fraction.

    Voxel centres within `radius` (periodic minimum image, voxel units) of a
    centre are inclusion (label 1); matrix label 0.
    """
    rng = np.random.default_rng(seed)
    c = np.arange(n) + .5
    inside = np.zeros((n, n, n), dtype=bool)
    centres = []
    count, total = 0, n ** 3
    reach = int(np.ceil(radius)) + 1
    while count / total < fraction:
        p = rng.uniform(0, n, size=3)
        centres.append(p)
        # Only the sphere's periodic bounding box can change; the distance
        # test is evaluated there exactly as over the whole grid.
        win = []
        for a in (2, 1, 0):  # z, y, x
            lo = int(np.floor(p[a])) - reach
            idx = np.unique(np.arange(lo, lo + 2 * reach + 2) % n)
            v = np.abs(c[idx] - p[a])
            win.append((idx, np.minimum(v, n - v)))
        (iz, dz), (iy, dy), (ix, dx) = win
        hit = (dz[:, None, None] ** 2 + dy[None, :, None] ** 2 + dx[None, None, :] ** 2) <= radius ** 2
        box = np.ix_(iz, iy, ix)
        count += int((hit & ~inside[box]).sum())
        inside[box] |= hit
    phase = inside.astype(np.uint8)
    materials = {0: dict(kind="isotropic", E=float(E[0]), nu=float(nu[0])),
                 1: dict(kind="isotropic", E=float(E[1]), nu=float(nu[1]))}
    return (phase, materials, np.ones(3), None), dict(spheres=len(centres), fraction=float(phase.mean()),
                                                     seed=seed)


def random_rotations(count, rng):
    """Uniform random rotations from Shoemake's quaternion method."""
    u1, u2, u3 = rng.uniform(size=(3, count))
    q = np.stack([np.sqrt(1 - u1) * np.sin(2 * np.pi * u2), np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
                  np.sqrt(u1) * np.sin(2 * np.pi * u3), np.sqrt(u1) * np.cos(2 * np.pi * u3)], axis=1)
    x, y, z, w = q.T
    R = np.empty((count, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - z * w)
    R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w)
    R[:, 2, 1] = 2 * (y * z + x * w)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def cubic_voigt(C11, C12, C44):
    C = np.zeros((6, 6))
    C[:3, :3] = C12
    C[np.arange(3), np.arange(3)] = C11
    C[np.arange(3, 6), np.arange(3, 6)] = C44
    return C


def voronoi_polycrystal(n=256, grains=200, seed=3, C11=168.4, C12=121.4, C44=75.4):
    """Periodic Voronoi polycrystal: grain label per voxel + one rotation per grain.

    Seeds uniform in [0,n)^3; voxel centres take the nearest seed under the
    periodic minimum image (ties to the lower seed index).  Each grain is a
    global-frame anisotropic phase whose tensor is the cubic tensor rotated by
    that grain's uniform random rotation (labels carry the per-grain frame).
    """
    from scipy.spatial import cKDTree
    rng = np.random.default_rng(seed)
    seeds = rng.uniform(0, n, size=(grains, 3))
    R = random_rotations(grains, rng)
    tree = cKDTree(seeds, boxsize=n)
    c = np.arange(n) + .5
    phase = np.empty((n, n, n), dtype=np.uint8 if grains <= 256 else np.uint16)
    for z in range(n):
        zz, yy, xx = np.meshgrid([c[z]], c, c, indexing="ij")
        pts = np.stack([xx.ravel(), yy.ravel(), zz.ravel()], axis=1)
        _, idx = tree.query(pts, k=1, workers=-1)
        phase[z] = idx.reshape(n, n)
    C = cubic_voigt(C11, C12, C44)
    rotated = unpack21(bond_rotate(C, R))
    materials = {g: dict(kind="anisotropic", C=rotated[g]) for g in range(grains)}
    return (phase, materials, np.ones(3), None), dict(grains=grains, seed=seed, rotations=R)


def woven_like(n=128, seed=2, tows=4, section=(4, 12), amplitude=6.0):
    """Woven-like cell: matrix (0), warp tows along x (1), weft tows along y (2).

    Tow centre-lines: warp j at y = (j+1/2) n/tows, z = n/2 + (-1)^j A sin(2 pi x/n);
    weft j at x = (j+1/2) n/tows, z = n/2 - (-1)^j A sin(2 pi y/n).  Elliptical
    sections `section` = (thickness, width) voxels.  Overlap rule: a voxel
    inside a warp and a weft tow is warp.  Tows are transversely isotropic
    (E1=230, E2=E3=15, G12=G13=27, nu12=nu13=0.2, nu23=0.3, G23=E2/(2(1+nu23))
    GPa) with axis1 = the local path tangent; matrix E=3.5 GPa, nu=0.35.
    `seed` is unused (the geometry is deterministic) and kept for a uniform API.
    """
    c = np.arange(n) + .5
    zc = n / 2
    a_t, a_w = section[0] / 2, section[1] / 2
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    phase = np.zeros((n, n, n), dtype=np.uint8)
    axis1 = np.zeros((n, n, n, 3))
    axis2 = np.zeros((n, n, n, 3))
    pitch = n / tows
    for j in range(tows):
        sgn = 1 if j % 2 == 0 else -1
        yc = (j + .5) * pitch
        zcw = zc + sgn * amplitude * np.sin(2 * np.pi * X / n)
        dy = (Y - yc + n / 2) % n - n / 2
        warp = (dy / a_w) ** 2 + ((Z - zcw) / a_t) ** 2 < 1
        slope = sgn * amplitude * 2 * np.pi / n * np.cos(2 * np.pi * X[warp] / n)
        norm = np.sqrt(1 + slope ** 2)
        phase[warp] = 1
        axis1[warp] = np.stack([1 / norm, np.zeros_like(norm), slope / norm], axis=1)
        axis2[warp] = [0., 1., 0.]
    for j in range(tows):
        sgn = 1 if j % 2 == 0 else -1
        xc = (j + .5) * pitch
        zcf = zc - sgn * amplitude * np.sin(2 * np.pi * Y / n)
        dx = (X - xc + n / 2) % n - n / 2
        weft = ((dx / a_w) ** 2 + ((Z - zcf) / a_t) ** 2 < 1) & (phase == 0)
        slope = -sgn * amplitude * 2 * np.pi / n * np.cos(2 * np.pi * Y[weft] / n)
        norm = np.sqrt(1 + slope ** 2)
        phase[weft] = 2
        axis1[weft] = np.stack([np.zeros_like(norm), 1 / norm, slope / norm], axis=1)
        axis2[weft] = [-1., 0., 0.]
    matrix_mask = phase == 0
    axis1[matrix_mask] = [1., 0., 0.]
    axis2[matrix_mask] = [0., 1., 0.]
    tow = dict(kind="orthotropic", frame="local", E1=230., E2=15., E3=15., nu12=.2, nu13=.2, nu23=.3,
               G12=27., G13=27., G23=15. / (2 * 1.3))
    materials = {0: dict(kind="isotropic", E=3.5, nu=.35), 1: dict(tow), 2: dict(tow)}
    return (phase, materials, np.ones(3), dict(axis1=axis1, axis2=axis2)), dict(
        tow_fraction=float((phase > 0).mean()))


def simp_density(n=96, sigma=4.0, seed=5, E_min=1e-3, E_max=70.0, nu=0.3, p=3.0):
    """SIMP density cell: rho in [0, 1] from seeded white noise N(0,1)
    convolved periodically with a Gaussian (sigma voxels), min-max scaled; the
    stiffness is E(rho) = E_min + rho^p (E_max - E_min), nu fixed.

    Returns (rho (n,n,n), unit-modulus isotropic 6x6 C1 with E = 1, E(rho),
    dE/drho); C_e = E(rho_e) C1.
    """
    from .materials import isotropic_voigt
    rng = np.random.default_rng(seed)
    noise = rng.normal(size=(n, n, n))
    k = np.fft.fftfreq(n)
    k2 = k[:, None, None] ** 2 + k[None, :, None] ** 2 + k[None, None, :] ** 2
    field = np.fft.ifftn(np.fft.fftn(noise) * np.exp(-2.0 * (np.pi * sigma) ** 2 * k2)).real
    rho = (field - field.min()) / (field.max() - field.min())
    E = E_min + rho ** p * (E_max - E_min)
    dE = p * rho ** (p - 1) * (E_max - E_min)
    return rho, isotropic_voigt(1.0, nu), E, dE


def random_inclusion_batch(count=64, seed=7, sizes=(32, 48, 64, 80, 96), fractions=(0.1, 0.4), radii=(2, 5),
                           contrasts=(10.0, 1000.0), nu=0.3, E_matrix=1.0):
    """Uneven batch of two-phase random-inclusion cells: per cell n ~ U(sizes),
    volume fraction ~ U[fractions], sphere radius ~ U{radii}, stiffness
    contrast E_inc/E_mat ~ log-uniform in `contrasts`, the same nu for both
    phases.  Returns a list of ((phase, materials, domain, None), meta)."""
    rng = np.random.default_rng(seed)
    cells = []
    for k in range(count):
        n = int(rng.choice(sizes))
        vf = float(rng.uniform(*fractions))
        radius = float(rng.integers(radii[0], radii[1] + 1))
        contrast = float(np.exp(rng.uniform(np.log(contrasts[0]), np.log(contrasts[1]))))
        (phase, mats, dom, _), meta = periodic_spheres(n, radius=radius, fraction=vf, seed=seed * 1000 + k,
                                                       E=(E_matrix, E_matrix * contrast), nu=(nu, nu))
        meta.update(n=n, radius=radius, contrast=contrast, target_fraction=vf)
        cells.append(((phase, mats, dom, None), meta))
    return cells
