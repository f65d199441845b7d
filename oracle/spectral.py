"""Oracle FI Fourier symbol of an isotropic reference medium and its Green
preconditioner, FP64 NumPy.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates
  * ``repro/bench/solver_comparison/fft_precision.py:62-92`` (dense 3x3
    tensor-product FE symbol, inverse per wavevector, k = 0 -> 0);
  * ``src/voxel_lmfem/_core/reference_medium/formula_kernels.cu:104-137`` and
    ``formula_adapter.py:32-68`` (diagonal + rank-1 Sherman-Morrison form from
    half-angle 1-D tables).

Symbol, per axis i with theta_i = 2 pi k_i / n_i:
    m_i = (2 + cos theta_i)/3,  D_i = (2 - 2 cos theta_i)/(h_i^2 m_i),
    v_i = sin theta_i/(h_i m_i),  d_i = D_i - v_i^2,
    K(k) = V prod_i m_i [diag(mu sum_j D_j + (lam+mu) d_i) + (lam+mu) v v^T].

C3D8R-HG control (``symbol_hg``, ``green_hg_dense``): with s_i = sin(theta_i/2),
c_i = cos(theta_i/2), w_x = (2/h_x) s_x c_y c_z (cyclic),
    K(k) = V (mu |w|^2 I + (lam+mu) w w^T) + s0 H(k) I,
    H = 8 (s_x^2 s_y^2 c_z^2 + s_y^2 s_z^2 c_x^2 + s_z^2 s_x^2 c_y^2 + s_x^2 s_y^2 s_z^2),
s0 the hourglass scale of the isotropic medium (femlib/c3d8r.py:48-59).

Pinning
  * ``symbol`` equals the periodic FI operator (oracle.fem, Gauss-B form) of a
    uniform isotropic cell applied to plane waves (independent derivation);
  * ``symbol_hg`` equals the assembled C3D8R-HG operator (oracle.fem dense
    element matrices) of a uniform isotropic cell applied to plane waves;
  * ``green_apply`` (Sherman-Morrison) equals the dense 3x3 inverse path;
  * K(k) G(k) = I on every k != 0.
"""
import numpy as np


def axis_tables(n, h, count=None):
    """Per-axis (m, D, v, d) at k = 0..count-1 (count defaults to n)."""
    count = n if count is None else count
    theta = 2 * np.pi * np.arange(count, dtype=np.float64) / n
    m = (2 + np.cos(theta)) / 3
    D = (2 - 2 * np.cos(theta)) / (h * h * m)
    v = np.sin(theta) / (h * m)
    # d = D - v^2 = 4 s^4 / (3 h^2 m^2) with s = sin(theta/2): no cancellation.
    s = np.sin(theta / 2)
    d = 4 * s ** 4 / (3 * h * h * m * m)
    return m, D, v, d


def symbol(shape, h, lam, mu):
    """Dense (Nz, Ny, Nx, 3, 3) symbol over the full spectrum."""
    nz, ny, nx = shape
    h = np.asarray(h, dtype=np.float64)
    V = float(np.prod(h))
    ax = [axis_tables(n, hh) for n, hh in zip((nx, ny, nz), h)]
    mx, Dx, vx, _ = ax[0]
    my, Dy, vy, _ = ax[1]
    mz, Dz, vz, _ = ax[2]
    m = mz[:, None, None] * my[None, :, None] * mx[None, None, :]
    D = [np.broadcast_to(Dx[None, None, :], (nz, ny, nx)),
         np.broadcast_to(Dy[None, :, None], (nz, ny, nx)),
         np.broadcast_to(Dz[:, None, None], (nz, ny, nx))]
    v = [np.broadcast_to(vx[None, None, :], (nz, ny, nx)),
         np.broadcast_to(vy[None, :, None], (nz, ny, nx)),
         np.broadcast_to(vz[:, None, None], (nz, ny, nx))]
    total = D[0] + D[1] + D[2]
    out = np.empty((nz, ny, nx, 3, 3))
    for i in range(3):
        for j in range(3):
            out[..., i, j] = (lam + mu) * v[i] * v[j]
        out[..., i, i] = mu * total + (lam + mu) * D[i]
    return out * (V * m)[..., None, None]


def dense_inverse(shape, h, lam, mu):
    """fft_precision.py:80-82: np.linalg.inv per wavevector, k=0 -> 0."""
    S = symbol(shape, h, lam, mu)
    inv = np.zeros_like(S)
    flat = inv.reshape(-1, 3, 3)
    flat[1:] = np.linalg.inv(S.reshape(-1, 3, 3)[1:])
    return inv


def green_dense(r, shape, h, lam, mu):
    """z = IFFT(K^-1 FFT(r)) with the dense inverse; r is (N, 3) node-major."""
    field = np.asarray(r, dtype=np.float64).reshape(*shape, 3)
    spectrum = np.fft.fftn(field, axes=(0, 1, 2))
    out = np.fft.ifftn(np.einsum("...ij,...j->...i", dense_inverse(shape, h, lam, mu), spectrum),
                       axes=(0, 1, 2)).real
    return out.reshape(-1, 3)


def green_coefficients(shape, h, lam, mu):
    """Diagonal inverses a_i and rank-1 coefficient beta on the half spectrum.

    formula_kernels.cu:113-126 form: K = diag(A) + b v v^T,
    K^-1 = diag(1/A) - beta (v/A)(v/A)^T, beta = b/(1 + b sum v_i^2/A_i).
    Returns (a (3,Nz,Ny,Nxh), beta (Nz,Ny,Nxh), v (3,Nz,Ny,Nxh)).
    """
    nz, ny, nx = shape
    nxh = nx // 2 + 1
    h = np.asarray(h, dtype=np.float64)
    V = float(np.prod(h))
    tx = axis_tables(nx, h[0], nxh)
    ty = axis_tables(ny, h[1])
    tz = axis_tables(nz, h[2])
    grid = lambda t, axis: (t[None, None, :] if axis == 0 else
                            t[None, :, None] if axis == 1 else t[:, None, None])
    m = [grid(tx[0], 0), grid(ty[0], 1), grid(tz[0], 2)]
    D = [grid(tx[1], 0), grid(ty[1], 1), grid(tz[1], 2)]
    v = [grid(tx[2], 0), grid(ty[2], 1), grid(tz[2], 2)]
    d = [grid(tx[3], 0), grid(ty[3], 1), grid(tz[3], 2)]
    shape_h = (nz, ny, nxh)
    t = V * m[0] * m[1] * m[2]
    total = D[0] + D[1] + D[2]
    A = np.stack([np.broadcast_to(t * (mu * total + (lam + mu) * d[i]), shape_h) for i in range(3)])
    vv = np.stack([np.broadcast_to(v[i], shape_h) for i in range(3)])
    b = np.broadcast_to((lam + mu) * t, shape_h)
    with np.errstate(divide="ignore", invalid="ignore"):
        a = np.where(A > 0, 1 / A, 0.0)
        den = 1 + b * (vv * vv * a).sum(0)
        beta = np.where(A[0] > 0, b / den, 0.0)
    a[:, 0, 0, 0] = 0
    beta = np.array(beta)
    beta[0, 0, 0] = 0
    return a, beta, vv


def green_apply(r, shape, h, lam, mu):
    """z = irfftn(G(k) rfftn(r)) with the Sherman-Morrison form; r (N,3)."""
    field = np.asarray(r, dtype=np.float64).reshape(*shape, 3)
    spec = np.fft.rfftn(np.moveaxis(field, -1, 0), axes=(1, 2, 3))
    a, beta, v = green_coefficients(shape, h, lam, mu)
    q = a * spec
    s = (v * q).sum(0)
    out = q - beta * (a * v) * s
    z = np.fft.irfftn(out, s=shape, axes=(1, 2, 3))
    return np.moveaxis(z, 0, -1).reshape(-1, 3)


def symbol_hg(shape, h, lam, mu, coef=0.1):
    """Dense (Nz, Ny, Nx, 3, 3) symbol of the C3D8R-HG element."""
    nz, ny, nx = shape
    h = np.asarray(h, dtype=np.float64)
    V = float(np.prod(h))
    hbar = h.sum() / 3
    s0 = coef * max((lam + 3 * mu) / 2, 1e-12) * V / (hbar * hbar)
    th = [2 * np.pi * np.arange(n) / n for n in (nx, ny, nz)]
    sx, sy, sz = (np.sin(t / 2) for t in th)
    cx, cy, cz = (np.cos(t / 2) for t in th)
    X = lambda a: a[None, None, :]
    Y = lambda a: a[None, :, None]
    Z = lambda a: a[:, None, None]
    w = [2 / h[0] * X(sx) * Y(cy) * Z(cz), 2 / h[1] * X(cx) * Y(sy) * Z(cz), 2 / h[2] * X(cx) * Y(cy) * Z(sz)]
    w = [np.broadcast_to(v, (nz, ny, nx)) for v in w]
    ww = w[0] ** 2 + w[1] ** 2 + w[2] ** 2
    Hk = 8 * (X(sx) ** 2 * Y(sy) ** 2 * Z(cz) ** 2 + Y(sy) ** 2 * Z(sz) ** 2 * X(cx) ** 2
              + Z(sz) ** 2 * X(sx) ** 2 * Y(cy) ** 2 + X(sx) ** 2 * Y(sy) ** 2 * Z(sz) ** 2)
    out = np.empty((nz, ny, nx, 3, 3))
    for i in range(3):
        for j in range(3):
            out[..., i, j] = V * (lam + mu) * w[i] * w[j]
        out[..., i, i] += V * mu * ww + s0 * Hk
    return out


def green_hg_dense(r, shape, h, lam, mu, coef=0.1):
    """z = IFFT(K_hg^-1 FFT(r)) with the dense per-wavevector inverse, k = 0 -> 0."""
    S = symbol_hg(shape, h, lam, mu, coef)
    inv = np.zeros_like(S)
    flat = inv.reshape(-1, 3, 3)
    flat[1:] = np.linalg.inv(S.reshape(-1, 3, 3)[1:])
    field = np.asarray(r, dtype=np.float64).reshape(*shape, 3)
    spectrum = np.fft.fftn(field, axes=(0, 1, 2))
    out = np.fft.ifftn(np.einsum("...ij,...j->...i", inv, spectrum), axes=(0, 1, 2)).real
    return out.reshape(-1, 3)
