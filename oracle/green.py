"""FI Green symbol and the partial Green-Jacobi preconditioner (test infrastructure only).

Restates ``solver/src/reference_medium/formula_adapter.py:32-68`` (tables,
v-tables), ``formula_kernels.cu:113-126`` (Sherman-Morrison inverse),
``solver/src/spectral/precision_kernels.cu:57-105`` (dense 6-coefficient
symbol with 1/N folded in, symmetric 3x3 multiply) and
``solver/experiments/adaptive_green_scaling/study.py:61-101`` (diagonal,
scaling, set_alpha). Pinned in ``tests/test_oracle_green.py`` against the
dense inverse of the impulse-response symbol of the assembled homogeneous K
(including Nyquist, odd and non-cubic grids), against the tensor-product
closed form, and against dense ``P s pinv(K0) s P``.
"""
import numpy as np
import scipy.fft

from . import hex8

# None: numpy.fft (the checker's default). An int: scipy.fft with that many workers (the same transforms,
# threaded), which the bench's CPU baseline sets while it times the oracle.
FFT_WORKERS = None


def angles(n, count):
    """theta_k = pi k / n, k < count (``formula_adapter.py:32-37``)."""
    return np.pi * np.arange(count, dtype=np.float64) / n


def symbol(grid, voxel, lam, mu):
    """Dense 6-coefficient Green symbol (6, nz, ny, nx//2+1), 1/N folded in.

    Order [G00, G11, G22, G01, G02, G12]; G = (D - beta (Dv)(Dv)^T)/N with
    D_ii = 1/(mu t sum_j (d_j + v_j^2) + (lam+mu) t d_i), t = V prod m_j,
    m = c^2 + s^2/3, v = 2sc/(h m), d = 4 s^4/(3 h^2 m^2),
    beta = (lam+mu) t / (1 + (lam+mu) t sum v_i^2 D_ii); zero at k = 0.
    """
    nz, ny, nx = grid
    hx, hy, hz = voxel
    axes = []
    for n, count, h in ((nx, nx // 2 + 1, hx), (ny, ny, hy), (nz, nz, hz)):
        th = angles(n, count)
        s, c = np.sin(th), np.cos(th)
        m = c * c + s * s / 3
        v = 2 * s * c / (h * m)
        d = 4 * s ** 4 / (3 * h * h * m * m)
        axes.append((m, v, d))
    (mx, vx, dx), (my, vy, dy), (mz, vz, dz) = axes
    shape = (nz, ny, nx // 2 + 1)
    M = [mx[None, None, :], my[None, :, None], mz[:, None, None]]
    V = [np.broadcast_to(vx[None, None, :], shape), np.broadcast_to(vy[None, :, None], shape),
         np.broadcast_to(vz[:, None, None], shape)]
    Dd = [dx[None, None, :], dy[None, :, None], dz[:, None, None]]
    t = hx * hy * hz * M[0] * M[1] * M[2]
    trace = sum(Dd[i] + V[i] * V[i] for i in range(3)) * t
    b = (lam + mu) * t
    zero = np.zeros(shape, bool)
    zero[0, 0, 0] = True
    inv = []
    den = np.ones(shape)
    for i in range(3):
        with np.errstate(divide='ignore'):
            q = 1.0 / (mu * trace + (lam + mu) * t * Dd[i])
        q = np.where(zero, 0.0, np.broadcast_to(q, shape))
        inv.append(q)
        den = den + b * V[i] * V[i] * q
    beta = np.where(zero, 0.0, np.broadcast_to(b / den, shape))
    invN = 1.0 / (nx * ny * nz)
    w = [inv[i] * V[i] for i in range(3)]
    out = np.empty((6,) + shape)
    out[0] = (inv[0] - beta * w[0] * w[0]) * invN
    out[1] = (inv[1] - beta * w[1] * w[1]) * invN
    out[2] = (inv[2] - beta * w[2] * w[2]) * invN
    out[3] = -beta * w[0] * w[1] * invN
    out[4] = -beta * w[0] * w[2] * invN
    out[5] = -beta * w[1] * w[2] * invN
    return out


def green_apply(v, G6, grid):
    """z = irfftn(G rfftn(v), norm='forward') per load (``contest_adapter.py:204-224``).

    ``v`` is AoS (L, 3N); the spectral multiply is the symmetric 3x3
    ``cached_inplace`` dense mode (``precision_kernels.cu:91-97``).
    """
    v = np.atleast_2d(np.asarray(v, float))
    L = v.shape[0]
    nz, ny, nx = grid
    field = v.reshape(L, nz, ny, nx, 3).transpose(0, 4, 1, 2, 3)
    if FFT_WORKERS:
        X = scipy.fft.rfftn(field, axes=(-3, -2, -1), workers=FFT_WORKERS)
    else:
        X = np.fft.rfftn(field, axes=(-3, -2, -1))
    x, y, z = X[:, 0], X[:, 1], X[:, 2]
    g0, g1, g2, g3, g4, g5 = G6
    Y = np.stack((g0 * x + g3 * y + g4 * z, g3 * x + g1 * y + g5 * z, g4 * x + g5 * y + g2 * z), axis=1)
    if FFT_WORKERS:
        out = scipy.fft.irfftn(Y, s=(nz, ny, nx), axes=(-3, -2, -1), norm='forward', workers=FFT_WORKERS)
    else:
        out = np.fft.irfftn(Y, s=(nz, ny, nx), axes=(-3, -2, -1), norm='forward')
    return out.transpose(0, 2, 3, 4, 1).reshape(L, -1)


def scale(a, alpha, C, voxel=None):
    """s = exp(-alpha/2 (log d - mean log d)) over all 3N entries (``resident.py:170-174``)."""
    q = np.log(hex8.diagonal(a, C, voxel))
    q = q - q.mean()
    return np.exp(-0.5 * alpha * q)


def precond(r, G6, grid, s=None):
    """M_alpha^-1 r = P(s * G(s * P r)); s=None is alpha = 0 (``adaptive_green_scaling/study.py:91-101``)."""
    r = np.atleast_2d(np.asarray(r, float))
    if s is None:
        return hex8.project(green_apply(hex8.project(r), G6, grid))
    return hex8.project(s * green_apply(hex8.project(r) * s, G6, grid))


def dense_symbol_inverse_apply(v, grid, voxel, lam, mu):
    """Independent Green apply: dense 3x3 inverse of the impulse-response symbol of K0.

    Restates ``solver/experiments/ladecky_comparison/study.py:210-224``: the
    homogeneous operator's response to unit impulses gives a 3x3 kernel whose
    full FFT is inverted per frequency (k = 0 set to zero).
    """
    v = np.atleast_2d(np.asarray(v, float))
    nz, ny, nx = grid
    n = nx * ny * nz
    C = hex8.isotropic(lam, mu)
    ones = np.ones(grid)
    kernel = np.empty((3, 3) + tuple(grid))
    for j in range(3):
        impulse = np.zeros((1, 3 * n))
        impulse[0, j] = 1.0
        force = hex8.apply(ones, impulse, C, voxel)
        kernel[:, j] = force.reshape(*grid, 3).transpose(3, 0, 1, 2)
    sym = np.fft.fftn(kernel, axes=(-3, -2, -1)).transpose(2, 3, 4, 0, 1)
    inv = np.zeros_like(sym)
    mask = np.ones(grid, bool)
    mask[0, 0, 0] = False
    inv[mask] = np.linalg.inv(sym[mask])
    out = []
    for row in v:
        rf = row.reshape(*grid, 3)
        rh = np.fft.fftn(rf, axes=(0, 1, 2))
        out.append(np.fft.ifftn(np.einsum('...ij,...j->...i', inv, rh), axes=(0, 1, 2)).real.reshape(-1))
    return np.array(out)
