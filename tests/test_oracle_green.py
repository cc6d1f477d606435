"""Pins the oracle's Green symbol and M_alpha against dense inverses (CPU)."""
import numpy as np
import pytest

from oracle import green, hex8

LAM, MU = 15 / 26, 5 / 13


def tensor_product_symbol(grid, voxel, lam, mu):
    """Closed-form FE symbol of K0 on the full frequency grid (restated from
    ``solver/experiments/solver_comparison/fft_precision.py:61-76``)."""
    nz, ny, nx = grid
    axes = []
    for n, h in ((nx, voxel[0]), (ny, voxel[1]), (nz, voxel[2])):
        th = 2 * np.pi * np.arange(n) / n
        mass = (2 + np.cos(th)) / 3
        axes.append((mass, (2 - 2 * np.cos(th)) / (h * h * mass), np.sin(th) / (h * mass)))
    out = np.empty((nz, ny, nx, 3, 3))
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                ks = (x, y, z)
                values = np.array([axes[j][1][k] for j, k in enumerate(ks)])
                v = np.array([axes[j][2][k] for j, k in enumerate(ks)])
                w = np.prod(voxel) * np.prod([axes[j][0][k] for j, k in enumerate(ks)])
                m = (lam + mu) * np.outer(v, v)
                np.fill_diagonal(m, mu * values.sum() + (lam + mu) * values)
                out[z, y, x] = w * m
    return out


@pytest.mark.parametrize('grid,voxel', [((3, 4, 5), (0.7, 1.1, 1.4)), ((4, 4, 4), None), ((6, 4, 8), None),
                                        ((5, 5, 5), None)])
def test_symbol_inverts_impulse_response_of_k0(grid, voxel):
    voxel = hex8.voxel_of(grid) if voxel is None else voxel
    G6 = green.symbol(grid, voxel, LAM, MU)
    rng = np.random.default_rng(5025)
    v = hex8.project(rng.normal(size=(2, 3 * int(np.prod(grid)))))
    z = green.green_apply(v, G6, grid)
    ref = green.dense_symbol_inverse_apply(v, grid, voxel, LAM, MU)
    assert np.linalg.norm(z - ref) / np.linalg.norm(ref) < 2e-12


@pytest.mark.parametrize('grid', [(4, 4, 4), (6, 4, 8)])
def test_symbol_times_tensor_product_symbol_is_identity(grid):
    voxel = hex8.voxel_of(grid)
    nz, ny, nx = grid
    A = tensor_product_symbol(grid, voxel, LAM, MU)[:, :, :nx // 2 + 1]
    G6 = green.symbol(grid, voxel, LAM, MU) * (nx * ny * nz)
    G = np.empty(A.shape)
    idx = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
    for k, (i, j) in enumerate(idx):
        G[..., i, j] = G[..., j, i] = G6[k]
    prod = np.einsum('...ij,...jk->...ik', G, A)
    eye = np.broadcast_to(np.eye(3), prod.shape).copy()
    eye[0, 0, 0] = 0
    assert np.abs(prod - eye).max() < 1e-12


def cosine_density(n, chi=100.0):
    """Planar cosine field of ``solver/experiments/smooth_contrast/study.py:44-53`` ('cosine', p = n)."""
    x = np.arange(n) / n
    xx, yy = np.meshgrid(x, x, indexing='xy')
    plane = .5 + .25 * (np.cos(2 * np.pi * (xx - yy)) + np.cos(2 * np.pi * (yy + xx))) + 1 / chi
    return np.broadcast_to(plane, (n, n, n)).copy()


def test_precond_matches_dense_pinv_form():
    """Dense M_alpha = P s pinv(K0) s P (``adaptive_green_scaling/study.py:205-235``)."""
    n = 4
    rho = cosine_density(n)
    C0 = hex8.isotropic(LAM, MU)
    N = rho.size
    K = hex8.dense_matrix(rho, C0)
    K0 = hex8.dense_matrix(np.ones_like(rho), C0)
    G = np.linalg.pinv(K0, rcond=1e-13)
    T = np.tile(np.eye(3), (N, 1))
    P = np.eye(3 * N) - T @ T.T / N
    Q = np.linalg.eigh(P)[1][:, 3:]
    ld = np.log(np.diag(K))
    ld -= ld.mean()
    G6 = green.symbol(rho.shape, hex8.voxel_of(rho.shape), LAM, MU)
    rng = np.random.default_rng(20261028)
    v = P @ rng.normal(size=3 * N)
    for alpha in (0., .25, .5, .75, 1.):
        s = np.exp(-.5 * alpha * ld)
        Bd = P @ (s[:, None] * G * s[None, :]) @ P
        sv = green.scale(rho, alpha, C0)
        assert np.abs(sv[0] - s).max() < 1e-13
        out = green.precond(v[None], G6, rho.shape, None if alpha == 0 else sv)[0]
        assert np.linalg.norm(out - Bd @ v) / np.linalg.norm(Bd @ v) < 1e-11
        assert np.linalg.norm(Bd - Bd.T) / np.linalg.norm(Bd) < 1e-11
        assert np.linalg.eigvalsh(Q.T @ Bd @ Q).min() > 0
