"""Periodic Q1/HEX8 operator, restated in numpy FP64 (test infrastructure only).

Conventions (reference-pinned):
- grids are (nz, ny, nx); node = (z*ny + y)*nx + x; nodes coincide with
  elements on the periodic grid (``hybrid_adapter.py:54-56``);
- displacements are AoS ``(L, 3N)``, node-major (``hybrid_adapter.py:74-80``);
- element node ``a`` sits at offset (a&1, (a>>1)&1, (a>>2)&1) in (x, y, z)
  (``exp/asg_gpu_resident_20261103/scalar_fi.cu:25``, ``backend.py:155``);
- Voigt order 11,22,33,23,13,12 with engineering shear (``docs/DATA.md:21``).
"""
from functools import lru_cache
from itertools import product

import numpy as np

OFFSETS = np.array([[a & 1, (a >> 1) & 1, (a >> 2) & 1] for a in range(8)])


def isotropic(lam, mu):
    """6x6 isotropic Voigt stiffness (``reproduction.py:59-62``)."""
    c = np.zeros((6, 6))
    c[:3, :3] = lam
    c[np.arange(3), np.arange(3)] += 2 * mu
    c[np.arange(3, 6), np.arange(3, 6)] = mu
    return c


def gauss_b_matrices(voxel):
    """Strain-displacement matrices B_q (8, 6, 24) at the 2x2x2 Gauss points.

    Restates ``solver/experiments/ladecky_comparison/study.py:160-168`` and
    ``exp/asg_candidates_20261103/candidates.py:204-228``: dN_a/dx =
    s_x (1 + s_y eta)(1 + s_z zeta) / (4 h_x) with s = 2*offset - 1.
    Pinned by ``tests/test_oracle_hex8.py`` (rigid modes, affine patch test,
    solid C^H = C0, dense brute force).
    """
    hx, hy, hz = voxel
    signs = 2 * OFFSETS - 1
    out = []
    g = 1 / np.sqrt(3)
    for zeta, eta, xi in product((-g, g), repeat=3):
        B = np.zeros((6, 24))
        for a, (sx, sy, sz) in enumerate(signs):
            dx = sx * (1 + sy * eta) * (1 + sz * zeta) / (4 * hx)
            dy = sy * (1 + sx * xi) * (1 + sz * zeta) / (4 * hy)
            dz = sz * (1 + sx * xi) * (1 + sy * eta) / (4 * hz)
            B[0, 3 * a] = dx
            B[1, 3 * a + 1] = dy
            B[2, 3 * a + 2] = dz
            B[3, 3 * a + 1] = dz
            B[3, 3 * a + 2] = dy
            B[4, 3 * a] = dz
            B[4, 3 * a + 2] = dx
            B[5, 3 * a] = dy
            B[5, 3 * a + 1] = dx
        out.append(B)
    return np.array(out)


def centroid_b(voxel):
    """Mean strain-displacement matrix (6, 24): gradients +-1/(4h) (``hybrid_kernels.cu:58-66``)."""
    return gauss_b_matrices(voxel).mean(0)


@lru_cache(None)
def _ke_cached(voxel, c_key):
    C = np.array(c_key).reshape(6, 6)
    B = gauss_b_matrices(voxel)
    volume = float(np.prod(voxel))
    return np.einsum('qsi,st,qtj->ij', B, C, B) * volume / 8


def element_stiffness(voxel, C):
    """Exact 24x24 Ke = sum_q B_q^T C B_q V/8 (``candidates.py:219-228``)."""
    return _ke_cached(tuple(float(v) for v in voxel), tuple(np.asarray(C, float).ravel()))


def voxel_of(grid):
    """Unit cell voxel (hx, hy, hz) = 1/(nx, ny, nz) (``adaptive_green_scaling/study.py:63``)."""
    nz, ny, nx = grid
    return (1.0 / nx, 1.0 / ny, 1.0 / nz)


def _gather(field, grid):
    """(L, nz, ny, nx, 3) -> (L, nz, ny, nx, 24): element e gets node e + offset_a."""
    parts = []
    for ox, oy, oz in OFFSETS:
        parts.append(np.roll(field, shift=(-oz, -oy, -ox), axis=(1, 2, 3)))
    return np.concatenate(parts, axis=-1)


def _scatter(forces):
    """(L, nz, ny, nx, 24) element forces -> (L, nz, ny, nx, 3) nodal sums."""
    out = np.zeros(forces.shape[:-1] + (3,))
    for a, (ox, oy, oz) in enumerate(OFFSETS):
        out += np.roll(forces[..., 3 * a:3 * a + 3], shift=(oz, oy, ox), axis=(1, 2, 3))
    return out


def apply(a, u, C, voxel=None, active=None):
    """q_l = sum_e a_e P_e^T Ke P_e u_l for every load (``scalar_fi.cu:14-36``).

    ``a`` is the (nz, ny, nx) SIMP factor, ``u`` is (L, 3N). Loads with
    ``active[l] == False`` return exact zeros, as the reference kernel skips
    them. Pinned against the dense assembly in ``tests/test_oracle_hex8.py``.
    """
    a = np.asarray(a, float)
    grid = a.shape
    voxel = voxel_of(grid) if voxel is None else voxel
    u = np.atleast_2d(np.asarray(u, float))
    L = u.shape[0]
    Ke = element_stiffness(voxel, C)
    field = u.reshape(L, *grid, 3)
    local = _gather(field, grid)
    forces = (local @ Ke.T) * a[None, ..., None]
    out = _scatter(forces).reshape(L, -1)
    if active is not None:
        out[~np.asarray(active, bool)] = 0.0
    return out


def macro_tensor(macro):
    """Symmetric strain tensor from engineering Voigt strain (``backend.py:158-160``)."""
    m = np.asarray(macro, float)
    return np.array([[m[0], m[5] / 2, m[4] / 2],
                     [m[5] / 2, m[1], m[3] / 2],
                     [m[4] / 2, m[3] / 2, m[2]]])


def affine(macro, voxel):
    """Element-local affine displacement chi (24,): chi_a = E (offset_a * h) (``backend.py:153-162``)."""
    E = macro_tensor(macro)
    pos = OFFSETS * np.asarray(voxel)[None, :]
    return (pos @ E.T).reshape(-1)


def rhs(a, macro, C, voxel=None):
    """f = -sum_e a_e Ke chi(macro), shape (1, 3N) (``hybrid_kernels.cu:49-68``, ``verify.py:24-25``)."""
    a = np.asarray(a, float)
    grid = a.shape
    voxel = voxel_of(grid) if voxel is None else voxel
    Ke = element_stiffness(voxel, C)
    chi = affine(macro, voxel)
    local = -(Ke @ chi)
    forces = a[None, ..., None] * local[None, None, None, None, :]
    return _scatter(forces).reshape(1, -1)


def energies(u, a_shape, macros, C, voxel=None):
    """Element energies e_l,e = (u_e + chi_l)^T Ke (u_e + chi_l) / V (``energy.cu:7-24``).

    ``u`` is (L, 3N), ``macros`` (L, 6); returns (L, N). The factor a_e is
    not applied (the reference multiplies afterwards, ``resident.py:213``).
    """
    grid = tuple(a_shape)
    voxel = voxel_of(grid) if voxel is None else voxel
    u = np.atleast_2d(np.asarray(u, float))
    L = u.shape[0]
    Ke = element_stiffness(voxel, C)
    volume = float(np.prod(voxel))
    local = _gather(u.reshape(L, *grid, 3), grid)
    chi = np.stack([affine(m, voxel) for m in np.atleast_2d(macros)])
    v = local + chi[:, None, None, None, :]
    e = np.einsum('...i,ij,...j->...', v, Ke, v) / volume
    return e.reshape(L, -1)


def avg_stress(a, u, macro, C, voxel=None):
    """Mean stress <a C0 (eps + B_bar u_e)> over elements (``hybrid_kernels.cu:69-106``)."""
    a = np.asarray(a, float)
    grid = a.shape
    voxel = voxel_of(grid) if voxel is None else voxel
    Bbar = centroid_b(voxel)
    local = _gather(np.asarray(u, float).reshape(1, *grid, 3), grid)[0]
    eps = local @ Bbar.T + np.asarray(macro, float)
    sig = (eps @ np.asarray(C).T) * a[..., None]
    return sig.reshape(-1, 6).mean(0)


def element_diagonal(voxel, C):
    """diag(Ke) as (8, 3) in this module's node order (``smooth_contrast/study.py:56-67``)."""
    return np.diag(element_stiffness(voxel, C)).reshape(8, 3)


def diagonal(a, C, voxel=None):
    """Nodal diag(K(a)) as (1, 3N): d_i = sum_a a_{i - o_a} ed_a (``resident.py:170-172``)."""
    a = np.asarray(a, float)
    grid = a.shape
    voxel = voxel_of(grid) if voxel is None else voxel
    ed = element_diagonal(voxel, C)
    d = np.zeros((*grid, 3))
    for k, (ox, oy, oz) in enumerate(OFFSETS):
        d += np.roll(a, shift=(oz, oy, ox), axis=(0, 1, 2))[..., None] * ed[k]
    return d.reshape(1, -1)


def dense_matrix(a, C, voxel=None):
    """Assembled K (3N, 3N) by explicit element loops; tiny grids only (``verify.py:8-17``)."""
    a = np.asarray(a, float)
    nz, ny, nx = a.shape
    voxel = voxel_of(a.shape) if voxel is None else voxel
    Ke = element_stiffness(voxel, C)
    n = nx * ny * nz
    K = np.zeros((3 * n, 3 * n))
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                ids = []
                for ox, oy, oz in OFFSETS:
                    node = (((z + oz) % nz) * ny + (y + oy) % ny) * nx + (x + ox) % nx
                    ids += [3 * node, 3 * node + 1, 3 * node + 2]
                K[np.ix_(ids, ids)] += a[z, y, x] * Ke
    return K


def project(v):
    """Zero-mean projection per load and component (``resident.py:28-31``)."""
    v = np.atleast_2d(np.asarray(v, float))
    q = v.reshape(v.shape[0], -1, 3)
    return (q - q.mean(1, keepdims=True)).reshape(v.shape)
