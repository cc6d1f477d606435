"""General-C21 periodic HEX8 operator, restated in numpy FP64 (test infrastructure only).

Element e has the stiffness of row ``index[e]`` of a (P, 21) pool (or row e
when ``index`` is None), 21 upper-triangle Voigt entries row-major
(``solver/src/materials/material_field.py:14``). Same grid, node and Voigt
conventions as ``oracle.hex8``.

* ``apply``: q_l = sum_e P_e^T Ke(C_e) P_e u_l, Ke from 2x2x2 Gauss
  (``solver/src/materials/hybrid_kernels.cu:25-48``);
* ``rhs``: f = -sum_e Ke(C_e) chi(macro) (``hybrid_kernels.cu:49-68``);
* ``avg_stress``: <C_e (macro + Bbar u_e)> (``hybrid_kernels.cu:69-106``);
* ``dense_matrix``: explicit element loops, tiny grids;
* ``laminate``: closed-form effective stiffness of layers normal to z.

Pins (tests/test_oracle_hybrid.py): rows a_e C0 reproduce the isotropic
oracle; the explicit dense assembly; rigid translations in the kernel;
``laminate`` against the solved C^H of z-layered media (the Q1 space holds
the exact piecewise-affine solution); the packing order against a literal.
"""
import numpy as np

from . import hex8 as _h

_ROWS = np.array([i for i in range(6) for j in range(i, 6)])
_COLS = np.array([j for i in range(6) for j in range(i, 6)])


def unpack(p):
    """(..., 21) -> symmetric (..., 6, 6)."""
    p = np.asarray(p, float)
    C = np.zeros(p.shape[:-1] + (6, 6))
    C[..., _ROWS, _COLS] = p
    C[..., _COLS, _ROWS] = p
    return C


def pack(C):
    """(..., 6, 6) -> (..., 21)."""
    return np.asarray(C, float)[..., _ROWS, _COLS]


def element_C(pool, index, grid):
    """(nz, ny, nx, 6, 6) per-element stiffness."""
    n = int(np.prod(grid))
    rows = np.arange(n) if index is None else np.asarray(index).reshape(-1)
    return unpack(np.asarray(pool, float).reshape(-1, 21)[rows]).reshape(*grid, 6, 6)


def element_matrices(Ce, voxel):
    """Ke(C_e) = sum_q B_q^T C_e B_q V/8 for every element: (..., 24, 24)."""
    B = _h.gauss_b_matrices(voxel)
    return np.einsum('qsi,...st,qtj->...ij', B, Ce, B) * float(np.prod(voxel)) / 8


def apply(pool, index, u, grid, voxel=None, active=None):
    """(L, 3N) -> (L, 3N); loads with active[l] False give zeros."""
    grid = tuple(grid)
    voxel = _h.voxel_of(grid) if voxel is None else voxel
    u = np.atleast_2d(np.asarray(u, float))
    L = u.shape[0]
    Ce = element_C(pool, index, grid).reshape(-1, 6, 6)
    B = _h.gauss_b_matrices(voxel)
    w = float(np.prod(voxel)) / 8
    local = _h._gather(u.reshape(L, *grid, 3), grid).reshape(L, -1, 24)
    forces = np.zeros_like(local)
    for q in range(8):                       # Ke u = sum_q B_q^T C (B_q u) V/8, without forming Ke
        eps = local @ B[q].T
        sig = np.einsum('nst,lnt->lns', Ce, eps)
        forces += (sig @ B[q]) * w
    out = _h._scatter(forces.reshape(L, *grid, 24)).reshape(L, -1)
    if active is not None:
        out[~np.asarray(active, bool)] = 0.0
    return out


def rhs(pool, index, macro, grid, voxel=None):
    """(1, 3N): the element forces of the affine field, negated and summed at nodes."""
    grid = tuple(grid)
    voxel = _h.voxel_of(grid) if voxel is None else voxel
    Ke = element_matrices(element_C(pool, index, grid), voxel)
    chi = _h.affine(macro, voxel)
    forces = -(Ke @ chi)
    return _h._scatter(forces[None]).reshape(1, -1)


def avg_stress(pool, index, u, macro, grid, voxel=None):
    """(6,) mean of C_e (macro + Bbar u_e)."""
    grid = tuple(grid)
    voxel = _h.voxel_of(grid) if voxel is None else voxel
    Bbar = _h.centroid_b(voxel)
    local = _h._gather(np.asarray(u, float).reshape(1, *grid, 3), grid)[0]
    eps = local @ Bbar.T + np.asarray(macro, float)
    sig = np.einsum('...st,...t->...s', element_C(pool, index, grid), eps)
    return sig.reshape(-1, 6).mean(0)


def dense_matrix(pool, index, grid, voxel=None):
    """Assembled K (3N, 3N) by explicit loops over elements and their nodes."""
    nz, ny, nx = grid
    voxel = _h.voxel_of(grid) if voxel is None else voxel
    Ce = element_C(pool, index, grid)
    n = nx * ny * nz
    K = np.zeros((3 * n, 3 * n))
    B = _h.gauss_b_matrices(voxel)
    V = float(np.prod(voxel))
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                Ke = sum(B[q].T @ Ce[z, y, x] @ B[q] for q in range(8)) * V / 8
                ids = []
                for ox, oy, oz in _h.OFFSETS:
                    node = (((z + oz) % nz) * ny + (y + oy) % ny) * nx + (x + ox) % nx
                    ids += [3 * node, 3 * node + 1, 3 * node + 2]
                K[np.ix_(ids, ids)] += Ke
    return K


def effective_tensor_dense(pool, index, grid, voxel=None):
    """C^H (6, 6) by the average-stress route with a dense solve (tiny grids)."""
    grid = tuple(grid)
    K = dense_matrix(pool, index, grid, voxel)
    out = np.zeros((6, 6))
    for J in range(6):
        E = np.zeros(6)
        E[J] = 1.0
        f = rhs(pool, index, E, grid, voxel)[0]
        u = np.linalg.lstsq(K, f, rcond=None)[0]
        u = _h.project(u[None])[0]
        out[:, J] = avg_stress(pool, index, u, E, grid, voxel)
    return 0.5 * (out + out.T)


def laminate(Cs, fractions):
    """Effective stiffness of perfectly bonded layers normal to x3.

    In-plane strains (Voigt 0, 1, 5) and out-of-plane stresses (2, 3, 4) are
    the same in every layer; eliminating the layer's out-of-plane strains gives
    C^H_bb = <C_bb^-1>^-1, C^H_ba = C^H_bb <C_bb^-1 C_ba>,
    C^H_aa = <C_aa - C_ab C_bb^-1 C_ba> + <C_ab C_bb^-1> C^H_ba.
    """
    a, b = [0, 1, 5], [2, 3, 4]
    w = np.asarray(fractions, float) / np.sum(fractions)
    Cs = np.asarray(Cs, float)
    inv_bb = np.array([np.linalg.inv(C[np.ix_(b, b)]) for C in Cs])
    Hbb = np.linalg.inv(np.einsum('k,kij->ij', w, inv_bb))
    mean_ib_ba = np.einsum('k,kij,kjl->il', w, inv_bb, Cs[:, b][:, :, a])
    Hba = Hbb @ mean_ib_ba
    mean_schur = np.einsum('k,kij->ij', w, Cs[:, a][:, :, a] - np.einsum('kij,kjl,klm->kim', Cs[:, a][:, :, b],
                                                                           inv_bb, Cs[:, b][:, :, a]))
    mean_ab_ib = np.einsum('k,kij,kjl->il', w, Cs[:, a][:, :, b], inv_bb)
    Haa = mean_schur + mean_ab_ib @ Hba
    H = np.zeros((6, 6))
    H[np.ix_(a, a)] = Haa
    H[np.ix_(b, a)] = Hba
    H[np.ix_(a, b)] = Hba.T
    H[np.ix_(b, b)] = Hbb
    return H
