"""Pins the oracle's HEX8 operator against closed forms and brute force (CPU)."""
import numpy as np
import pytest

from oracle import hex8

LAM, MU = 15 / 26, 5 / 13
C0 = hex8.isotropic(LAM, MU)


def tensor_product_ke(voxel, lam, mu):
    """Exact Q1 stiffness from 1-D stiffness/mass/derivative factors (independent of Gauss B)."""
    hs = voxel
    K1 = [np.array([[1., -1.], [-1., 1.]]) / h for h in hs]
    M1 = [h * np.array([[1 / 3, 1 / 6], [1 / 6, 1 / 3]]) for h in hs]
    # D1[a, b] = int phi_a' phi_b
    D1 = [np.array([[-.5, -.5], [.5, .5]]) for _ in hs]

    def integral(i, j, a, b):
        """int d_i N_a d_j N_b over the element."""
        val = 1.0
        for d in range(3):
            ad, bd = hex8.OFFSETS[a][d], hex8.OFFSETS[b][d]
            if d == i and d == j:
                val *= K1[d][ad, bd]
            elif d == i:
                val *= D1[d][ad, bd]
            elif d == j:
                val *= D1[d][bd, ad]
            else:
                val *= M1[d][ad, bd]
        return val

    Ke = np.zeros((24, 24))
    for a in range(8):
        for b in range(8):
            for i in range(3):
                for j in range(3):
                    v = lam * integral(i, j, a, b) + mu * integral(j, i, a, b)
                    if i == j:
                        v += mu * sum(integral(k, k, a, b) for k in range(3))
                    Ke[3 * a + i, 3 * b + j] = v
    return Ke


@pytest.mark.parametrize('voxel', [(0.25, 0.25, 0.25), (0.7, 1.1, 1.4)])
def test_ke_matches_tensor_product_closed_form(voxel):
    Ke = hex8.element_stiffness(voxel, C0)
    ref = tensor_product_ke(voxel, LAM, MU)
    assert np.abs(Ke - ref).max() < 1e-13 * np.abs(ref).max()


def test_ke_symmetric_psd_with_rigid_null_space():
    voxel = (0.7, 1.1, 1.4)
    Ke = hex8.element_stiffness(voxel, C0)
    assert np.abs(Ke - Ke.T).max() < 1e-15 * np.abs(Ke).max()
    w = np.linalg.eigvalsh(Ke)
    assert w.min() > -1e-13 * w.max()
    assert int((w > 1e-10 * w.max()).sum()) == 18
    pos = hex8.OFFSETS * np.array(voxel)
    rigid = []
    for d in range(3):
        t = np.zeros((8, 3))
        t[:, d] = 1
        rigid.append(t.ravel())
    for (i, j) in [(0, 1), (0, 2), (1, 2)]:
        r = np.zeros((8, 3))
        r[:, i] = -pos[:, j]
        r[:, j] = pos[:, i]
        rigid.append(r.ravel())
    for v in rigid:
        assert np.abs(Ke @ v).max() < 1e-13 * np.abs(Ke).max()


@pytest.mark.parametrize('voxel', [(0.25, 0.25, 0.25), (0.7, 1.1, 1.4)])
def test_modal_ke_is_block_diagonal(voxel):
    """In the Hadamard basis of the eight corners (the modes the row-march kernels compute in), Ke couples a DOF
    (corner mode m, component c) only with DOFs of the same reflection class: the parity of m along each axis,
    flipped along the component's own axis. Eight 3x3 blocks, 45 nonzeros in all; the dense contraction on
    matrix cores multiplies all 576 entries."""
    Ke = hex8.element_stiffness(voxel, C0)
    h1 = np.array([[1.0, 1.0], [1.0, -1.0]])
    T = np.kron(np.kron(np.kron(h1, h1), h1), np.eye(3))   # corner o = x + 2 y + 4 z, DOF 3 o + c
    M = T @ Ke @ T.T
    nonzero = np.abs(M) > 1e-12 * np.abs(M).max()
    cls = [tuple(((o >> i) & 1) ^ int(c == i) for i in range(3)) for o in range(8) for c in range(3)]
    assert all(cls[i] == cls[j] for i, j in zip(*np.nonzero(nonzero)))
    assert int(nonzero.sum()) == 45


def test_affine_patch_equals_centroid_b():
    voxel = (0.7, 1.1, 1.4)
    Ke = hex8.element_stiffness(voxel, C0)
    Bbar = hex8.centroid_b(voxel)
    for m in np.eye(6):
        chi = hex8.affine(m, voxel)
        ref = Bbar.T @ (C0 @ m) * np.prod(voxel)
        assert np.abs(Ke @ chi - ref).max() < 1e-13 * np.abs(ref).max()
        # element energy of the affine field is eps^T C eps
        e = chi @ Ke @ chi / np.prod(voxel)
        assert abs(e - m @ C0 @ m) < 1e-13


@pytest.mark.parametrize('grid,voxel', [((3, 4, 5), (0.7, 1.1, 1.4)), ((4, 4, 4), None), ((2, 3, 2), None)])
def test_apply_matches_dense_assembly(grid, voxel):
    rng = np.random.default_rng(731)
    a = np.exp(rng.uniform(-4, 0, grid))
    K = hex8.dense_matrix(a, C0, voxel)
    u = rng.normal(size=(3, 3 * a.size))
    ref = u @ K.T
    out = hex8.apply(a, u, C0, voxel)
    assert np.abs(out - ref).max() < 1e-14 * np.abs(ref).max()
    assert np.abs(K - K.T).max() < 1e-15 * np.abs(K).max()
    d = hex8.diagonal(a, C0, voxel)[0]
    assert np.abs(d - np.diag(K)).max() < 1e-14 * d.max()
    masked = hex8.apply(a, u, C0, voxel, active=[True, False, True])
    assert (masked[1] == 0).all() and np.array_equal(masked[0], out[0])


def test_translation_invariance_and_rhs():
    rng = np.random.default_rng(5)
    grid = (4, 4, 4)
    a = np.exp(rng.uniform(-4, 0, grid))
    u = rng.normal(size=(1, 3 * a.size))
    shifted = (u.reshape(-1, 3) + np.array([1., -2., .5])).reshape(1, -1)
    base = hex8.apply(a, u, C0)
    assert np.abs(hex8.apply(a, shifted, C0) - base).max() < 1e-13 * np.abs(base).max()
    # rhs = -sum_e a_e Ke chi_e, assembled densely from the global affine field restricted per element
    voxel = hex8.voxel_of(grid)
    K = hex8.dense_matrix(a, C0)
    Ke = hex8.element_stiffness(voxel, C0)
    for m in np.eye(6):
        f = hex8.rhs(a, m, C0)[0]
        ref = np.zeros(3 * a.size)
        chi = hex8.affine(m, voxel)
        nz, ny, nx = grid
        for z in range(nz):
            for y in range(ny):
                for x in range(nx):
                    ids = []
                    for ox, oy, oz in hex8.OFFSETS:
                        node = (((z + oz) % nz) * ny + (y + oy) % ny) * nx + (x + ox) % nx
                        ids += [3 * node, 3 * node + 1, 3 * node + 2]
                    ref[ids] -= a[z, y, x] * (Ke @ chi)
        assert np.abs(f - ref).max() < 1e-14 * np.abs(ref).max()
        assert np.abs(f.reshape(-1, 3).sum(0)).max() < 1e-14 * np.abs(f).sum()
    homogeneous = hex8.rhs(np.full(grid, .13), np.ones(6), C0)
    assert np.abs(homogeneous).max() < 1e-15
    del K


def test_energies_and_stress_of_affine_field():
    grid = (3, 4, 5)
    a = np.full(grid, 1.0)
    u = np.zeros((6, 3 * a.size))
    e = hex8.energies(u, grid, np.eye(6), C0)
    for l in range(6):
        assert np.allclose(e[l], C0[l, l], rtol=1e-13)
    s = hex8.avg_stress(a, u[0], np.eye(6)[3], C0)
    assert np.allclose(s, C0[3], atol=1e-15)
