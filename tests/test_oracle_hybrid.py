"""Pins of the general-C21 oracle and of the host-side material builders (CPU only)."""
import numpy as np
import pytest
import torch

from oracle import hex8, hybrid
from pgj import materials as mat


def spd(rng, scale=1.0):
    A = rng.standard_normal((6, 6))
    return scale * (A @ A.T + 6 * np.eye(6))


def random_rotation(rng):
    q, r = np.linalg.qr(rng.standard_normal((3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def test_pack_order_literal():
    C = np.array([[10 * min(i, j) + max(i, j) for j in range(6)] for i in range(6)], float)
    want = [0, 1, 2, 3, 4, 5, 11, 12, 13, 14, 15, 22, 23, 24, 25, 33, 34, 35, 44, 45, 55]
    assert hybrid.pack(C).tolist() == want
    assert mat.pack_c21(C).tolist() == want
    assert np.array_equal(mat.unpack_c21(want), C)
    assert mat.PAIRS[:7] == ((0, 0), (0, 1), (0, 2), (0, 3), (0, 4), (0, 5), (1, 1))


def test_isotropic_rows_reproduce_scalar_oracle():
    rng = np.random.default_rng(1)
    grid = (3, 4, 5)
    a = rng.uniform(0.1, 1.0, grid)
    C0 = hex8.isotropic(15 / 26, 5 / 13)
    pool = a.reshape(-1, 1) * hybrid.pack(C0)[None]
    u = rng.standard_normal((2, 3 * a.size))
    np.testing.assert_allclose(hybrid.apply(pool, None, u, grid), hex8.apply(a, u, C0), rtol=0, atol=1e-13)
    m = rng.standard_normal(6)
    np.testing.assert_allclose(hybrid.rhs(pool, None, m, grid), hex8.rhs(a, m, C0), rtol=0, atol=1e-13)
    np.testing.assert_allclose(hybrid.avg_stress(pool, None, u[:1], m, grid), hex8.avg_stress(a, u[:1], m, C0),
                               rtol=1e-13, atol=1e-14)


def test_dense_assembly_symmetry_and_rigid_modes():
    rng = np.random.default_rng(2)
    grid = (3, 2, 3)
    pool = np.stack([hybrid.pack(spd(rng)) for _ in range(4)])
    index = rng.integers(0, 4, int(np.prod(grid)))
    K = hybrid.dense_matrix(pool, index, grid)
    u = rng.standard_normal((3, K.shape[0]))
    np.testing.assert_allclose(hybrid.apply(pool, index, u, grid), u @ K.T, rtol=0, atol=1e-11 * np.abs(K).max())
    np.testing.assert_allclose(K, K.T, atol=1e-12 * np.abs(K).max())
    assert np.linalg.eigvalsh(K).min() > -1e-9 * np.abs(K).max()
    for d in range(3):
        t = np.zeros((1, K.shape[0]))
        t[0, d::3] = 1.0
        assert np.abs(hybrid.apply(pool, index, t, grid)).max() < 1e-12 * np.abs(K).max()
    # homogeneous medium: an affine field is in equilibrium
    same = np.zeros_like(index)
    assert np.abs(hybrid.rhs(pool, same, rng.standard_normal(6), grid)).max() < 1e-12 * np.abs(pool).max()


def test_inactive_loads_are_zero():
    rng = np.random.default_rng(3)
    grid = (2, 3, 2)
    pool = hybrid.pack(spd(rng))[None]
    index = np.zeros(int(np.prod(grid)), int)
    u = rng.standard_normal((3, 3 * int(np.prod(grid))))
    out = hybrid.apply(pool, index, u, grid, active=[True, False, True])
    assert np.all(out[1] == 0.0)
    np.testing.assert_array_equal(out[[0, 2]], hybrid.apply(pool, index, u[[0, 2]], grid))


@pytest.mark.parametrize('layers', [(0, 0, 1, 2), (1, 0)])
def test_laminate_closed_form(layers):
    """z-layers: the Q1 space holds the exact piecewise-affine cell solution, so C^H is the laminate formula."""
    rng = np.random.default_rng(4)
    Cs = [spd(rng), spd(rng, 0.2), spd(rng, 3.0)]
    grid = (len(layers), 2, 2)
    pool = np.stack([hybrid.pack(C) for C in Cs])
    index = np.repeat(np.array(layers), 4)
    H = hybrid.effective_tensor_dense(pool, index, grid, voxel=(0.5, 0.5, 1.0 / len(layers)))
    frac = [layers.count(k) for k in range(3)]
    used = [k for k in range(3) if frac[k]]
    want = hybrid.laminate([Cs[k] for k in used], [frac[k] for k in used])
    np.testing.assert_allclose(H, want, rtol=1e-10, atol=1e-10 * np.abs(want).max())


def test_laminate_homogeneous_limit():
    rng = np.random.default_rng(5)
    C = spd(rng)
    np.testing.assert_allclose(hybrid.laminate([C, C], [1, 3]), C, rtol=1e-12)


# ---------------------------------------------------------------------------- materials
def test_isotropic_and_orthotropic_stiffness():
    np.testing.assert_allclose(mat.isotropic_voigt(1.0, 0.3), hex8.isotropic(15 / 26, 5 / 13), rtol=1e-14)
    E, nu = 2.5, 0.2
    G = E / (2 * (1 + nu))
    ortho = mat.orthotropic_voigt(E, E, E, nu, nu, nu, G, G, G)
    np.testing.assert_allclose(ortho, mat.isotropic_voigt(E, nu), rtol=1e-12, atol=1e-12)
    C = mat.material_stiffness(dict(kind='orthotropic', E1=10, E2=2, E3=1, nu12=0.3, nu13=0.2, nu23=0.25,
                                    G12=1.5, G13=0.8, G23=0.6))
    S = np.linalg.inv(C)
    assert np.isclose(1 / S[0, 0], 10) and np.isclose(-S[0, 1] / S[0, 0], 0.3) and np.isclose(1 / S[5, 5], 1.5)


def test_rotation_preserves_energy_and_isotropy():
    rng = np.random.default_rng(6)
    C = spd(rng)
    R = np.stack([random_rotation(rng) for _ in range(5)])
    rotated = mat.unpack_c21(mat.rotate_c21(C, R))
    T = mat.voigt_to_tensor(C)
    for k in range(5):
        e = rng.standard_normal((3, 3))
        e = e + e.T
        eg = R[k] @ e @ R[k].T
        Tg = mat.voigt_to_tensor(rotated[k])
        assert np.isclose(np.einsum('ij,ijkl,kl->', eg, Tg, eg), np.einsum('ij,ijkl,kl->', e, T, e), rtol=1e-12)
    np.testing.assert_allclose(mat.unpack_c21(mat.rotate_c21(C, np.eye(3)[None]))[0], C, rtol=1e-14)
    iso = mat.isotropic_voigt(3.0, 0.27)
    np.testing.assert_allclose(mat.unpack_c21(mat.rotate_c21(iso, R)), np.broadcast_to(iso, (5, 6, 6)),
                               atol=1e-13)
    # a quarter turn about z exchanges the 1 and 2 axes
    Rz = np.array([[[0.0, -1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0]]])
    Cz = mat.unpack_c21(mat.rotate_c21(C, Rz))[0]
    assert np.isclose(Cz[0, 0], C[1, 1]) and np.isclose(Cz[1, 1], C[0, 0]) and np.isclose(Cz[3, 3], C[4, 4])


def test_axis_frames_match_matrices():
    rng = np.random.default_rng(7)
    R = np.stack([random_rotation(rng) for _ in range(4)])
    axes = dict(axis1=R[:, :, 0] * 2.0, axis2=R[:, :, 1] + 0.3 * R[:, :, 0])
    np.testing.assert_allclose(mat.orientation_frames(axes, (4,), np.arange(4)), R, atol=1e-13)
    np.testing.assert_allclose(mat.orientation_frames(R, (4,), np.arange(4)), R)
    with pytest.raises(ValueError, match='right-handed'):
        mat.orientation_frames(-np.eye(3), (2,), np.arange(2))
    with pytest.raises(ValueError, match='collinear'):
        mat.orientation_frames(dict(axis1=[1, 0, 0], axis2=[2, 0, 0]), (2,), np.arange(2))


def _materials():
    return {0: dict(kind='isotropic', E=1.0, nu=0.3),
            3: dict(kind='orthotropic', frame='local', E1=10, E2=2, E3=1, nu12=0.3, nu13=0.2, nu23=0.25,
                    G12=1.5, G13=0.8, G23=0.6),
            7: dict(kind='anisotropic', C=mat.isotropic_voigt(4.0, 0.1) + 0.1 * np.ones((6, 6)))}


def test_hybrid_pool_layout_and_dense_equivalence():
    rng = np.random.default_rng(8)
    phase = rng.choice([0, 3, 7], size=(3, 4, 5))
    R = np.stack([random_rotation(rng) for _ in range(phase.size)]).reshape(*phase.shape, 3, 3)
    pool, index, meta = mat.build_hybrid_material(phase, _materials(), orientation=R, device='cpu')
    dense = mat.build_material_field(phase, _materials(), orientation=R, device='cpu', chunk_size=7)
    local = int((phase == 3).sum())
    assert pool.shape == (3 + local, 21) and index.shape == (1, phase.size) and index.dtype == torch.int32
    assert meta['pool_rows'] == 3 + local and meta['n_local_voxels'] == local
    assert meta['labels'] == [0, 3, 7] and meta['dense_bytes'] == phase.size * 21 * 8
    np.testing.assert_array_equal(pool[index[0].long()].numpy(), dense[0].numpy())
    flat = phase.reshape(-1)
    idx = index[0].numpy()
    assert np.all(idx[flat == 0] == 0) and np.all(idx[flat == 7] == 2)
    assert np.array_equal(idx[flat == 3], np.arange(3, 3 + local))
    C3 = mat.material_stiffness(_materials()[3])
    k = np.flatnonzero(flat == 3)[0]
    np.testing.assert_allclose(pool[3].numpy(), mat.rotate_c21(C3, R.reshape(-1, 3, 3)[k:k + 1])[0], rtol=1e-14)


def test_global_phases_need_no_orientation_and_errors():
    phase = np.zeros((2, 2, 3), np.int32)
    pool, index, meta = mat.build_hybrid_material(phase, _materials(), device='cpu')
    assert pool.shape == (3, 21) and meta['n_local_voxels'] == 0
    with pytest.raises(ValueError, match='require orientation'):
        mat.build_hybrid_material(phase + 3, _materials(), device='cpu')
    with pytest.raises(ValueError, match='Undefined material labels'):
        mat.build_hybrid_material(phase + 1, _materials(), device='cpu')
    with pytest.raises(ValueError, match='nonempty integer'):
        mat.build_hybrid_material(phase.astype(float), _materials(), device='cpu')
    with pytest.raises(ValueError, match='chunk_size'):
        mat.build_material_field(phase, _materials(), device='cpu', chunk_size=0)
    with pytest.raises(ValueError, match='positive definite'):
        mat.material_stiffness(dict(kind='anisotropic', C=-np.eye(6)))
    with pytest.raises(ValueError, match='symmetric'):
        mat.material_stiffness(dict(kind='anisotropic', C=np.eye(6) + np.triu(np.ones((6, 6)), 1)))
    with pytest.raises(ValueError, match='frame'):
        mat.material_stiffness(dict(kind='isotropic', E=1, nu=0.3, frame='other'))
    with pytest.raises(ValueError, match='integers'):
        mat.phase_table({'a': dict(kind='isotropic', E=1, nu=0.3)})
    batched = mat.build_material_field(np.zeros((2, 2, 2, 2), np.int32), _materials(), device='cpu')
    assert batched.shape == (2, 8, 21)


def test_hybrid_resident_recurrence_equals_scalar_for_isotropic_rows():
    """The resident recurrence over C21 rows a_e C0 with the Green medium C0 is the scalar recurrence: same
    iterations, solutions to rounding."""
    from oracle import pcg
    C0 = hex8.isotropic(15 / 26, 5 / 13)
    a = np.random.default_rng(1).uniform(0.1, 1, (6, 6, 6))
    pool = a.reshape(-1, 1) * hybrid.pack(C0)[None]
    loads = np.eye(6)[[0, 3]]
    ref = pcg.ResidentPCG(a, C0, loads, batch=2)
    ri, _ = ref.solve(1e-10)
    hyb = pcg.HybridResidentPCG(pool, None, a.shape, hex8.voxel_of(a.shape), loads, batch=2,
                                lam_mu=(15 / 26, 5 / 13))
    hi, _ = hyb.solve(1e-10)
    assert [i['iterations'] for i in ri] == [i['iterations'] for i in hi]
    assert np.abs(ref.solutions - hyb.solutions).max() < 1e-14 * np.abs(ref.solutions).max()
