"""The oracle still reproduces the committed golden fixtures (CPU).

tests/golden/make_golden.py wrote them; this pins the oracle against drift so
that the GPU tests that read the fixtures keep checking the same numbers.
"""
import os

import numpy as np
import pytest

from oracle import design, green, hex8, pcg

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')
LAM, MU = 15 / 26, 5 / 13
C0 = hex8.isotropic(LAM, MU)


def load(name):
    with np.load(os.path.join(HERE, name + '.npz'), allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def test_operator_fixture():
    g = load('operator_4x5x6')
    a, u = g['a'], g['u']
    voxel = hex8.voxel_of(a.shape)
    np.testing.assert_allclose(hex8.apply(a, u, C0, voxel), g['Ku'], rtol=0, atol=1e-13 * np.abs(g['Ku']).max())
    rhs = np.concatenate([hex8.rhs(a, m, C0, voxel) for m in np.eye(6)])
    np.testing.assert_allclose(rhs, g['rhs'], rtol=0, atol=1e-14 * np.abs(g['rhs']).max())
    np.testing.assert_allclose(hex8.diagonal(a, C0, voxel), g['diag'], rtol=1e-14)


def test_green_fixture():
    g = load('green_3x4x5')
    grid = g['a'].shape
    G6 = green.symbol(grid, hex8.voxel_of(grid), LAM, MU)
    np.testing.assert_allclose(G6, g['symbol'], rtol=0, atol=1e-14 * np.abs(g['symbol']).max())
    z = green.precond(g['r'], G6, grid, g['s'])
    np.testing.assert_allclose(z, g['z_half'], rtol=0, atol=1e-13 * np.abs(g['z_half']).max())


def test_solve_fixture():
    g = load('solve_16_alpha05')
    rho = design.filtered_noise(16, 20261104, 2)
    np.testing.assert_array_equal(rho, g['rho'])
    a = design.simp(rho, 1e-4)
    ref = pcg.ResidentPCG(a, C0, np.eye(6), batch=1, s=green.scale(a, .5, C0))
    infos, _ = ref.solve(rtol=1e-8)
    assert [i['iterations'] for i in infos] == g['iterations'].tolist()
    CH = design.effective_tensor(a, ref.solutions, C0)
    np.testing.assert_allclose(CH, g['CH'], rtol=0, atol=1e-13 * np.abs(g['CH']).max())


def test_oc_fixture():
    g = load('oc_cone_8')
    oc = design.PaperOC(8, 0.35, emin=1e-6, move=0.2)
    ph, dp, _ = oc.physical(g['x'])
    new, count, _, _ = oc.update(g['x'], ph, dp, g['energy'])
    assert count == int(g['count'])
    np.testing.assert_array_equal(new, g['new'])


@pytest.mark.parametrize('name', ['operator_4x5x6', 'green_3x4x5', 'solve_16_alpha05', 'oc_cone_8'])
def test_fixtures_hold_plain_arrays(name):
    g = load(name)
    assert all(v.dtype.kind in 'fiu' for v in g.values())


def test_solve64_fixture_is_consistent():
    """The 64^3 oracle solves (minutes of CPU to rerun): certified, C^H symmetric, SPD and the same at every
    alpha (the preconditioner changes the iterates, not the solution)."""
    g = load('solve_64_alpha')
    H = g['CH_a05']
    for tag in ('a0', 'a05', 'a1'):
        assert np.all(g['relative_residual_' + tag] <= 1e-8)
        assert g['iterations_' + tag].shape == (6,)
        np.testing.assert_allclose(g['CH_' + tag], H, rtol=0, atol=1e-9 * np.abs(H).max())
    assert np.abs(H - H.T).max() <= 1e-12 * np.abs(H).max()
    assert np.linalg.eigvalsh(H).min() > 0
    # Green (alpha 0) needs several times the iterations of the scaled preconditioners at this contrast
    assert g['iterations_a0'].min() > 4 * g['iterations_a05'].max()
