"""Resident PCG over the general-C21 operator (HybridResident) vs the oracle recurrence and closed forms (GPU).

Tolerances:
- iteration counts within 1 of the oracle's recurrence over the same operator and Green reference medium
  (rounding of the partial sums, as in the scalar solver's tests);
- C^H 1e-10 relative against the oracle's solutions' mean stresses;
- every certified residual is rechecked on the CPU with the oracle's apply.
"""
import numpy as np
import pytest
import torch

import pgj
from pgj import materials as mat
from pgj.certify import CertificationError
from pgj.hybrid import HybridFIOperator, HybridResident, compute_effective_stiffness_resident, enable_fi_reference
from oracle import hex8, hybrid, pcg

pytestmark = pytest.mark.gpu

from test_gpu_hybrid import multiphase, random_rotation, rel, spd  # noqa: E402

CASES = [((8, 8, 8), None, 'parallel'), ((8, 8, 8), None, 'serial'), ((6, 8, 10), (0.7, 1.1, 1.4), 'parallel'),
         ((16, 16, 16), None, 'parallel')]


@pytest.mark.parametrize('grid,voxel,mode', CASES, ids=[f'{g}-{m}' for g, _, m in CASES])
def test_hybrid_resident_matches_oracle(grid, voxel, mode):
    print(f'hybrid resident {grid} {mode}', flush=True)
    rng, pool, index, vox = multiphase(grid, 11, voxel)
    op = enable_fi_reference(HybridFIOperator(pool, index, grid, vox))
    ref_C0 = np.asarray(op._formula_reference_C0)
    res = HybridResident(op, load_mode=mode).capture()
    infos, checks = res.solve(rtol=1e-10)
    pool_h, index_h = pool.cpu().numpy(), index.cpu().numpy()
    ref = pcg.HybridResidentPCG(pool_h, index_h, grid, vox, np.eye(6), batch=1 if mode == 'serial' else 6,
                                lam_mu=(ref_C0[0, 1], ref_C0[3, 3]))
    rinfo, _ = ref.solve(rtol=1e-10)
    for gi, oi in zip(infos, rinfo):
        assert gi['converged'] and gi['relative_residual'] <= 1e-10
        assert abs(gi['iterations'] - oi['iterations']) <= 1, (gi, oi)
        assert gi['restarts'] == oi['restarts'] == 0
    u = res.solutions.cpu().numpy()
    for l in range(6):
        f = hybrid.rhs(pool_h, index_h, np.eye(6)[l], grid, vox)
        r = f - hybrid.apply(pool_h, index_h, u[l:l + 1], grid, vox)
        assert np.linalg.norm(r) / np.linalg.norm(f) <= 1.0001e-10, l
    H = res.effective_tensor()
    Href = np.stack([hybrid.avg_stress(pool_h, index_h, ref.solutions[j:j + 1], np.eye(6)[j], grid, vox)
                     for j in range(6)], axis=1)
    assert rel(H, Href) < 1e-10
    assert np.abs(H - H.T).max() < 1e-9 * np.abs(H).max()


def test_hybrid_resident_laminate_closed_form():
    """Rotated orthotropic / anisotropic / isotropic z-layers: C^H equals the laminate closed form."""
    rng = np.random.default_rng(21)
    grid = (16, 8, 8)
    layers = np.array([0] * 6 + [1] * 4 + [2] * 6, np.int32)
    phase = np.broadcast_to(layers[:, None, None], grid).copy()
    materials = {0: dict(kind='isotropic', E=1.0, nu=0.3),
                 1: dict(kind='orthotropic', frame='local', E1=10, E2=2, E3=1, nu12=0.3, nu13=0.2, nu23=0.25,
                         G12=1.5, G13=0.8, G23=0.6),
                 2: dict(kind='anisotropic', C=spd(rng, 0.3))}
    pool, index, meta = mat.build_hybrid_material(phase, materials, orientation=random_rotation(rng), device='cuda')
    op = enable_fi_reference(HybridFIOperator(pool, index, grid, hex8.voxel_of(grid)))
    C, infos = compute_effective_stiffness_resident(op, tol=1e-11)
    assert all(i['converged'] for i in infos)
    pool_h = pool.cpu().numpy()
    want = hybrid.laminate([hybrid.unpack(pool_h[0]), hybrid.unpack(pool_h[3]), hybrid.unpack(pool_h[2])],
                           [6, 4, 6])
    assert rel(C[0].numpy(), want) < 1e-9


def test_hybrid_resident_at_64_matches_generic_driver():
    """The 64^3 three-phase composite of the hybrid bench: the resident C^H against the generic driver's
    (compute_effective_stiffness_batched, one load at a time), both certified to 1e-10."""
    import bench
    n = 64
    print('hybrid resident 64^3 composite', flush=True)
    phase, materials, orientation = bench.composite(n, 20261104, 4)
    pool, index, _ = mat.build_hybrid_material(phase, materials, orientation=orientation)
    op = enable_fi_reference(HybridFIOperator(pool, index, (n, n, n), (1.0 / n,) * 3))
    res = HybridResident(op).capture()
    infos, _ = res.solve(1e-10)
    assert all(i['converged'] and i['relative_residual'] <= 1e-10 for i in infos)
    C, info = pgj.compute_effective_stiffness_batched(op, tol=1e-10, max_iter=4000)
    print('  resident iterations', [i['iterations'] for i in infos], 'generic',
          [info[0][J]['iters'] for J in range(6)], flush=True)
    assert rel(res.effective_tensor(), np.asarray(C[0])) < 1e-8


def test_hybrid_resident_cap_raises_and_warm_start():
    grid = (8, 8, 8)
    _, pool, index, vox = multiphase(grid, 5)
    op = enable_fi_reference(HybridFIOperator(pool, index, grid, vox))
    res = HybridResident(op, cap=3).capture()
    with pytest.raises(CertificationError):
        res.solve(1e-10)
    res = HybridResident(op).capture()
    first, _ = res.solve(1e-9)
    again, _ = res.solve(1e-9)   # from the certified solutions
    assert all(b['iterations'] <= a['iterations'] for a, b in zip(first, again))
    with pytest.raises(ValueError):
        HybridResident(HybridFIOperator(pool, index, grid, vox))
