"""Mixed-precision form: FP32 inner PCG inside FP64 refinement (GPU).

Every load passes the FP64 true-residual certification; C^H agrees with the
plain FP64 path and the oracle to 1e-8 and the element energies (the
sensitivity input) to 1e-6, on a 16^3 and a 64^3 design; two designs in one
set give the same answers as each alone.
"""
import numpy as np
import pytest
import torch

import pgj
from pgj.mixed import MixedResident
from oracle import design, green, hex8, pcg

pytestmark = pytest.mark.gpu
LAM, MU = 15 / 26, 5 / 13
C0 = hex8.isotropic(LAM, MU)


def density(n, emin, radius, seed=20261104):
    return design.simp(design.filtered_noise(n, seed, radius), emin)


def rel(a, b):
    a = a.cpu().numpy() if torch.is_tensor(a) else np.asarray(a)
    b = b.cpu().numpy() if torch.is_tensor(b) else np.asarray(b)
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize('n,emin,radius,alpha', [(16, 1e-4, 2, .5), (64, 1e-6, 4, .5)])
def test_mixed_matches_fp64(n, emin, radius, alpha):
    print(f'mixed n={n} emin={emin}', flush=True)
    a = density(n, emin, radius)
    m = MixedResident(n).capture()
    m.material(0, torch.tensor(a, device='cuda'))
    m.select(0, alpha)
    infos, stats = m.solve(rtol=1e-8)
    assert all(i['converged'] and i['relative_residual'] <= 1e-8 for i in infos)
    assert stats['outer'] >= 2
    ref = pgj.Resident(n, load_mode='parallel', batch=6).capture()
    ref.material(torch.tensor(a, device='cuda'))
    ref.select(alpha)
    ref.solve(rtol=1e-8)
    assert rel(m.effective_tensor(0), ref.effective_tensor()) < 1e-8
    assert rel(m.energies(0), ref.energies()) < 1e-6
    u = m.solutions(0).cpu().numpy()
    for l in (0, 4):   # the FP64 residual, rechecked on the CPU
        f = hex8.rhs(a, np.eye(6)[l], C0)
        assert np.linalg.norm(f - hex8.apply(a, u[l:l + 1], C0)) / np.linalg.norm(f) <= 1.0001e-8


def test_mixed_matches_oracle_and_designs_are_independent():
    n = 16
    fields = [density(n, 1e-4, 2, seed) for seed in (3, 4)]
    both = MixedResident(n, designs=2).capture()
    for d, a in enumerate(fields):
        both.material(d, torch.tensor(a, device='cuda'))
        both.select(d, .5)
    both.solve(rtol=1e-10)
    for d, a in enumerate(fields):
        alone = MixedResident(n)
        alone.material(0, torch.tensor(a, device='cuda'))
        alone.select(0, .5)
        alone.solve(rtol=1e-10)
        assert rel(both.effective_tensor(d), alone.effective_tensor(0)) < 1e-12
        ref = pcg.ResidentPCG(a, C0, np.eye(6), batch=6, s=green.scale(a, .5, C0))
        ref.solve(rtol=1e-10)
        assert rel(both.effective_tensor(d), design.effective_tensor(a, ref.solutions, C0)) < 1e-9


def test_mixed_rejects_runtime_only_grids():
    with pytest.raises(ValueError):
        MixedResident(12)
