"""Radix-8 FFT plans and the autotuner (GPU).

Every pass of the resident PCG iteration can run its compile-time Stockham
transforms radix 8 first (64 = 8 x 8) instead of radix 4 first. The plan is a
library setting per (cube, precision, pass); with radix 8 everywhere the plain
solve must still match the oracle recurrence (iterations +-1, C^H to 1e-10) and
the FP32-inner refinement must still certify with C^H equal to the FP64 one.
Each test restores the default plans.
"""
import numpy as np
import pytest
import torch

import pgj
from oracle import design, green, hex8, pcg
from pgj import _lib, tune
from pgj.mixed import MixedResident

pytestmark = pytest.mark.gpu
C0 = hex8.isotropic(15 / 26, 5 / 13)


def _set_all(grid, es, plan):
    for pas in range(3):
        _lib.check(_lib.lib().pgj_fft_set_plan(grid, es, pas, plan))


@pytest.mark.parametrize('n', [16, 24, 32, 48, 64])
def test_radix8_plans_match_oracle(n):
    print(f'radix-8 plans, FP64 n={n}', flush=True)
    a = design.simp(design.filtered_noise(n, 20261104 + n, 2), 1e-4)
    r = pgj.Resident(n, load_mode='parallel', batch=6)
    _set_all(r.op.grid, 8, 1)
    try:
        assert tune.fft_plans(r.op.grid, 8) == (1, 1, 1)
        r.capture()
        r.material(torch.tensor(a, device='cuda'))
        r.select('0.5')
        infos, _ = r.solve(rtol=1e-10)
    finally:
        _set_all(r.op.grid, 8, 0)
    ref = pcg.ResidentPCG(a, C0, np.eye(6), batch=6, s=green.scale(a, .5, C0))
    rinfo, _ = ref.solve(rtol=1e-10)
    for gi, oi in zip(infos, rinfo):
        assert gi['converged'] and abs(gi['iterations'] - oi['iterations']) <= 1, (gi, oi)
    H = r.effective_tensor().cpu().numpy()
    Href = design.effective_tensor(a, ref.solutions, C0)
    assert np.abs(H - Href).max() / np.abs(Href).max() < 1e-10


@pytest.mark.parametrize('n', [96, 128])
def test_radix8_plans_large_certify(n):
    print(f'radix-8 plans, FP64 n={n}', flush=True)
    a = design.simp(design.filtered_noise(n, 7, 4), 1e-4)
    r = pgj.Resident(n, loads=np.eye(6)[[1, 5]], load_mode='parallel', batch=2)
    _set_all(r.op.grid, 8, 1)
    try:
        r.capture()
        r.material(torch.tensor(a, device='cuda'))
        r.select('0.5')
        infos, _ = r.solve()
    finally:
        _set_all(r.op.grid, 8, 0)
    assert all(i['converged'] for i in infos)
    u = r.solutions.cpu().numpy()
    for l, m in enumerate((1, 5)):
        f = hex8.rhs(a, np.eye(6)[m], C0)
        res = np.linalg.norm(f - hex8.apply(a, u[l:l + 1], C0)) / np.linalg.norm(f)
        assert res <= 1.0001e-8, (n, m, res)


@pytest.mark.parametrize('n', [32, 64])
def test_radix8_plans_fp32_inner(n):
    print(f'radix-8 plans, FP32 inner n={n}', flush=True)
    a = torch.tensor(design.simp(design.filtered_noise(n, 11, 2), 1e-6), device='cuda')
    ref = pgj.Resident(n, load_mode='parallel', batch=6).capture()
    ref.material(a)
    ref.select(.5)
    ref.solve(1e-10)
    H = ref.effective_tensor().cpu().numpy()
    m = MixedResident(n)
    _set_all(m.op.grid, 4, 1)
    try:
        m.capture()
        m.material(0, a)
        m.select(0, .5)
        infos, _ = m.solve(1e-10)
    finally:
        _set_all(m.op.grid, 4, 0)
    assert all(i['converged'] for i in infos)
    Hm = m.effective_tensor(0).cpu().numpy()
    assert np.abs(Hm - H).max() / np.abs(H).max() < 1e-9


def test_tuner_choices_certify():
    print('tuner at 32^3', flush=True)
    n = 32
    a = torch.tensor(design.simp(design.filtered_noise(n, 3, 2), 1e-4), device='cuda')
    r = pgj.Resident(n, load_mode='parallel', batch=6).capture()
    try:
        choice = r.tune(reps=2)
        assert set(choice['fft']) == {'fxy', 'fz', 'fxyi'}
        assert tune.fft_plans(r.op.grid, 8) == tuple(choice['fft'][k]['plan'] for k in ('fxy', 'fz', 'fxyi'))
        r.material(a)
        r.select('0.5')
        infos, _ = r.solve()
        assert all(i['converged'] for i in infos)
    finally:
        tune.reset(r.op.grid, 3, 8)
