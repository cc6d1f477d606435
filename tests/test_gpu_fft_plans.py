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
        _set_all(r.op.grid, 8, -1)
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
        _set_all(r.op.grid, 8, -1)
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
        _set_all(m.op.grid, 4, -1)
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


def _plan_z(grid, es, plan):
    _lib.check(_lib.lib().pgj_fft_set_plan(grid, es, 1, plan))


def test_register_z_pass_preconditioner_matches_oracle():
    """The 64^3 z pass with its transforms in registers (plan 2): the first direction p0 = P s G s P r of a
    6-load set (two 3-load launches) and of a 1-load set against the oracle's preconditioner."""
    print('register z pass: 64^3 preconditioner', flush=True)
    n = 64
    a = design.simp(design.filtered_noise(n, 20261104, 4), 1e-6)
    rng = np.random.default_rng(17)
    r6 = rng.standard_normal((6, 3 * n ** 3))
    G6 = green.symbol((n, n, n), hex8.voxel_of((n, n, n)), 15 / 26, 5 / 13)
    ref = green.precond(r6, G6, (n, n, n), green.scale(a, .5, C0))
    for kw, loads in ((dict(load_mode='parallel', batch=6), 6), (dict(load_mode='serial'), 1)):
        res = pgj.Resident(n, **kw)
        _plan_z(res.op.grid, 8, 2)
        try:
            assert tune.fft_plans(res.op.grid, 8) == (0, 2, 0)
            res.material(torch.tensor(a, device='cuda'))
            res.select('0.5')
            res._setup(range(res.batch), 1e-8)
            res.r.copy_(res.to_soa(torch.tensor(r6[:loads], device='cuda')))
            res.initialize()
            p0 = res.to_aos(res.p0).cpu().numpy()
        finally:
            _plan_z(res.op.grid, 8, -1)
        err = np.abs(p0 - ref[:loads]).max() / np.abs(ref[:loads]).max()
        print(f'  {loads} load(s): rel err {err:.2e}', flush=True)
        assert err < 1e-13


def test_register_z_pass_solve_matches_oracle_64():
    """Cold 6-load solve at the 64^3 workload (alpha 0.5, parallel, two 3-load halves) with the register z pass:
    iterations and C^H against the oracle's committed solves, every residual rechecked on the CPU."""
    import test_gpu_configs as cfg
    print('register z pass: 64^3 cold solve', flush=True)
    a = cfg._a64()
    g = cfg._golden()
    r = pgj.Resident(64, load_mode='parallel', batch=6)
    _plan_z(r.op.grid, 8, 2)
    try:
        r.capture()
        r.material(torch.tensor(a, device='cuda'))
        r.select('0.5')
        infos, _ = r.solve(rtol=1e-8)
    finally:
        _plan_z(r.op.grid, 8, -1)
    its = np.array([i['iterations'] for i in infos])
    H = r.effective_tensor().cpu().numpy()
    err = np.abs(H - g['CH_a05']).max() / np.abs(g['CH_a05']).max()
    print(f'  iterations {its.tolist()} oracle {g["iterations_a05"].tolist()}, C^H rel err {err:.2e}', flush=True)
    assert all(i['converged'] and i['relative_residual'] <= 1e-8 for i in infos)
    assert np.abs(its - g['iterations_a05']).max() <= cfg.ITER_SPREAD_64['a05']
    assert err <= 1e-10
    u = r.solutions.cpu().numpy()
    for l in range(6):
        f = hex8.rhs(a, np.eye(6)[l], C0)
        res = np.linalg.norm(f - hex8.apply(a, u[l:l + 1], C0)) / np.linalg.norm(f)
        assert res <= 1.0001e-8, (l, res)


def test_register_z_pass_fp32_inner_64():
    """The FP32-inner refinement at 64^3 with the FP32 register z pass certifies with the FP64 C^H."""
    print('register z pass: FP32 inner 64^3', flush=True)
    n = 64
    a = torch.tensor(design.simp(design.filtered_noise(n, 11, 2), 1e-6), device='cuda')
    ref = pgj.Resident(n, load_mode='parallel', batch=6).capture()
    ref.material(a)
    ref.select(.5)
    ref.solve(1e-10)
    H = ref.effective_tensor().cpu().numpy()
    m = MixedResident(n)
    _plan_z(m.op.grid, 4, 2)
    try:
        m.capture()
        m.material(0, a)
        m.select(0, .5)
        infos, _ = m.solve(1e-10)
    finally:
        _plan_z(m.op.grid, 4, -1)
    assert all(i['converged'] for i in infos)
    Hm = m.effective_tensor(0).cpu().numpy()
    assert np.abs(Hm - H).max() / np.abs(H).max() < 1e-9


def test_register_z_pass_plan_is_64_only():
    r = pgj.Resident(32, load_mode='parallel', batch=6)
    with pytest.raises(ValueError, match='plan 2'):
        _plan_z(r.op.grid, 8, 2)
    r64 = pgj.Resident(64, load_mode='parallel', batch=6)
    with pytest.raises(ValueError, match='plan 2'):
        _lib.check(_lib.lib().pgj_fft_set_plan(r64.op.grid, 8, 0, 2))
