"""The reference's own callers, replayed through pgj's Resident / PaperResident constructors (GPU).

- examples/shear_design.py:46-64: Study I, ``Resident(backend, n, load_mode=mode)``,
  ``outer.evaluate(beta)`` -> ``material`` -> ``select(method)`` -> ``solve`` ->
  ``energy() -> (moduli, sens)`` -> ``outer.step`` -> ``advance``, at n = 8 for 3
  steps with methods green, gj, 0.5 in both load modes. Every combination
  designs the same structure (all solves certified to 1e-8); the reference's
  release smoke run of exactly this (validation/gpu_smoke.json) records a
  largest relative objective difference of 6.6e-10 and density difference of
  2.0e-9 across the six combinations. The 0.5/serial run is also compared step
  by step with the oracle's Study-I loop.
- exp/asg_sctop_reproduction_20261105/verify.py:86-93: Study II,
  ``PaperResident(b, 6, 'auxetic', .2, load_mode=mode)``, one solve at rtol
  1e-10 and one OC step; serial and parallel agree to 1e-8 (the reference's
  assertion).
"""
import numpy as np
import pytest
import torch

import pgj
from oracle import design, green, hex8, pcg

pytestmark = pytest.mark.gpu


def _shear_design(n, steps, method, mode, seed=20261104):
    backend = pgj.Candidates()
    resident = pgj.Resident(backend, n, load_mode=mode)
    resident.capture()
    resident.outer.set_initial(pgj.pilot.initial(n, seed))
    resident.solutions.zero_()
    rows = []
    for k in range(steps):
        beta = pgj.pilot.CONFIG['beta_stages'][min(k // 25, 3)]
        resident.outer.evaluate(beta)
        resident.material(resident.outer.a)
        resident.select(method)
        infos, checks = resident.solve()
        moduli, energy = resident.energy()
        resident.outer.step(energy)
        assert moduli.shape == (3,) and energy.shape == (n, n, n)
        rows.append(dict(objective=float(moduli.mean()), iterations=[i['iterations'] for i in infos],
                         worst=max(i['relative_residual'] for i in infos), x=resident.outer.x.cpu().numpy().copy(),
                         new=resident.outer.new.cpu().numpy().copy()))
        assert all(i['converged'] for i in infos)
        if k < steps - 1:
            resident.outer.advance()
    return rows


def test_shear_design_replay_agrees_across_methods_and_modes():
    print('shear_design replay n=8', flush=True)
    n, steps = 8, 3
    runs = {(m, mode): _shear_design(n, steps, m, mode) for mode in ('serial', 'parallel')
            for m in ('green', 'gj', '0.5')}
    base = runs[('0.5', 'serial')]
    obj_diff = dens_diff = 0.0
    for key, rows in runs.items():
        for a, b in zip(rows, base):
            obj_diff = max(obj_diff, abs(a['objective'] - b['objective']) / abs(b['objective']))
            dens_diff = max(dens_diff, float(np.abs(a['new'] - b['new']).max()))
            assert a['worst'] <= 1e-8
    print(f'shear_design replay: objective rel diff {obj_diff:.2e}, density diff {dens_diff:.2e}', flush=True)
    assert obj_diff <= 1e-9
    assert dens_diff <= 1e-8
    # serial and parallel with one method: same per-load arithmetic up to partial-sum grouping
    for m in ('green', 'gj', '0.5'):
        for a, b in zip(runs[(m, 'serial')], runs[(m, 'parallel')]):
            assert abs(a['objective'] - b['objective']) <= 1e-9 * abs(b['objective'])
            assert np.abs(a['new'] - b['new']).max() <= 2e-9


def test_shear_design_matches_oracle_study1_loop():
    """The 0.5 / serial run step by step against the oracle's Study-I loop (ResidentPCG + GaussianOC)."""
    n, steps = 8, 3
    rows = _shear_design(n, steps, '0.5', 'serial')
    C0 = hex8.isotropic(4 / 3, 1.0)
    loads = np.eye(6)[3:]
    oc = design.GaussianOC(n, volume=.3, move=.1, emin=1e-4)
    x = pgj.pilot.initial(n, 20261104)
    sol = np.zeros((3, 3 * n ** 3))
    for k in range(steps):
        beta = pgj.pilot.CONFIG['beta_stages'][min(k // 25, 3)]
        ph, dp, a = oc.physical(x, beta)
        ref = pcg.ResidentPCG(a, C0, loads, batch=1, s=green.scale(a, .5, C0))
        ref.solutions[:] = sol
        infos, _ = ref.solve(rtol=1e-8)
        sol = ref.solutions.copy()
        e = hex8.energies(sol, a.shape, loads, C0)
        moduli = (e * a.reshape(1, -1)).mean(1)
        new, g, dv = oc.update(x, ph, dp, e.mean(0).reshape(a.shape), beta)
        its = [i['iterations'] for i in infos]
        got = rows[k]
        print(f'step {k}: iterations {got["iterations"]} oracle {its}', flush=True)
        assert max(abs(p - q) for p, q in zip(got['iterations'], its)) <= 1
        assert abs(got['objective'] - moduli.mean()) <= 1e-9 * abs(moduli.mean())
        np.testing.assert_allclose(got['x'], x, rtol=0, atol=2e-8)
        np.testing.assert_allclose(got['new'], new, rtol=0, atol=2e-8)
        x = new


def test_paper_resident_serial_parallel_scheduling():
    print('PaperResident scheduling n=6', flush=True)
    b = pgj.PaperBackend()
    out = []
    for mode in ('serial', 'parallel'):
        r = pgj.PaperResident(b, 6, 'auxetic', .2, load_mode=mode)
        r.capture()
        r.outer.set_initial(pgj.pilot.paper_initial(6, 1))
        r.outer.evaluate()
        r.material(r.outer.a)
        r.select('0.5')
        infos, _ = r.solve(1e-10)
        assert len(infos) == 6 and all(i['converged'] for i in infos)
        f, e, diagonal = r.energy(10)
        assert e.shape == (6, 6, 6) and diagonal.shape == (6,)
        r.outer.step(e)
        out.append((float(f), r.outer.new.cpu().numpy()))
        H = r.effective_tensor().cpu().numpy()
        assert np.abs(H - H.T).max() <= 1e-12 * np.abs(H).max()
    assert abs(out[0][0] - out[1][0]) < 1e-8
    assert np.abs(out[0][1] - out[1][1]).max() < 1e-8


def test_reference_constructor_errors():
    with pytest.raises(TypeError):
        pgj.Resident(pgj.Candidates())
    with pytest.raises(TypeError):
        pgj.Resident(8, np.eye(6))
    with pytest.raises(ValueError):
        pgj.PaperResident(pgj.PaperBackend(), 6, 'stiff', .2)
    bad = pgj.Candidates()
    bad.C0 = np.arange(36.0).reshape(6, 6)
    with pytest.raises(ValueError):
        pgj.Resident(bad, 8)


def test_replay_state_call_sequence(tmp_path):
    """examples/replay_state.py:38-59 through pgj: a Study-II state file with the archive's keys (a, rho, u0),
    SelectorResident(backend, n, objective, .2, rmin=3.2, emin=eta, load_mode), material + outer.ph from the
    file, and for each method: select, warm solve from u0, reset to u0, timed select + solve, C^H. Every method
    certifies, selects the oracle's alpha and gives the same C^H; C^H matches the oracle's solve."""
    from oracle import select as osel
    from pgj import checkpoint
    n, eta = 16, 1e-6
    outer = pgj.outer.PaperOuter(n, .2, rmin=3.2, emin=eta).capture()
    outer.set_initial(torch.tensor(pgj.density.filtered_noise(n, 5, 2), device='cuda'))
    outer.evaluate()
    a, rho = outer.a.cpu().numpy(), outer.ph.cpu().numpy()
    C0 = hex8.isotropic(15 / 26, 5 / 13)
    warm = pcg.ResidentPCG(a, C0, np.eye(6), batch=6, s=green.scale(a, .5, C0))
    warm.solve(rtol=1e-4)
    path = tmp_path / 'state.npz'
    checkpoint.save_state(path, a=a, rho=rho, u0=warm.solutions)
    state = np.load(path, allow_pickle=False)
    for mode in ('serial', 'parallel'):
        print(f'replay_state call sequence, {mode}', flush=True)
        r = pgj.SelectorResident(pgj.PaperBackend(), n, 'shear', .2, rmin=3.2, emin=eta, load_mode=mode)
        r.capture()
        r.material(torch.as_tensor(state['a'], device='cuda'))
        r.outer.ph.copy_(torch.as_tensor(state['rho'], device='cuda'))
        u0 = torch.as_tensor(state['u0'], device='cuda')
        Cs = {}
        for method in ('gj', 'eta', 'compact8', 'gray', '0.5'):
            r.select(method)
            r.solutions.copy_(u0)
            r.solve()
            r.solutions.copy_(u0)
            r.select(method)
            loads, checks = r.solve()
            assert len(loads) == 6 and all(i['converged'] for i in loads)
            Cs[method] = r.effective_tensor().cpu().numpy()
            want = {'gj': 1.0, 'eta': osel.eta_alpha(eta), 'compact8': osel.compact8_alpha(a),
                    'gray': osel.gray_alpha(a, rho), '0.5': 0.5}[method]
            assert abs(float(r.alpha) - want) <= 1e-6, (method, float(r.alpha), want)
        H = Cs['0.5']
        for method, C in Cs.items():
            assert np.abs(C - H).max() <= 1e-9 * np.abs(H).max(), method
        ref = pcg.ResidentPCG(a, C0, np.eye(6), batch=3, s=green.scale(a, .5, C0))
        ref.solutions[:] = warm.solutions
        ref.solve(rtol=1e-8)
        Href = design.effective_tensor(a, ref.solutions, C0)
        assert np.abs(H - Href).max() <= 1e-10 * np.abs(Href).max()
