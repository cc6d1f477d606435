"""Parity at the benchmark configurations (GPU).

- 64^3 (Emin 1e-6, radius-4 noise, seed 20261104), alpha 0, 0.5 and 1, serial
  and parallel load modes: cold solves against the oracle's committed
  results (tests/golden/solve_64_alpha.npz: per-load iterations within the
  spread rounding alone produces, ITER_SPREAD_64, and C^H to 1e-10), every residual rechecked on the CPU at alpha 0.5, C^H independent
  of alpha (1e-9) and of the load mode.
- Ten warm-started Study-II design steps (volume 0.35, move 0.2): at 16^3 step
  by step against the oracle's loop (iterations +-1, equal bisection counts,
  densities within DENSITY_BOUND_16), and at 64^3 serial against parallel.
- 48 designs of n in {24, 32, 48, 64} in one batched graph against solo solves.
- The FP32-inner refinement at 128^3 / Emin 1e-9 against the FP64 path.
- The server at sizes {32, 64, 128} and Emin {1e-4, 1e-6, 1e-9}: each result the
  same alone and in the mix.
"""
import os

import numpy as np
import pytest
import torch

import pgj
from oracle import design, green, hex8, pcg
from pgj.batched import BatchedDesigns
from pgj.mixed import MixedResident
from pgj.outer import PaperOuter
from pgj.serve import Client, Server

pytestmark = pytest.mark.gpu
LAM, MU = 15 / 26, 5 / 13
C0 = hex8.isotropic(LAM, MU)
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden', 'solve_64_alpha.npz')
# OC densities over ten design steps, GPU vs oracle at 16^3. Each step's solves are certified to 1e-8; the GPU
# and numpy iterates differ by rounding, so once the iteration counts of a load differ by one (or rounding has
# grown over a hundred iterations) the solutions differ at the 1e-9 level and the exponent-0.3 update moves a
# density by a few 1e-10 (measured: 4e-15 to 1e-13 in steps 0-3, then at most 4.5e-10).
DENSITY_BOUND_16 = 1e-9


def _golden():
    with np.load(GOLDEN, allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def _a64():
    return pgj.density.simp(pgj.density.filtered_noise(64, 20261104, 4), 1e-6)


@pytest.fixture(scope='module')
def runs64():
    """Cold 6-load solves at 64^3 for alpha in {0, 0.5, 1} x {serial, parallel}."""
    a = torch.tensor(_a64(), device='cuda')
    out = {}
    for mode in ('serial', 'parallel'):
        r = pgj.Resident(64, load_mode=mode, batch=6 if mode == 'parallel' else None).capture()
        for tag, alpha in (('a0', 0.0), ('a05', 0.5), ('a1', 1.0)):
            print(f'64^3 cold solve alpha {alpha} {mode}', flush=True)
            r.solutions.zero_()
            r.material(a)
            r.select(str(alpha))
            infos, _ = r.solve(rtol=1e-8)
            out[(tag, mode)] = dict(infos=infos, H=r.effective_tensor().cpu().numpy(),
                                    u=r.solutions.cpu().numpy() if (tag, mode) == ('a05', 'parallel') else None)
    return out


# Per-load iteration spread at 64^3 that rounding alone produces: the GPU's own serial and parallel runs differ
# only in the grouping of their partial sums and still stop up to 3 iterations apart at alpha 0 (about 600
# iterations; measured 618/621, 580/574, ...) and up to 2 apart at alpha 0.5 (about 87), where the true
# residual crosses 1e-8 on a plateau of the CG residual. The same spread separates the GPU from the numpy
# recurrence; at 16^3 and 32^3 (50-170 iterations) the counts stay within 1 (test_gpu_resident.py).
ITER_SPREAD_64 = {'a0': 3, 'a05': 2, 'a1': 2}


@pytest.mark.parametrize('tag', ['a0', 'a05', 'a1'])
@pytest.mark.parametrize('mode', ['serial', 'parallel'])
def test_64_cold_solve_matches_oracle(runs64, tag, mode):
    g = _golden()
    run = runs64[(tag, mode)]
    its = np.array([i['iterations'] for i in run['infos']])
    Href = g['CH_' + tag]
    err = np.abs(run['H'] - Href).max() / np.abs(Href).max()
    print(f'64^3 {tag} {mode}: iterations {its.tolist()} oracle {g["iterations_" + tag].tolist()}, '
          f'C^H rel err {err:.2e}', flush=True)
    assert all(i['converged'] and i['relative_residual'] <= 1e-8 for i in run['infos'])
    assert np.abs(its - g['iterations_' + tag]).max() <= ITER_SPREAD_64[tag]
    assert err <= 1e-10


def test_64_residuals_rechecked_on_cpu(runs64):
    a = _a64()
    u = runs64[('a05', 'parallel')]['u']
    for l in range(6):
        f = hex8.rhs(a, np.eye(6)[l], C0)
        res = np.linalg.norm(f - hex8.apply(a, u[l:l + 1], C0)) / np.linalg.norm(f)
        assert res <= 1.0001e-8, (l, res)


def test_64_ch_independent_of_alpha_and_mode(runs64):
    H = runs64[('a05', 'parallel')]['H']
    for key, run in runs64.items():
        assert np.abs(run['H'] - H).max() <= 1e-9 * np.abs(H).max(), key
    for tag in ('a0', 'a05', 'a1'):
        s = [i['iterations'] for i in runs64[(tag, 'serial')]['infos']]
        p = [i['iterations'] for i in runs64[(tag, 'parallel')]['infos']]
        assert max(abs(x - y) for x, y in zip(s, p)) <= ITER_SPREAD_64[tag], (tag, s, p)


def _gpu_design_steps(n, emin, radius, steps, mode, seed=20261104):
    res = pgj.Resident(n, load_mode=mode, batch=6 if mode == 'parallel' else None).capture()
    outer = PaperOuter(n, .35, rmin=3.0, beta=4.0, eta=.5, emin=emin, move=.2).capture()
    outer.set_initial(torch.tensor(pgj.density.filtered_noise(n, seed, radius), device='cuda'))
    rows = []
    for k in range(steps):
        outer.evaluate()
        res.material(outer.a)
        res.select(0.5)
        infos, _ = res.solve(rtol=1e-8)
        obj, local, _ = res.objective_energy('shear')
        count = outer.step(local)
        rows.append(dict(iterations=[i['iterations'] for i in infos], bisections=count, objective=float(obj),
                         new=outer.new.cpu().numpy().copy(), worst=max(i['relative_residual'] for i in infos)))
        outer.advance()
    return rows


def test_ten_design_steps_16_match_oracle():
    n, emin, steps = 16, 1e-6, 10
    print('16^3 ten design steps (GPU)', flush=True)
    rows = _gpu_design_steps(n, emin, 2, steps, 'parallel')
    oc = design.PaperOC(n, .35, emin=emin, move=.2)
    x = design.filtered_noise(n, 20261104, 2)
    sol = np.zeros((6, 3 * n ** 3))
    worst = 0.0
    for k in range(steps):
        ph, dp, a = oc.physical(x)
        ref = pcg.ResidentPCG(a, C0, np.eye(6), batch=6, s=green.scale(a, .5, C0))
        ref.solutions[:] = sol
        infos, _ = ref.solve(rtol=1e-8)
        sol = ref.solutions.copy()
        obj, local, _ = design.shear_objective(a, sol, C0)
        new, count, _, _ = oc.update(x, ph, dp, local)
        got = rows[k]
        its = [i['iterations'] for i in infos]
        dev = float(np.abs(got['new'] - new).max())
        worst = max(worst, dev)
        print(f'step {k}: iterations {got["iterations"]} oracle {its}, bisections {got["bisections"]} / {count}, '
              f'density dev {dev:.2e}, objective rel {abs(got["objective"] - obj) / abs(obj):.2e}', flush=True)
        assert got['worst'] <= 1e-8
        assert max(abs(p - q) for p, q in zip(got['iterations'], its)) <= 1
        assert got['bisections'] == count
        assert abs(got['objective'] - obj) <= 1e-9 * abs(obj)
        x = new
    assert worst <= DENSITY_BOUND_16


def test_ten_design_steps_64_serial_vs_parallel():
    """The reference's scheduling check (exp/asg_sctop_reproduction_20261105/verify.py:86-93) over ten steps."""
    runs = {}
    for mode in ('serial', 'parallel'):
        print(f'64^3 ten design steps {mode}', flush=True)
        runs[mode] = _gpu_design_steps(64, 1e-6, 4, 10, mode)
    obj = dens = 0.0
    for s, p in zip(runs['serial'], runs['parallel']):
        assert s['worst'] <= 1e-8 and p['worst'] <= 1e-8
        assert s['bisections'] == p['bisections']
        obj = max(obj, abs(s['objective'] - p['objective']) / abs(p['objective']))
        dens = max(dens, float(np.abs(s['new'] - p['new']).max()))
    print(f'64^3 serial vs parallel over ten steps: objective rel {obj:.2e}, density {dens:.2e}', flush=True)
    # the reference's release check of serial vs parallel (validation/gpu_smoke.json: 6.6e-10 and 2.0e-9)
    assert obj <= 1e-9
    assert dens <= 5e-9


def test_batched_48_designs_match_solo():
    sizes = {24: 12, 32: 12, 48: 12, 64: 12}
    rng = np.random.default_rng(20261104)
    designs = {n: [(pgj.density.simp(pgj.density.filtered_noise(n, 20261104 + 7 * n + d, 2 if n < 48 else 4), 1e-4),
                    float(rng.choice([.25, .5, .75]))) for d in range(cnt)] for n, cnt in sizes.items()}
    print('batched 48 designs', flush=True)
    bd = BatchedDesigns(sizes).capture()
    for n, g in bd.groups.items():
        for d, (a, alpha) in enumerate(designs[n]):
            g.material(d, torch.tensor(a, device='cuda'))
            g.select(d, alpha)
    infos, _ = bd.solve(1e-8)
    assert all(i['converged'] for rows in infos.values() for i in rows)
    worst = 0.0
    solos = {}
    for n, g in bd.groups.items():
        solo = solos.setdefault(n, pgj.Resident(n, load_mode='parallel', batch=6).capture())
        for d, (a, alpha) in enumerate(designs[n]):
            solo.solutions.zero_()
            solo.material(torch.tensor(a, device='cuda'))
            solo.select(alpha)
            sinfo, _ = solo.solve(1e-8)
            Hs = solo.effective_tensor().cpu().numpy()
            Hb = g.effective_tensor(d).cpu().numpy()
            worst = max(worst, float(np.abs(Hb - Hs).max() / np.abs(Hs).max()))
            bits = [i['iterations'] for i in infos[n][6 * d:6 * d + 6]]
            assert max(abs(x - y['iterations']) for x, y in zip(bits, sinfo)) <= 2, (n, d)
    print(f'batched vs solo C^H max rel {worst:.2e}', flush=True)
    assert worst <= 1e-10


def test_mixed_128_emin_1e9_against_fp64():
    n = 128
    a = pgj.density.simp(pgj.density.filtered_noise(n, 20261104, 4), 1e-9)
    at = torch.tensor(a, device='cuda')
    print('128^3 Emin 1e-9 mixed', flush=True)
    m = MixedResident(n, inner_rtol=1e-4).capture()
    m.material(0, at)
    m.select(0, .5)
    infos, stats = m.solve(1e-8)
    assert all(i['converged'] and i['relative_residual'] <= 1e-8 for i in infos)
    Hm = m.effective_tensor(0).cpu().numpy()
    em = m.energies(0)
    um = m.solutions(0).cpu().numpy()
    print(f'128^3 mixed: outer {stats["outer"]}, inner {stats["inner"]}', flush=True)
    del m
    torch.cuda.empty_cache()
    print('128^3 Emin 1e-9 FP64', flush=True)
    r = pgj.Resident(n, load_mode='parallel', batch=6).capture()
    r.material(at)
    r.select(.5)
    rinfos, _ = r.solve(1e-8)
    Hr = r.effective_tensor().cpu().numpy()
    er = r.energies()
    ch_err = float(np.abs(Hm - Hr).max() / np.abs(Hr).max())
    # shear sensitivity density -sum_{l=3..5} e_l, max relative error against the FP64 solve
    sm, sr = -em[3:].sum(0), -er[3:].sum(0)
    sens_err = float(((sm - sr).abs().max() / sr.abs().max()).item())
    print(f'128^3 mixed vs FP64: C^H {ch_err:.2e}, sensitivity {sens_err:.2e}', flush=True)
    assert ch_err <= 1e-8
    assert sens_err <= 1e-6
    for l in (0, 4):
        f = hex8.rhs(a, np.eye(6)[l], C0)
        res = np.linalg.norm(f - hex8.apply(a, um[l:l + 1], C0)) / np.linalg.norm(f)
        assert res <= 1.0001e-8, (l, res)


SERVE_SPECS = [(32, 1e-4, .3, 31), (64, 1e-6, .6, 32), (128, 1e-9, .5, 33)]


def _serve_clients():
    return [Client(i, n, emin, alpha, seed, 2) for i, (n, emin, alpha, seed) in enumerate(SERVE_SPECS)]


def test_serve_sizes_32_64_128_same_alone_and_in_mix():
    sizes, budget = (32, 64, 128), 24 * 2 ** 30
    arrivals = sorted([(0.001 * k, 0) for k in range(2)] + [(0.002 + 0.001 * k, 1) for k in range(2)]
                      + [(0.004 + 0.001 * k, 2) for k in range(2)])
    mix = _serve_clients()
    print('serve 32/64/128 mix', flush=True)
    server = Server(sizes, budget, max_slots=2, precision='mixed')
    done = server.run(mix, arrivals)
    assert len(done) == 6 and server.unserved == 0 and not server.failed
    for i in range(3):
        alone = _serve_clients()
        print(f'serve 32/64/128 alone {i}', flush=True)
        solo = Server(sizes, budget, max_slots=2, precision='mixed')
        solo.run(alone, [(0.001 * k, i) for k in range(2)])
        for ha, hb in zip(mix[i].history, alone[i].history):
            assert ha['iterations'] == hb['iterations'] and ha['bisections'] == hb['bisections']
            assert ha['max_rel'] <= 1e-8
            np.testing.assert_allclose(ha['CH'], hb['CH'], rtol=0, atol=1e-12 * np.abs(hb['CH']).max())
        assert float((mix[i].x - alone[i].x).abs().max()) <= 1e-12
