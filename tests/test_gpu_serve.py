"""Continuous-batching server: results are the same alone and in the mix (GPU).

Three clients of two grid sizes and three Emin values submit staggered
design-step requests; each one's C^H, iteration counts and densities after
every step must equal those of the same client served alone (same slot
capacity), to 1e-12. One client is also compared with the plain resident path
(bench.py's step) to the plain-path tolerance.
"""
import numpy as np
import pytest
import torch

import pgj
from pgj.outer import PaperOuter
from pgj.serve import Client, Server, latency_stats

pytestmark = pytest.mark.gpu
SIZES = (16, 24)
BUDGET = 4 * 2 ** 30
SPECS = [(16, 1e-4, .5, 11), (24, 1e-6, .25, 12), (16, 1e-9, .75, 13)]
STEPS = 3


def clients():
    return [Client(i, n, emin, alpha, seed, STEPS) for i, (n, emin, alpha, seed) in enumerate(SPECS)]


@pytest.mark.parametrize('precision', ['fp64', 'mixed'])
def test_same_alone_and_in_mix(precision):
    mix = clients()
    # staggered arrivals: client 1 and 2 join while client 0 is solving
    arrivals = sorted([(0.0 + 0.001 * k, 0) for k in range(STEPS)] + [(0.002 + 0.001 * k, 1) for k in range(STEPS)]
                      + [(0.004 + 0.001 * k, 2) for k in range(STEPS)])
    print(f'serve mix {precision}', flush=True)
    server = Server(SIZES, BUDGET, max_slots=4, precision=precision)
    done = server.run(mix, arrivals)
    assert len(done) == 3 * STEPS and server.unserved == 0
    stats = latency_stats(done)
    assert stats['count'] == 9 and stats['p99'] >= stats['p50'] > 0
    for i in range(3):
        alone = clients()
        print(f'serve alone {i} {precision}', flush=True)
        solo = Server(SIZES, BUDGET, max_slots=4, precision=precision)
        solo.run(alone, [(0.001 * k, i) for k in range(STEPS)])
        a, b = mix[i].history, alone[i].history
        assert len(a) == len(b) == STEPS
        for ha, hb in zip(a, b):
            assert ha['iterations'] == hb['iterations']
            assert ha['bisections'] == hb['bisections']
            assert ha['max_rel'] <= 1e-8
            np.testing.assert_allclose(ha['CH'], hb['CH'], rtol=0, atol=1e-12 * np.abs(hb['CH']).max())
        assert torch.equal(mix[i].x, alone[i].x) or \
            float((mix[i].x - alone[i].x).abs().max()) < 1e-12


def test_client_matches_plain_path():
    c = clients()[:1]
    server = Server((16,), BUDGET, max_slots=2)
    server.run(c, [(0.0, 0), (0.001, 0), (0.002, 0)])
    n, emin, alpha, seed = SPECS[0]
    res = pgj.Resident(n, loads=np.eye(6), load_mode='parallel', batch=6).capture()
    outer = PaperOuter(n, 0.35, rmin=3.0, beta=4.0, eta=0.5, emin=emin, move=0.2).capture()
    outer.set_initial(Client(0, n, emin, alpha, seed, 1).x)
    for k in range(STEPS):
        outer.evaluate()
        res.material(outer.a)
        res.select(alpha)
        infos, _ = res.solve(rtol=1e-8)
        _, local, _ = res.objective_energy('shear')
        H = res.effective_tensor().cpu().numpy()
        count = outer.step(local)
        outer.advance()
        h = c[0].history[k]
        assert count == h['bisections']
        for j in range(6):
            assert abs(infos[j]['iterations'] - h['iterations'][j]) <= 1
        np.testing.assert_allclose(h['CH'], H, rtol=0, atol=1e-10 * np.abs(H).max())
    assert float((c[0].x - outer.x).abs().max()) < 1e-8


def test_sequential_baseline_serves_in_order():
    from pgj.serve import SequentialServer
    c = clients()[:2]
    base = SequentialServer((16, 24))
    done = base.run(c, [(0.0, 0), (0.0, 1), (0.001, 0)])
    assert len(done) == 3 and base.unserved == 0
    assert [len(x.history) for x in c] == [2, 1]
    assert all(h['max_rel'] <= 1e-8 for x in c for h in x.history)


def test_closed_loop_clients_submit_after_each_answer():
    """think=...: each client's next request arrives a think time after its previous one was answered, so no
    request waits behind its own client's step; every step is certified."""
    c = clients()[:2]
    server = Server(SIZES, BUDGET, max_slots=2)
    done = server.run(c, [(0.0, 0), (0.0, 1)], think=lambda client: 0.002)
    assert len(done) == 2 * STEPS and server.unserved == 0
    for client in c:
        reqs = sorted((r for r in done if r.client is client), key=lambda r: r.arrival)
        assert len(reqs) == STEPS
        for prev, nxt in zip(reqs, reqs[1:]):
            assert nxt.arrival >= prev.end + 0.002 - 1e-9
    assert all(h['max_rel'] <= 1e-8 for x in c for h in x.history)
