"""Certification failure paths on the GPU, against the oracle (GPU).

The reference stops a load at its iteration cap or when its recurrence breaks
down (non-finite or non-positive p.q or r.z latches ``bad``), restarts an
uncertified group from the certified iterate at most twice, and raises
``RuntimeError(('true residual certification failed', info))``
(exp/asg_gpu_resident_20261103/resident.py:40-51, :186-209). Each case here
makes the solver fail, not the GPU: a cap of 3 iterations, a NaN in the SIMP
field (the load never activates, two restarts), and a NaN in the scale vector
(the first iteration latches ``bad``). Every kernel still runs to its end.
"""
import numpy as np
import pytest
import torch

import pgj
from oracle import design, green, hex8, pcg
from pgj.serve import Client, SequentialServer, Server

pytestmark = pytest.mark.gpu
LAM, MU = 15 / 26, 5 / 13
C0 = hex8.isotropic(LAM, MU)


def _oracle_failure(a, loads, batch, s, rtol, cap=None, monkeypatch=None):
    if cap is not None:
        monkeypatch.setattr(pcg, 'CAP', cap)
    ref = pcg.ResidentPCG(a, C0, loads, batch=batch, s=s)
    with pytest.raises(RuntimeError) as info:
        ref.solve(rtol=rtol)
    return info.value.args[0][1]


def _gpu_failure(r, rtol):
    with pytest.raises(pgj.CertificationError) as info:
        r.solve(rtol=rtol)
    assert isinstance(info.value, RuntimeError)
    assert info.value.args[0][0] == 'true residual certification failed'
    return info.value.info


def _same(gpu, ref):
    assert len(gpu) == len(ref)
    for g, o in zip(gpu, ref):
        assert g['iterations'] == o['iterations'], (g, o)
        assert g['restarts'] == o['restarts'], (g, o)
        assert g['converged'] == o['converged'] is False, (g, o)


def test_iteration_cap_stops_without_restart(monkeypatch):
    print('certify: cap 3 at 6^3', flush=True)
    n = 6
    a = design.simp(design.filtered_noise(n, 7, 1), 1e-6)
    loads = np.eye(6)[:3]
    r = pgj.Resident(n, loads=loads, load_mode='parallel', batch=3, cap=3).capture()
    r.material(torch.tensor(a, device='cuda'))
    r.select('green')
    gpu = _gpu_failure(r, 1e-10)
    ref = _oracle_failure(a, loads, 3, None, 1e-10, cap=3, monkeypatch=monkeypatch)
    _same(gpu, ref)
    assert all(g['iterations'] == 3 and g['restarts'] == 0 for g in gpu)
    # the same solver with the default cap certifies
    r2 = pgj.Resident(n, loads=loads, load_mode='parallel', batch=3).capture()
    r2.material(torch.tensor(a, device='cuda'))
    r2.select('green')
    infos, _ = r2.solve(rtol=1e-10)
    assert all(i['converged'] and i['iterations'] > 3 for i in infos)


def test_nan_material_restarts_twice_then_raises():
    print('certify: NaN in a at 8^3', flush=True)
    n = 8
    a = design.simp(design.filtered_noise(n, 5, 1), 1e-4)
    a[2, 3, 4] = np.nan
    loads = np.eye(6)[3:4]
    r = pgj.Resident(n, loads=loads, load_mode='serial').capture()
    r.material(torch.tensor(a, device='cuda'))
    r.select('gj')
    gpu = _gpu_failure(r, 1e-8)
    ref = _oracle_failure(a, loads, 1, green.scale(a, 1.0, C0), 1e-8)
    _same(gpu, ref)
    assert gpu[0]['iterations'] == 0 and gpu[0]['restarts'] == 2
    assert np.isnan(gpu[0]['rhs_norm']) and np.isnan(gpu[0]['threshold'])


def test_nan_scale_latches_bad_at_first_iteration():
    print('certify: NaN in s at 8^3', flush=True)
    n = 8
    a = design.simp(design.filtered_noise(n, 3, 1), 1e-4)
    loads = np.eye(6)[3:4]
    r = pgj.Resident(n, loads=loads, load_mode='serial').capture()
    r.material(torch.tensor(a, device='cuda'))
    r.select('0.5')
    s = green.scale(a, 0.5, C0)
    # node 5, component x: SoA index 5 on the device, AoS index 15 in the oracle
    r.scale[0, 5] = float('nan')
    s[0, 15] = np.nan
    gpu = _gpu_failure(r, 1e-8)
    assert bool(r.bad.any())
    ref = _oracle_failure(a, loads, 1, s, 1e-8)
    _same(gpu, ref)
    assert gpu[0]['iterations'] == 1 and gpu[0]['restarts'] == 0


@pytest.mark.parametrize('precision', ['fp64', 'mixed'])
def test_servers_record_failed_steps_and_go_on(precision):
    """A cap of 3 fails every step: both servers record the failures, free the slot and keep the design."""
    print(f'certify: servers with cap 3 ({precision})', flush=True)
    spec = (16, 1e-6, .5, 21)
    arrivals = [(0.0, 0), (0.001, 0)]
    mk = lambda: [Client(0, *spec, steps=2)]
    clients = mk()
    x0 = clients[0].x.clone()
    seq = SequentialServer((16,), cap=3)
    assert seq.run(clients, arrivals) == []
    assert len(seq.failed) == 2 and clients[0].failed == 2 and clients[0].done == 0
    assert torch.equal(clients[0].x, x0)
    clients = mk()
    srv = Server((16,), 2 ** 30, max_slots=2, precision=precision, cap=3, max_outer=2)
    assert srv.run(clients, arrivals) == []
    assert len(srv.failed) == 2 and clients[0].failed == 2 and clients[0].done == 0
    assert clients[0].slot is None and sorted(srv.free[16]) == [0, 1] and not srv.solving[16]
    assert torch.equal(clients[0].x, x0)
