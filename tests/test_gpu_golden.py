"""GPU outputs vs the committed golden fixtures (no oracle run on the device host).

Operator and Green fixtures within 1e-13 / 1e-12 relative; the 16^3 plain
method (Emin 1e-4, box radius 2, alpha 0.5, serial, rtol 1e-8) within 1e-10 on
C^H and +-1 iteration per load; one cone OC update bit-for-bit in count and
within 1e-12 on the densities.
"""
import os

import numpy as np
import pytest
import torch

import pgj
from pgj.outer import PaperOuter

pytestmark = pytest.mark.gpu
HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')


def load(name):
    with np.load(os.path.join(HERE, name + '.npz'), allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def dev(x):
    return torch.tensor(np.ascontiguousarray(x), device='cuda')


def close(got, want, rel):
    got = got.cpu().numpy() if torch.is_tensor(got) else np.asarray(got)
    err = np.abs(got.reshape(want.shape) - want).max() / max(np.abs(want).max(), 1e-300)
    assert err <= rel, err


def test_operator_golden():
    g = load('operator_4x5x6')
    op = pgj.Operator(dev(g['a']))
    close(op.apply_batched(dev(g['u'])), g['Ku'], 1e-13)
    close(op.apply_reference(dev(g['u'])), g['Ku'], 1e-13)
    close(torch.cat([op.compute_rhs(m) for m in np.eye(6)]), g['rhs'], 1e-14)
    close(pgj.diagonal(op), g['diag'], 1e-14)
    stress = np.concatenate([op.compute_avg_stress(dev(g['u'][:1]), m) for m in np.eye(6)])
    close(stress, g['stress'], 1e-13)


def test_green_golden():
    g = load('green_3x4x5')
    op = pgj.build(dev(g['a']))
    close(op._contest_coefficient, g['symbol'].reshape(6, -1), 1e-14)
    close(op.apply_precond(dev(g['r'])), g['z_green'], 1e-12)
    pgj.set_alpha(op, .5)
    close(op._scale, g['s'], 1e-14)
    close(op.apply_precond(dev(g['r'])), g['z_half'], 1e-12)


@pytest.mark.parametrize('mode', ['serial', 'parallel'])
def test_plain_16_golden(mode):
    g = load('solve_16_alpha05')
    r = pgj.Resident(16, load_mode=mode).capture()
    r.material(dev(g['a']))
    r.select('0.5')
    infos, _ = r.solve(rtol=1e-8)
    for info, want in zip(infos, g['iterations']):
        assert info['converged'] and info['relative_residual'] <= 1e-8
        assert abs(info['iterations'] - int(want)) <= 1
    close(r.effective_tensor(), g['CH'], 1e-10)


def test_oc_golden():
    g = load('oc_cone_8')
    oc = PaperOuter(8, 0.35, emin=1e-6, move=0.2)
    oc.set_initial(g['x'])
    oc.evaluate()
    count = oc.step(dev(g['energy']))
    assert count == int(g['count'])
    close(oc.g, g['g'], 1e-12)
    close(oc.dv, g['dv'], 1e-12)
    assert np.abs(oc.new.cpu().numpy() - g['new']).max() < 1e-12
