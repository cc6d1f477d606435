"""Replay of a saved design state (GPU): a certified solution stored in the reference's .npz format warm-starts
a fresh resident solver, which certifies it without iterating and reproduces C^H."""
import numpy as np
import pytest
import torch

import pgj
from pgj import checkpoint
from oracle import design

pytestmark = pytest.mark.gpu


def test_replay_of_a_certified_state(tmp_path):
    n = 16
    rho = design.filtered_noise(n, 11, 2)
    a = design.simp(rho, 1e-4)
    res = pgj.Resident(n, load_mode='parallel', batch=6).capture()
    res.material(torch.tensor(a, device='cuda'))
    res.select(.5)
    infos, _ = res.solve(rtol=1e-8)
    assert all(i['converged'] for i in infos)
    H = res.effective_tensor().cpu().numpy()
    path = tmp_path / 'state_012.npz'
    checkpoint.save_state(path, x=rho, a=a, physical=rho, beta=4.0, previous=np.zeros((6, 3 * n ** 3)),
                          solution=res.solutions, energy=np.zeros((n, n, n)), J=float(H[3, 3]))
    state = checkpoint.load_state(path)
    state.pop('previous')                       # warm start from the certified solution itself
    fresh = pgj.Resident(n, load_mode='parallel', batch=6).capture()
    again = checkpoint.replay(fresh, state, .5)
    assert all(i['converged'] and i['iterations'] == 0 for i in again)
    np.testing.assert_allclose(fresh.effective_tensor().cpu().numpy(), H, rtol=0, atol=1e-14 * np.abs(H).max())
    with pytest.raises(ValueError):                 # a three-load warm start does not fit a six-load solver
        checkpoint.replay(fresh, dict(a=a, previous=np.zeros((3, 3 * n ** 3))), .5)
