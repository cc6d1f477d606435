"""Design-state checkpoints in the reference's .npz format: round trip, aliases, validation (CPU)."""
import numpy as np
import pytest

from pgj import checkpoint


def state(n=4, L=6, seed=0):
    rng = np.random.default_rng(seed)
    return dict(x=rng.random((n, n, n)), a=rng.random((n, n, n)), physical=rng.random((n, n, n)), beta=4.0,
                previous=rng.standard_normal((L, 3 * n ** 3)), solution=rng.standard_normal((L, 3 * n ** 3)),
                energy=rng.random((n, n, n)), J=-0.25)


def test_round_trip(tmp_path):
    s = state()
    path = tmp_path / 'state_000.npz'
    checkpoint.save_state(path, **s)
    back = checkpoint.load_state(path)
    assert set(back) == set(s)
    for k, v in s.items():
        np.testing.assert_array_equal(back[k], v)
    with np.load(path, allow_pickle=False) as z:        # the reference's reader (docs/DATA.md:23)
        assert sorted(z.files) == sorted(s)


def test_older_key_names(tmp_path):
    s = state()
    path = tmp_path / 'old.npz'
    np.savez(path, x=s['x'], a=s['a'], rho=s['physical'], u0=s['previous'], beta=s['beta'])
    back = checkpoint.load_state(path)
    np.testing.assert_array_equal(back['physical'], s['physical'])
    np.testing.assert_array_equal(back['previous'], s['previous'])
    checkpoint.save_state(tmp_path / 'new.npz', rho=s['physical'], u0=s['previous'])
    assert sorted(checkpoint.load_state(tmp_path / 'new.npz')) == ['physical', 'previous']
    # the reference's Study-II readers index the old names (examples/replay_state.py:43-45)
    with np.load(tmp_path / 'new.npz', allow_pickle=False) as z:
        np.testing.assert_array_equal(z['rho'], s['physical'])
        np.testing.assert_array_equal(z['u0'], s['previous'])
    with pytest.raises(ValueError):
        checkpoint.save_state(tmp_path / 'dup.npz', rho=s['physical'], physical=s['physical'])


@pytest.mark.parametrize('bad', [dict(a=np.zeros((4, 4, 5))), dict(a=np.zeros((4, 4, 4)), x=np.zeros((5, 5, 5))),
                                 dict(a=np.zeros((4, 4, 4)), previous=np.zeros((6, 3 * 4 ** 3 + 3))),
                                 dict(solution=np.zeros(12))])
def test_shapes_are_checked(tmp_path, bad):
    with pytest.raises(ValueError):
        checkpoint.save_state(tmp_path / 'bad.npz', **bad)


def test_object_arrays_are_refused(tmp_path):
    with pytest.raises(ValueError):
        checkpoint.save_state(tmp_path / 'obj.npz', note=np.array([{'k': 1}], dtype=object))
