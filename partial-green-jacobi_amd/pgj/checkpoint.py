"""Design-state checkpoints in the reference's ``.npz`` format (replay and serving persistence).

Keys (``exp/asg_topopt_20261103/study.py:176-177``, ``docs/DATA.md:15-23``):
``x`` design density, ``a`` stiffness multiplier, ``physical`` (older
archives: ``rho``) filtered/projected density, all (n, n, n) in (z, y, x)
order; ``beta`` projection sharpness; ``previous`` (older: ``u0``) the warm
start and ``solution`` the certified displacements, load-major (L, 3 n^3)
with three components per node; ``energy`` and ``J``. Files are written
compressed and read with ``allow_pickle=False``; a missing key means the field
was not saved.
"""
import numpy as np
import torch

ALIASES = {'rho': 'physical', 'u0': 'previous'}
FIELDS = ('x', 'a', 'physical')
VECTORS = ('previous', 'solution')


def _host(v):
    return v.detach().cpu().numpy() if torch.is_tensor(v) else np.asarray(v)


def save_state(path, **fields):
    """Write the given fields (tensors or arrays) to ``path`` (``np.savez_compressed``) under the names given, so
    a Study-II state saved with ``rho``/``u0`` stays readable by the reference's own readers
    (``examples/replay_state.py:43-45`` read ``state['rho']`` and ``state['u0']``). A field given under both
    its names is refused."""
    out, canonical = {}, {}
    for key, value in fields.items():
        if value is None:
            continue
        name = ALIASES.get(key, key)
        if name in canonical:
            raise ValueError(f'{key!r} given twice (as {canonical[name]!r} and {key!r})')
        canonical[name] = key
        out[key] = _host(value)
        if out[key].dtype == object:
            raise ValueError(f'{key!r}: object arrays are not stored (the files are read without pickle)')
    check_state({ALIASES.get(k, k): v for k, v in out.items()})
    np.savez_compressed(path, **out)


def load_state(path):
    """The fields of a state file as numpy arrays under their current names (``rho`` -> ``physical``,
    ``u0`` -> ``previous``); shapes are checked."""
    with np.load(path, allow_pickle=False) as z:
        out = {}
        for key in z.files:
            name = ALIASES.get(key, key)
            if name in out:
                raise ValueError(f'{path}: both {key!r} and {name!r} present')
            out[name] = z[key]
    check_state(out)
    return out


def check_state(state):
    """Raise ``ValueError`` unless fields are (n, n, n) and vectors (L, 3 n^3) of one grid."""
    n = None
    for key in FIELDS:
        if key in state:
            shape = np.shape(state[key])
            if len(shape) != 3 or len(set(shape)) != 1:
                raise ValueError(f'{key!r} must be (n, n, n), got {shape}')
            if n is not None and shape[0] != n:
                raise ValueError(f'{key!r} is on a {shape[0]}^3 grid, expected {n}^3')
            n = shape[0]
    for key in VECTORS:
        if key in state:
            shape = np.shape(state[key])
            if len(shape) != 2 or (n is not None and shape[1] != 3 * n ** 3) or shape[1] % 3:
                raise ValueError(f'{key!r} must be (L, 3 n^3), got {shape}')


def replay(resident, state, alpha):
    """Solve the state's design on ``resident`` (a ``pgj.Resident`` of the state's grid) from its warm start
    (``previous``, else ``solution``, else zero); returns the solve's infos. ``alpha`` is passed to ``select``."""
    a = torch.as_tensor(state['a'], dtype=torch.float64, device=resident.op.device)
    resident.material(a)
    resident.select(alpha)
    start = state.get('previous', state.get('solution'))
    if start is None:
        resident.solutions.zero_()
    else:
        start = torch.as_tensor(start, dtype=torch.float64, device=resident.op.device)
        if start.shape != resident.solutions.shape:
            raise ValueError(f'warm start {tuple(start.shape)} does not fit {tuple(resident.solutions.shape)}')
        resident.solutions.copy_(start)
    infos, _ = resident.solve()
    return infos
