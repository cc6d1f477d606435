"""Launch-shape autotuner: HEX8 row-march shapes and FFT radix plans, timed on a set's own buffers.

Counterpart of the reference's variant and tile sweeps (``solver/src/hex8_modal/gpu_study.py:174-222``,
``modal_adapter.py:28-41``). Every candidate computes the same operator, so a choice changes only speed and
rounding. A choice is recorded in the library (``pgj_hex8_set_shape``, ``pgj_fft_set_plan``) and holds for later
launches and graph captures of that grid; tune before capturing.
"""
import torch

from . import _lib

STAGE_FXY, STAGE_FZ, STAGE_FXYI = 1, 2, 3


def time_launch(fn, reps):
    """Mean device milliseconds of ``fn()`` over ``reps`` back-to-back launches after one warm launch."""
    fn()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def row_march_candidates(n, es):
    """(R, W, ZC, plane) shapes; the FP64 plane-carry march runs at most 4 row waves per block."""
    return [(1, w, zc, p) for p in (0, 1) for w in (2, 4, 8) for zc in (1, 2, 4, 8, 16)
            if n % zc == 0 and not (p and w > 4 and es == 8)]


def _fastest(candidates, select, launch, reps, rounds):
    """{candidate: best mean ms} over ``rounds`` interleaved rounds (each round times every candidate once), so a
    transient slowdown of the device lands on one round of every candidate rather than on all of one."""
    times = {}
    for _ in range(rounds):
        for cand in candidates:
            select(cand)
            t = time_launch(launch, reps)
            times[cand] = min(times.get(cand, t), t)
    return times


def tune_row_march(grid, L, es, launch_hex8, reps=6, candidates=None, rounds=3):
    """Fastest row-march shape of ``launch_hex8()`` (a HEX8 stage over L loads); recorded and returned."""
    lib = _lib.lib()
    times = _fastest(list(candidates or row_march_candidates(grid.nx, es)),
                     lambda shape: _lib.check(lib.pgj_hex8_set_shape(grid, L, es, *shape)), launch_hex8, reps, rounds)
    shape = min(times, key=times.get)
    _lib.check(lib.pgj_hex8_set_shape(grid, L, es, *shape))
    return dict(shape=shape, ms=times[shape])


def tune_fft(grid, es, launch_stage, reps=6, rounds=3):
    """Radix plan (0: radix 4 first, 1: radix 8 first; for the z pass at 64^3 also 2: the transforms in
    registers) of each FFT pass, by timing ``launch_stage(k)`` for stages 1 (forward xy), 2 (z and Green
    multiply) and 3 (inverse xy) in interleaved rounds; recorded and returned."""
    lib = _lib.lib()
    out = {}
    for pas, stage in enumerate((STAGE_FXY, STAGE_FZ, STAGE_FXYI)):
        plans = (0, 1, 2) if pas == 1 and grid.nx == 64 else (0, 1)
        times = _fastest(plans, lambda plan: _lib.check(lib.pgj_fft_set_plan(grid, es, pas, plan)),
                         lambda: launch_stage(stage), reps, rounds)
        plan = min(times, key=times.get)
        _lib.check(lib.pgj_fft_set_plan(grid, es, pas, plan))
        out[('fxy', 'fz', 'fxyi')[pas]] = dict(plan=plan, ms={str(k): v for k, v in times.items()})
    return out


class heterogeneous:
    """Context that fills a material field with a fixed heterogeneous one (values in [0.05, 1]) and restores it.

    Tuning times stage launches with every load active. On a homogeneous material the Green preconditioner is
    exact and a load converges in one iteration, after which its launches exit at once and time nothing."""

    def __init__(self, field):
        self.field = field

    def __enter__(self):
        self.saved = self.field.clone()
        g = torch.Generator(device=self.field.device).manual_seed(20261104)
        self.field.copy_(0.05 + 0.95 * torch.rand(self.field.shape, generator=g, device=self.field.device,
                                                  dtype=self.field.dtype))
        return self.field

    def __exit__(self, *exc):
        self.field.copy_(self.saved)
        return False


def fft_plans(grid, es):
    """The radix plans in use for the three passes of a cube in precision es."""
    out = (_lib.c_int * 3)()
    _lib.check(_lib.lib().pgj_fft_get_plan(grid, es, out))
    return tuple(out)


def reset(grid, L=None, es=8):
    """Default shapes and plans again (row-march shape for L loads per launch, when given)."""
    lib = _lib.lib()
    for pas in range(3):
        _lib.check(lib.pgj_fft_set_plan(grid, es, pas, -1))
    if L is not None:
        _lib.check(lib.pgj_hex8_set_shape(grid, L, es, 0, 0, 0, 0))
