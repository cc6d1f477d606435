"""Resident multi-load PCG on the GPU vs the oracle recurrence (GPU)."""
import numpy as np
import pytest
import torch

import pgj
from oracle import design, green, hex8, pcg

pytestmark = pytest.mark.gpu
LAM, MU = 15 / 26, 5 / 13
C0 = hex8.isotropic(LAM, MU)


def density(n, emin, radius, seed=20261104):
    return design.simp(design.filtered_noise(n, seed, radius), emin)


@pytest.mark.parametrize('n,emin,alpha,mode', [(6, 1e-4, .5, 'serial'), (8, 1e-4, .5, 'parallel'),
                                               (8, 1e-6, 0., 'parallel'), (16, 1e-4, .5, 'serial'),
                                               (16, 1e-6, 1., 'parallel'), (24, 1e-4, .5, 'parallel'),
                                               (32, 1e-6, .5, 'parallel')])
def test_resident_matches_oracle(n, emin, alpha, mode):
    a = density(n, emin, 2 if n >= 8 else 1)
    r = pgj.Resident(n, load_mode=mode).capture()
    r.material(torch.tensor(a, device='cuda'))
    r.select(str(alpha))
    infos, checks = r.solve(rtol=1e-10)
    s = None if alpha == 0 else green.scale(a, alpha, C0)
    ref = pcg.ResidentPCG(a, C0, np.eye(6), batch=1 if mode == 'serial' else 3, s=s)
    rinfo, _ = ref.solve(rtol=1e-10)
    for gi, oi in zip(infos, rinfo):
        assert gi['converged'] and gi['relative_residual'] <= 1e-10
        assert abs(gi['iterations'] - oi['iterations']) <= 1, (gi, oi)
        assert gi['restarts'] == oi['restarts']
    u = r.solutions.cpu().numpy()
    # the certified residual, recomputed independently on the CPU
    for l in range(6):
        f = hex8.rhs(a, np.eye(6)[l], C0)
        res = np.linalg.norm(f - hex8.apply(a, u[l:l + 1], C0)) / np.linalg.norm(f)
        assert res <= 1.0001e-10, (l, res)
    H = r.effective_tensor().cpu().numpy()
    Href = design.effective_tensor(a, ref.solutions, C0)
    assert np.abs(H - Href).max() / np.abs(Href).max() < 1e-10


def test_zero_rhs_and_warm_start():
    n = 8
    r = pgj.Resident(n, load_mode='parallel', batch=6).capture()
    r.material(torch.full((n, n, n), .13, dtype=torch.float64, device='cuda'))
    r.select('0.5')
    infos, checks = r.solve()
    # one group of six loads: a single host check finds no active load
    assert all(i['iterations'] == 0 and i['converged'] for i in infos) and checks == 1
    a = density(n, 1e-4, 2)
    r.material(torch.tensor(a, device='cuda'))
    r.select('0.5')
    infos, _ = r.solve()
    first = [i['iterations'] for i in infos]
    infos2, _ = r.solve()   # warm start from the certified solutions
    assert all(i['iterations'] <= f for i, f in zip(infos2, first))
    H = r.effective_tensor().cpu().numpy()
    assert np.abs(H - H.T).max() < 1e-12 * np.abs(H).max()


def test_solid_effective_tensor_is_c0():
    n = 4
    r = pgj.Resident(n, load_mode='parallel', batch=6).capture()
    r.material(torch.ones((n, n, n), dtype=torch.float64, device='cuda'))
    r.select('0.5')
    r.solve(1e-11)
    H = r.effective_tensor().cpu().numpy()
    assert np.abs(H - C0).max() < 1e-11


def test_resident_64_certifies():
    """64^3 workload (Emin 1e-6, radius-4 noise): every load certified, residual rechecked on the CPU."""
    n = 64
    a = density(n, 1e-6, 4)
    r = pgj.Resident(n, load_mode='parallel', batch=6).capture()
    r.material(torch.tensor(a, device='cuda'))
    r.select('0.5')
    infos, _ = r.solve()
    assert all(i['converged'] for i in infos)
    u = r.solutions.cpu().numpy()
    for l in (0, 3):
        f = hex8.rhs(a, np.eye(6)[l], C0)
        res = np.linalg.norm(f - hex8.apply(a, u[l:l + 1], C0)) / np.linalg.norm(f)
        assert res <= 1.0001e-8, (l, res)
    H = r.effective_tensor().cpu().numpy()
    assert np.abs(H - H.T).max() < 1e-10 * np.abs(H).max()
    assert np.linalg.eigvalsh(H).min() > 0


def test_resident_64_z_chunk_widths_agree():
    """A 6-load 64^3 FP64 workspace runs the z pass on 17-column kx chunks (its 11-column blocks would need more
    than one round on 256 CUs); a 1-load workspace keeps 11-column chunks. One 6-load launch per stage, the two 3-load
    halves and six 1-load solves certify with the same iteration counts (+-2: the r.z partials regroup) and C^H."""
    from pgj import _lib
    n = 64
    grid = _lib.make_grid((n, n, n), (1 / n,) * 3, LAM, MU)
    assert _lib.lib().pgj_zpass_wide(grid, 6, 8) == 1     # 1152 narrow blocks > 1024 slots of 256 CUs
    assert _lib.lib().pgj_zpass_wide(grid, 1, 8) == 0
    assert _lib.lib().pgj_zpass_wide(grid, 6, 4) == 0     # the FP32 pass has no wide instance
    a = torch.tensor(density(n, 1e-6, 4), device='cuda')
    runs = []
    for kw in (dict(load_mode='parallel', batch=6, streams=1), dict(load_mode='parallel', batch=6, streams=2),
               dict(load_mode='serial')):
        r = pgj.Resident(n, **kw).capture()
        r.material(a)
        r.select('0.5')
        infos, _ = r.solve()
        assert all(i['converged'] and i['relative_residual'] <= 1e-8 for i in infos), kw
        runs.append((np.array([i['iterations'] for i in infos]), r.effective_tensor().cpu().numpy()))
    its, H = runs[2]
    for other_its, other_H in runs[:2]:
        assert np.abs(other_its - its).max() <= 2, (other_its, its)
        assert np.abs(other_H - H).max() <= 1e-9 * np.abs(H).max()


def test_resident_two_x_per_lane_matches_oracle():
    """nx = 96 (two x per lane in the row-march PCG kernel, runtime FFT) vs the oracle recurrence."""
    shape = (6, 8, 96)
    a = design.simp(np.random.default_rng(3).uniform(.05, 1, shape), 1e-4)
    r = pgj.Resident(shape, load_mode='parallel', batch=3).capture()
    r.material(torch.tensor(a, device='cuda'))
    r.select('0.5')
    infos, _ = r.solve(rtol=1e-10)
    ref = pcg.ResidentPCG(a, C0, np.eye(6), batch=3, s=green.scale(a, .5, C0))
    rinfo, _ = ref.solve(rtol=1e-10)
    for gi, oi in zip(infos, rinfo):
        assert gi['converged'] and abs(gi['iterations'] - oi['iterations']) <= 1, (gi, oi)
    H = r.effective_tensor().cpu().numpy()
    Href = design.effective_tensor(a, ref.solutions, C0)
    assert np.abs(H - Href).max() / np.abs(Href).max() < 1e-10


@pytest.mark.parametrize('n', [96, 128])
def test_resident_large_certifies(n):
    """Compile-time FFT and two-x-per-lane HEX8 at 96^3 / 128^3: one load group certified, rechecked on the CPU."""
    print(f'resident large n={n}', flush=True)
    a = density(n, 1e-4, 4)
    r = pgj.Resident(n, loads=np.eye(6)[[0, 3]], load_mode='parallel', batch=2).capture()
    r.material(torch.tensor(a, device='cuda'))
    r.select('0.5')
    infos, _ = r.solve()
    assert all(i['converged'] for i in infos)
    u = r.solutions.cpu().numpy()
    for l, m in enumerate((0, 3)):
        f = hex8.rhs(a, np.eye(6)[m], C0)
        res = np.linalg.norm(f - hex8.apply(a, u[l:l + 1], C0)) / np.linalg.norm(f)
        assert res <= 1.0001e-8, (n, m, res)


@pytest.mark.parametrize('n', [16, 32])
@pytest.mark.parametrize('plane', [0, 1], ids=['node', 'plane'])
def test_row_march_variants_in_pcg(n, plane):
    """The resident iteration with either HEX8 row-march variant matches the oracle recurrence."""
    from pgj import _lib
    a = density(n, 1e-4, 2)
    r = pgj.Resident(n, load_mode='parallel', batch=6)
    lib = _lib.lib()
    _lib.check(lib.pgj_hex8_set_shape(r.op.grid, 6, 8, 1, 4, 4, plane))
    try:
        r.capture()
        r.material(torch.tensor(a, device='cuda'))
        r.select('0.5')
        infos, _ = r.solve(rtol=1e-10)
        ref = pcg.ResidentPCG(a, C0, np.eye(6), batch=6, s=green.scale(a, .5, C0))
        rinfo, _ = ref.solve(rtol=1e-10)
        for gi, oi in zip(infos, rinfo):
            assert gi['converged'] and abs(gi['iterations'] - oi['iterations']) <= 1
        H = r.effective_tensor().cpu().numpy()
        assert np.abs(H - design.effective_tensor(a, ref.solutions, C0)).max() / np.abs(H).max() < 1e-10
    finally:
        _lib.check(lib.pgj_hex8_set_shape(r.op.grid, 6, 8, 0, 0, 0, 0))


def test_64_wide_and_narrow_preconditioner_match_oracle():
    """The first direction p0 = P s G s P r of a 6-load set (wide z chunks) and of a 1-load set (narrow) against
    the oracle's preconditioner on the same residuals."""
    from pgj import _lib
    n = 64
    a = density(n, 1e-6, 4)
    rng = np.random.default_rng(17)
    r6 = rng.standard_normal((6, 3 * n ** 3))
    s = green.scale(a, .5, C0)
    G6 = green.symbol((n, n, n), hex8.voxel_of((n, n, n)), LAM, MU)
    ref = green.precond(r6, G6, (n, n, n), s)
    for kw, loads in ((dict(load_mode='parallel', batch=6, streams=1), 6), (dict(load_mode='serial'), 1)):
        res = pgj.Resident(n, **kw)
        assert _lib.lib().pgj_zpass_wide(res.op.grid, res.batch, 8) == (1 if loads == 6 else 0)
        res.material(torch.tensor(a, device='cuda'))
        res.select('0.5')
        res._setup(range(res.batch), 1e-8)
        res.r.copy_(res.to_soa(torch.tensor(r6[:loads], device='cuda')))
        res.initialize()
        p0 = res.to_aos(res.p0).cpu().numpy()
        err = np.abs(p0 - ref[:loads]).max() / np.abs(ref[:loads]).max()
        print(f'64^3 preconditioner, {loads} load(s): rel err {err:.2e}', flush=True)
        assert err < 1e-13


def test_64_batched_multi_round_matches_solo():
    """Two 64^3 designs in one batched set (12 loads: wide z chunks over more than one round of blocks) against
    each design solved alone in the serial load mode (narrow chunks)."""
    from pgj.batched import BatchedDesigns
    n = 64
    designs = [(density(n, 1e-4, 4, seed=41 + d), alpha) for d, alpha in enumerate((.25, .75))]
    bd = BatchedDesigns({n: 2}).capture()
    g = bd.groups[n]
    for d, (a, alpha) in enumerate(designs):
        g.material(d, torch.tensor(a, device='cuda'))
        g.select(d, alpha)
    infos, _ = bd.solve(1e-8)
    solo = pgj.Resident(n, load_mode='serial').capture()
    for d, (a, alpha) in enumerate(designs):
        solo.solutions.zero_()
        solo.material(torch.tensor(a, device='cuda'))
        solo.select(alpha)
        sinfo, _ = solo.solve(1e-8)
        Hs = solo.effective_tensor().cpu().numpy()
        Hb = g.effective_tensor(d).cpu().numpy()
        assert np.abs(Hb - Hs).max() <= 1e-10 * np.abs(Hs).max()
        bits = [i['iterations'] for i in infos[n][6 * d:6 * d + 6]]
        assert max(abs(x - y['iterations']) for x, y in zip(bits, sinfo)) <= 2
