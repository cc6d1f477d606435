"""Hand-written FFT Green preconditioner vs the oracle (GPU)."""
import numpy as np
import pytest
import torch

import pgj
from oracle import design, green, hex8

pytestmark = pytest.mark.gpu
LAM, MU = 15 / 26, 5 / 13
C0 = hex8.isotropic(LAM, MU)

GRIDS = [(4, 4, 4), (6, 6, 6), (8, 8, 8), (16, 16, 16), (3, 4, 5), (5, 5, 5), (6, 4, 8), (12, 12, 12),
         (24, 24, 24), (32, 32, 32), (7, 5, 9)]


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


def test_symbol_matches_oracle():
    for grid in [(4, 4, 4), (3, 4, 5), (8, 6, 10)]:
        op = pgj.build(torch.ones(grid, dtype=torch.float64, device='cuda'))
        got = op._contest_coefficient.cpu().numpy().reshape(6, *grid[:2], grid[2] // 2 + 1)
        ref = green.symbol(grid, hex8.voxel_of(grid), LAM, MU)
        assert np.abs(got - ref).max() < 1e-14 * np.abs(ref).max(), grid


@pytest.mark.parametrize('grid', GRIDS, ids=[str(g) for g in GRIDS])
@pytest.mark.parametrize('alpha', [0.0, 0.5, 1.0])
def test_precond_matches_oracle(grid, alpha):
    rng = np.random.default_rng(20261024)
    a = np.exp(rng.uniform(-6, 0, grid))
    op = pgj.Operator(torch.tensor(a, device='cuda')).enable_green()
    pgj.set_alpha(op, alpha)
    L = 3
    r = rng.normal(size=(L, 3 * a.size))
    z = op.apply_precond(torch.tensor(r, device='cuda')).cpu().numpy()
    s = None if alpha == 0 else green.scale(a, alpha, C0)
    ref = green.precond(r, green.symbol(grid, hex8.voxel_of(grid), LAM, MU), grid, s)
    assert rel(z, ref) < 1e-13
    one = op.apply_precond(torch.tensor(r[2:3], device='cuda')).cpu().numpy()
    # a load alone and in a batch: the block shape follows the load count (CU coverage), so the partial
    # sums of the projection regroup; the results agree to rounding
    assert np.abs(one[0] - z[2]).max() <= 1e-14 * np.abs(z[2]).max()


def test_precond_properties_at_64():
    """64^3: symmetry, projection (zero mean), agreement with the oracle on one load."""
    n = 64
    a = design.simp(design.filtered_noise(n, 20261104, 4), 1e-6)
    op = pgj.Operator(torch.tensor(a, device='cuda')).enable_green()
    pgj.set_alpha(op, .5)
    g = torch.Generator(device='cuda').manual_seed(7)
    u = torch.randn((2, 3 * n ** 3), generator=g, device='cuda', dtype=torch.float64)
    Mu = op.apply_precond(u)
    sym = float(((u[0] * Mu[1]).sum() - (u[1] * Mu[0]).sum()).abs() / (Mu.norm() * u.norm()))
    assert sym < 1e-14
    assert float(Mu.reshape(2, -1, 3).mean(1).abs().max()) < 1e-15 * float(Mu.abs().max())
    s = green.scale(a, .5, C0)
    ref = green.precond(u[:1].cpu().numpy(), green.symbol((n,) * 3, (1 / n,) * 3, LAM, MU), (n,) * 3, s)
    assert rel(Mu[:1].cpu().numpy(), ref) < 1e-13


@pytest.mark.parametrize('n', [16, 24, 32, 48, 64, 96, 128])
def test_static_kernels_match_runtime(n):
    """The resident's FFT passes (SoA state) agree with the generic preconditioner apply on AoS vectors, which
    runs the same compile-time passes on SoA copies; and with the oracle to 48^3 and at 128^3."""
    print(f'static vs runtime n={n}', flush=True)
    a = design.simp(design.filtered_noise(n, 5 + n, 2), 1e-4)
    res = pgj.Resident(n, load_mode='parallel', batch=3)
    res.material(torch.tensor(a, device='cuda'))
    res.select(0.5)
    s_aos = res.to_aos(res.scale)
    rng = np.random.default_rng(n)
    r = torch.tensor(rng.standard_normal((3, 3 * n ** 3)), device='cuda')
    res._setup(range(3), 1e-8)
    res.r.copy_(res.to_soa(r))
    res.initialize()
    z = res.to_aos(res.p0).cpu().numpy()   # the first direction is M^-1 r
    shat_rt = res.op.scale_spectrum(s_aos)
    assert rel(res.scale_hat.cpu().numpy().reshape(-1), shat_rt.cpu().numpy().reshape(-1)) < 1e-13
    ref = res.op.green_precond(r, s_aos, shat_rt).cpu().numpy()
    assert rel(z, ref) < 1e-12
    if n <= 48 or n == 128:
        G6 = green.symbol(a.shape, hex8.voxel_of(a.shape), LAM, MU)
        assert rel(z, green.precond(r.cpu().numpy(), G6, a.shape, green.scale(a, 0.5, C0))) < 1e-12


@pytest.mark.parametrize('grid', [(8, 8, 8), (5, 6, 7), (4, 6, 9)], ids=str)
def test_factored_symbol_and_protocol_attributes(grid):
    """The (D, beta) form and v tables rebuild the plan's dense symbol; the C21 rows a C0 give the same K."""
    from pgj.hybrid import HybridFIOperator
    rng = np.random.default_rng(3)
    a = rng.uniform(0.2, 1.0, grid)
    op = pgj.build(torch.tensor(a, device='cuda'))
    D, beta = op._fft_symbol_inv[:3].cpu().numpy(), op._fft_symbol_inv[3].cpu().numpy()
    vx, vy, vz = (t.cpu().numpy() for t in op._formula_vtables)
    nz, ny, nx = grid
    assert vx.shape == (nx // 2 + 1,) and vy.shape == (ny,) and vz.shape == (nz,)
    v = np.stack(np.broadcast_arrays(vx[None, None, :], vy[None, :, None], vz[:, None, None])).reshape(3, -1)
    w = D * v
    pairs = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
    G = np.stack([(D[i] if i == j else 0.0) - beta * w[i] * w[j] for i, j in pairs])
    sym = op._contest_coefficient.cpu().numpy() * a.size
    assert np.abs(G - sym).max() <= 1e-13 * np.abs(sym).max()
    assert op._contest_plan is op._plan and op._hybrid_indirect is False
    hop = HybridFIOperator(op._hybrid_coeff.contiguous(), op._hybrid_index, grid, op._voxel, dense=True)
    u = torch.tensor(rng.normal(size=(1, 3 * a.size)), device='cuda')
    ref = op.apply(u).cpu().numpy()
    assert np.abs(hop.apply(u).cpu().numpy() - ref).max() <= 1e-13 * np.abs(ref).max()


def test_set_alpha_attaches_the_centred_log_diagonal():
    rng = np.random.default_rng(8)
    a = rng.uniform(0.05, 1.0, (6, 6, 6))
    op = pgj.build(torch.tensor(a, device='cuda'))
    pgj.set_alpha(op, 0.5)
    d = hex8.diagonal(a, C0)
    q = np.log(d) - np.log(d).mean()
    assert np.abs(op._log_diagonal.cpu().numpy() - q).max() < 1e-13
    s = np.exp(-0.25 * q)
    assert np.abs(op._scale.cpu().numpy() - s).max() < 1e-13


@pytest.mark.parametrize('n', [16, 24, 32, 48, 96, 128])
def test_precond_static_cubes_match_oracle(n):
    """Cube grids run the generic preconditioner apply through the compile-time FFT passes, which read r and s
    and write z in the (L, 3N) node-interleaved layout."""
    rng = np.random.default_rng(n)
    a = np.exp(rng.uniform(-5, 0, (n, n, n)))
    op = pgj.Operator(torch.tensor(a, device='cuda')).enable_green()
    pgj.set_alpha(op, 0.5)
    r = rng.normal(size=(2, 3 * n ** 3))
    z = op.apply_precond(torch.tensor(r, device='cuda')).cpu().numpy()
    grid = (n, n, n)
    ref = green.precond(r, green.symbol(grid, hex8.voxel_of(grid), LAM, MU), grid, green.scale(a, 0.5, C0))
    assert rel(z, ref) < 1e-12
    one = op.apply_precond(torch.tensor(r[1:2], device='cuda')).cpu().numpy()
    assert np.array_equal(one[0], z[1])   # the compile-time passes have one block shape for any load count


@pytest.mark.parametrize('n', [32, 48, 64, 96])
def test_mixed_apply_scaled_matches_oracle(n):
    """pgj_precond_apply_mixed (FP32 arithmetic on FP64 vectors) with the FP32-rounded scale and its spectrum,
    two loads: within FP32 rounding of the oracle's FP64 M_alpha (2-norm relative 1e-5, and above 1e-10, so the
    FP32 passes ran)."""
    from pgj import _lib
    rng = np.random.default_rng(n + 5)
    a = np.exp(rng.uniform(-5, 0, (n, n, n)))
    op = pgj.Operator(torch.tensor(a, device='cuda')).enable_green()
    pgj.set_alpha(op, 0.5)
    lib = _lib.lib()
    plan32 = torch.empty(lib.pgj_plan32_bytes(op.grid), dtype=torch.uint8, device='cuda')
    _lib.call('pgj_plan32_init', op.grid, _lib.ptr(op._plan), _lib.ptr(plan32), _lib.stream())
    s32 = op._scale.float().contiguous()
    shat32 = torch.empty(2 * op._scale_hat.numel(), dtype=torch.float32, device='cuda')
    _lib.call('pgj_to_f32', 2 * op._scale_hat.numel(), _lib.ptr(op._scale_hat), _lib.ptr(shat32), _lib.stream())
    r = rng.normal(size=(2, 3 * n ** 3))
    rt = torch.tensor(r, device='cuda')
    z = torch.empty_like(rt)
    ws = torch.empty(lib.pgj_precond_ws_bytes(op.grid, 2), dtype=torch.uint8, device='cuda')
    _lib.call('pgj_precond_apply_mixed', op.grid, 2, _lib.ptr(plan32), _lib.ptr(rt), _lib.ptr(s32), _lib.ptr(shat32),
              _lib.ptr(z), _lib.ptr(ws), _lib.stream())
    grid = (n, n, n)
    ref = green.precond(r, green.symbol(grid, hex8.voxel_of(grid), LAM, MU), grid, green.scale(a, 0.5, C0))
    err = rel(z.cpu().numpy(), ref)
    assert 1e-10 < err < 1e-5, err
    with pytest.raises(ValueError):
        odd = pgj.Operator(torch.ones((6, 8, 10), dtype=torch.float64, device='cuda')).enable_green()
        _lib.call('pgj_precond_apply_mixed', odd.grid, 1, _lib.ptr(plan32), _lib.ptr(rt), None, None, _lib.ptr(z),
                  _lib.ptr(ws), _lib.stream())
