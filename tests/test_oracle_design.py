"""Pins the oracle's C^H, sensitivities, filters and OC updates (CPU)."""
import numpy as np
import pytest
from scipy.ndimage import uniform_filter

from oracle import design, green, hex8, pcg

LAM, MU = 15 / 26, 5 / 13
C0 = hex8.isotropic(LAM, MU)


def solve_all(a, alpha=.5, rtol=1e-12, batch=6):
    s = green.scale(a, alpha, C0)
    solver = pcg.ResidentPCG(a, C0, np.eye(6), batch=batch, s=s)
    solver.solve(rtol=rtol)
    return solver.solutions


def test_solid_effective_tensor_is_c0():
    n = 4
    a = np.ones((n, n, n))
    u = solve_all(a)
    H = design.effective_tensor(a, u, C0)
    assert np.abs(H - C0).max() < 1e-11


def laminate_tensor(layers):
    """Exact laminate C^H for layers normal to x with C = a C0 (standard lamination formula)."""
    n_idx = [0, 4, 5]   # Voigt components carrying the traction on x-normal planes
    t_idx = [1, 2, 3]
    Cs = [a * C0 for a in layers]
    inv_nn = [np.linalg.inv(C[np.ix_(n_idx, n_idx)]) for C in Cs]
    A = np.linalg.inv(np.mean(inv_nn, axis=0))
    H = np.zeros((6, 6))
    mean = lambda f: np.mean([f(C, i) for C, i in zip(Cs, inv_nn)], axis=0)
    Cnt = A @ mean(lambda C, i: i @ C[np.ix_(n_idx, t_idx)])
    H[np.ix_(n_idx, n_idx)] = A
    H[np.ix_(n_idx, t_idx)] = Cnt
    H[np.ix_(t_idx, n_idx)] = Cnt.T
    tt = mean(lambda C, i: C[np.ix_(t_idx, t_idx)] - C[np.ix_(t_idx, n_idx)] @ i @ C[np.ix_(n_idx, t_idx)])
    H[np.ix_(t_idx, t_idx)] = tt + mean(lambda C, i: C[np.ix_(t_idx, n_idx)] @ i) @ Cnt
    return H


def test_laminate_effective_tensor_matches_closed_form():
    n = 6
    layers = np.array([1.0, .3, .05, 1.0, .6, .01])
    a = np.broadcast_to(layers[None, None, :], (n, n, n)).copy()
    H = design.effective_tensor(a, solve_all(a), C0)
    ref = laminate_tensor(layers)
    assert np.abs(H - ref).max() < 1e-10 * np.abs(ref).max()
    assert np.abs(H - H.T).max() < 1e-13


def test_random_effective_tensor_bounds():
    n = 4
    a = design.simp(design.filtered_noise(n, 11, 1), 1e-4)
    H = design.effective_tensor(a, solve_all(a), C0)
    assert np.abs(H - H.T).max() < 1e-12
    assert np.linalg.eigvalsh(H).min() > 0
    voigt = a.mean() * C0
    reuss = np.linalg.inv(np.mean([np.linalg.inv(v * C0) for v in a.ravel()], axis=0))
    assert np.linalg.eigvalsh(voigt - H).min() > -1e-12
    assert np.linalg.eigvalsh(H - reuss).min() > -1e-12


def test_filtered_noise_definition():
    rho = design.filtered_noise(8, 20261104, 2)
    noise = np.random.default_rng(20261104).random((8, 8, 8))
    f = uniform_filter(noise, size=5, mode='wrap')
    ref = 1e-3 + (1 - 1e-3) * (f - f.min()) / (f.max() - f.min())
    assert np.abs(rho - ref).max() < 1e-13
    assert rho.min() == pytest.approx(1e-3, abs=1e-15) and rho.max() == pytest.approx(1.0, abs=1e-15)


def filter_matrix(n, rmin):
    """All-pairs clipped cone matrix (``exp/asg_sctop_reproduction_20261105/verify.py:12-17``)."""
    xyz = np.array([[x, y, z] for z in range(n) for y in range(n) for x in range(n)])
    H = np.maximum(0, rmin - np.linalg.norm(xyz[:, None, :] - xyz[None, :, :], axis=-1))
    return H, H.sum(1)


def test_cone_filter_and_oc_match_dense_filter_matrix():
    n, rmin, emin, volume = 4, 2.1, 1e-9, .2
    oc = design.PaperOC(n, volume, rmin=rmin, emin=emin, move=.1)
    H, hs = filter_matrix(n, rmin)
    assert np.abs(oc.hs.ravel() - hs).max() < 1e-12
    rng = np.random.default_rng(51)
    x = rng.uniform(.15, .7, (n, n, n))
    ph, dp, a = oc.physical(x)
    t = (H @ x.ravel()) / hs
    den = np.tanh(2.) + np.tanh(2.)
    q = np.tanh(4 * (t - .5))
    assert np.abs(ph.ravel() - (np.tanh(2.) + q) / den).max() < 1e-12
    e = -rng.uniform(.1, 1, (n, n, n))
    new, count, g, dv = oc.update(x, ph, dp, e)
    df = (H @ ((dp * 3 * (1 - emin) * ph ** 2 * e).ravel() / hs) / n ** 3).reshape(x.shape)
    dg = (H @ (dp.ravel() / hs)).reshape(x.shape)
    assert np.abs(g - df).max() < 1e-12 * np.abs(df).max()
    lo, hi, cnt = 0., 1e9, 0
    while (hi - lo) / (hi + lo) > 1e-4 and hi > 1e-40:
        mid = .5 * (lo + hi)
        cand = np.maximum(.001, np.maximum(x - .1, np.minimum(1, np.minimum(x + .1, x * np.maximum(1e-10, -df / dg / mid) ** .3))))
        rho = (np.tanh(2.) + np.tanh(4 * ((H @ cand.ravel()) / hs - .5))) / den
        lo, hi = (mid, hi) if rho.mean() > volume else (lo, mid)
        cnt += 1
    assert cnt == count
    assert np.abs(new - cand).max() < 1e-12


def test_sensitivity_matches_finite_difference():
    n, emin = 4, 1e-9
    oc = design.PaperOC(n, .2, rmin=2.1, emin=emin)
    rng = np.random.default_rng(51)
    x = rng.uniform(.15, .7, (n, n, n))

    def objective(xx):
        ph, dp, a = oc.physical(xx)
        u = solve_all(a, rtol=1e-13)
        J, local, _ = design.shear_objective(a, u, C0)
        return J, ph, dp, local

    J, ph, dp, local = objective(x)
    _, _, g, _ = oc.update(x, ph, dp, local)
    direction = rng.normal(size=x.shape)
    direction /= np.linalg.norm(direction)
    h = 1e-5
    fd = (objective(x + h * direction)[0] - objective(x - h * direction)[0]) / (2 * h)
    analytic = float((g * direction).sum())
    assert abs(fd - analytic) / max(abs(fd), abs(analytic)) < 1e-5


def test_gaussian_oc_filter_and_bisection():
    n = 8
    oc = design.GaussianOC(n)
    radius = int(4 * oc.sigma + .5)
    w = design.gaussian_weights(oc.sigma, radius)
    delta = np.zeros((n, n, n))
    delta[0, 0, 0] = 1
    out = oc.filter(delta)
    full = np.zeros(n)
    for k, wk in zip(range(-radius, radius + 1), w):
        full[k % n] += wk
    assert np.abs(out - full[:, None, None] * full[None, :, None] * full[None, None, :]).max() < 1e-15
    assert np.abs(oc.filter(np.full((n, n, n), .3)) - .3).max() < 1e-15
    rng = np.random.default_rng(739)
    x = rng.uniform(.2, .4, (n, n, n))
    ph, dp, a = oc.physical(x, 2.)
    new, g, dv = oc.update(x, ph, dp, rng.uniform(.01, 2, (n, n, n)), 2.)
    assert np.all(new >= np.maximum(0, x - .1) - 1e-15) and np.all(new <= np.minimum(1, x + .1) + 1e-15)
    assert abs(oc.physical(new, 2.)[0].mean() - .3) < 1e-6
