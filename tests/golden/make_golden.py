"""Writes the small golden fixtures under tests/golden/ from the oracle.

Run from the repository root: ``python tests/golden/make_golden.py`` (the
small fixtures), or name fixtures, e.g. ``python tests/golden/make_golden.py
solve_64_alpha`` (the 64^3 solves take minutes of CPU). Every
input is seeded; outputs are FP64 arrays in compressed .npz files (no pickled
objects). The reference repository holds no numeric fixtures for this path, so
these record the oracle's outputs (the oracle itself is pinned by closed forms
and brute force in tests/test_oracle_*.py); the GPU tests compare against them
without running the oracle, and the CPU tests check that the oracle still
reproduces them.
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import design, green, hex8, pcg  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
LAM, MU = 15 / 26, 5 / 13
C0 = hex8.isotropic(LAM, MU)


def operator_fixture():
    """Non-cubic 4x5x6 grid: K u, rhs, diagonal, energies, mean stress."""
    rng = np.random.default_rng(101)
    grid = (4, 5, 6)
    a = 1e-3 + rng.random(grid)
    N = int(np.prod(grid))
    u = rng.standard_normal((2, 3 * N))
    macros = np.eye(6)
    voxel = hex8.voxel_of(grid)
    return dict(a=a, u=u, Ku=hex8.apply(a, u, C0, voxel),
                rhs=np.concatenate([hex8.rhs(a, m, C0, voxel) for m in macros]),
                diag=hex8.diagonal(a, C0, voxel),
                energies=hex8.energies(u[:1].repeat(6, 0), grid, macros, C0, voxel),
                stress=np.array([hex8.avg_stress(a, u[:1], m, C0, voxel) for m in macros]).reshape(6, 6))


def green_fixture():
    """Odd/even mixed 3x4x5 grid: dense symbol, P s G s P r at alpha 0.5 and P G P r."""
    rng = np.random.default_rng(202)
    grid = (3, 4, 5)
    a = 1e-2 + rng.random(grid)
    voxel = hex8.voxel_of(grid)
    G6 = green.symbol(grid, voxel, LAM, MU)
    r = rng.standard_normal((2, 3 * int(np.prod(grid))))
    s = green.scale(a, 0.5, C0, voxel)
    return dict(a=a, r=r, symbol=G6, s=s, z_half=green.precond(r, G6, grid, s), z_green=green.precond(r, G6, grid))


def solve_fixture():
    """Plain method at 16^3: Emin 1e-4, box radius 2, alpha 0.5, rtol 1e-8, zero start, serial loads."""
    n = 16
    rho = design.filtered_noise(n, 20261104, 2)
    a = design.simp(rho, 1e-4)
    s = green.scale(a, 0.5, C0)
    ref = pcg.ResidentPCG(a, C0, np.eye(6), batch=1, s=s)
    t = time.time()
    infos, checks = ref.solve(rtol=1e-8)
    print(f'16^3 solve: {time.time() - t:.1f} s, iterations {[i["iterations"] for i in infos]}')
    CH = design.effective_tensor(a, ref.solutions, C0)
    return dict(rho=rho, a=a, CH=CH, iterations=np.array([i['iterations'] for i in infos]),
                relative_residual=np.array([i['relative_residual'] for i in infos]),
                restarts=np.array([i['restarts'] for i in infos]), host_checks=np.array(checks))


def oc_fixture():
    """One cone-filter OC update at 8^3 (volume 0.35, move 0.2, Emin 1e-6) from seeded energies."""
    rng = np.random.default_rng(303)
    n = 8
    x = 0.2 + 0.3 * rng.random((n, n, n))
    energy = -rng.random((n, n, n))
    oc = design.PaperOC(n, 0.35, emin=1e-6, move=0.2)
    ph, dp, _ = oc.physical(x)
    new, count, g, dv = oc.update(x, ph, dp, energy)
    return dict(x=x, energy=energy, new=new, count=np.array(count), g=g, dv=dv)


def solve64_fixture():
    """Plain method at 64^3 (Emin 1e-6, box radius 4, seed 20261104, rtol 1e-8, zero start) at alpha 0, 0.5
    and 1: per-load iterations, relative residuals and C^H. The recurrences are per load, so these are the
    serial and the parallel load mode's numbers alike (tests/test_oracle_pcg.py checks serial = parallel)."""
    n = 64
    a = design.simp(design.filtered_noise(n, 20261104, 4), 1e-6)
    out = {}
    for tag, alpha in (('a0', 0.0), ('a05', 0.5), ('a1', 1.0)):
        s = green.scale(a, alpha, C0)
        ref = pcg.ResidentPCG(a, C0, np.eye(6), batch=6, s=s)
        t = time.time()
        infos, _ = ref.solve(rtol=1e-8)
        its = [i['iterations'] for i in infos]
        print(f'64^3 alpha {alpha}: {time.time() - t:.1f} s, iterations {its}', flush=True)
        out['iterations_' + tag] = np.array(its)
        out['relative_residual_' + tag] = np.array([i['relative_residual'] for i in infos])
        out['CH_' + tag] = design.effective_tensor(a, ref.solutions, C0)
    return out


FIXTURES = {'operator_4x5x6': operator_fixture, 'green_3x4x5': green_fixture, 'solve_16_alpha05': solve_fixture,
            'oc_cone_8': oc_fixture}
# minutes of CPU: written only when named on the command line
LARGE = {'solve_64_alpha': solve64_fixture}


if __name__ == '__main__':
    chosen = {k: v for k, v in {**FIXTURES, **LARGE}.items() if k in sys.argv[1:]} or FIXTURES
    for name, fn in chosen.items():
        path = os.path.join(HERE, name + '.npz')
        np.savez_compressed(path, **fn())
        print(name, os.path.getsize(path), 'bytes')
