"""Material/load backends in the shape the reference's resident solvers take.

The reference passes a ``backend`` object to ``Resident(backend, n, chunk,
load_mode)`` (``exp/asg_gpu_resident_20261103/resident.py:61``) and
``PaperResident(backend, n, objective, volume, ...)``
(``exp/asg_sctop_reproduction_20261105/reproduction.py:121-122``). The solver
reads two things from it: the isotropic reference stiffness ``C0`` (6x6 Voigt,
engineering shear) and the macro strains ``loads`` (L, 6). Any object with
those two attributes works here; ``Candidates`` and ``PaperBackend`` are the
two the reference builds:

- ``Candidates``: Study I, lambda = 4/3, mu = 1 (``exp/asg_topopt_20261103/backend.py:57-60``),
  the three shear strains (``backend.py:102``);
- ``PaperBackend``: Study II, E = 1, nu = 0.3, i.e. lambda = 15/26, mu = 5/13,
  all six unit strains (``reproduction.py:56-67``).
"""
import numpy as np


def isotropic_c0(lam, mu):
    """6x6 Voigt stiffness of an isotropic medium (engineering shear)."""
    C = np.zeros((6, 6))
    C[:3, :3] = lam
    C[np.arange(3), np.arange(3)] += 2 * mu
    C[np.arange(3, 6), np.arange(3, 6)] = mu
    return C


def lame(C0):
    """(lambda, mu) of an isotropic 6x6 C0; ValueError for any other stiffness (the kernels take a * C0(lam, mu))."""
    C0 = np.asarray(C0, dtype=np.float64)
    if C0.shape != (6, 6):
        raise ValueError('C0 must be a 6x6 Voigt stiffness')
    lam, mu = float(C0[0, 1]), float(C0[3, 3])
    if not np.allclose(C0, isotropic_c0(lam, mu), rtol=1e-14, atol=1e-14 * abs(C0).max()):
        raise ValueError('the scalar-a resident kernels need an isotropic C0 (a * C0(lambda, mu))')
    return lam, mu


class Backend:
    """``C0`` and ``loads`` of one study."""

    def __init__(self, lam, mu, loads):
        self.C0 = isotropic_c0(lam, mu)
        self.loads = np.atleast_2d(np.asarray(loads, dtype=np.float64))


class Candidates(Backend):
    """Study I: lambda = 4/3, mu = 1, the three shear strains 23, 13, 12."""

    def __init__(self):
        super().__init__(4 / 3, 1.0, np.eye(6)[3:])


class PaperBackend(Backend):
    """Study II: E = 1, nu = 0.3 (lambda = 15/26, mu = 5/13), the six unit strains."""

    def __init__(self):
        super().__init__(15 / 26, 5 / 13, np.eye(6))
