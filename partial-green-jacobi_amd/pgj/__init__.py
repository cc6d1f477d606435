"""pgj: MI355X-native partial Green-Jacobi homogenization hot path.

Host-side mirror of the reference's operator / preconditioner / PCG / resident
interfaces over the HIP C ABI in ``libpgj.so`` (declared in ``include/pgj.h``).
"""
from . import _lib, density, pcg, pilot
from .backend import Candidates, PaperBackend
from .certify import CertificationError
from .driver import compute_effective_stiffness_batched
from .operator import Operator, build, diagonal, scale, set_alpha
from .pcg import certified_solve, solve_pcg_batched
from .resident import PaperResident, Resident, SelectorResident

__all__ = ['Candidates', 'CertificationError', 'Operator', 'PaperBackend', 'PaperResident', 'Resident',
           'SelectorResident', 'build', 'certified_solve', 'compute_effective_stiffness_batched', 'diagonal', 'pilot',
           'scale', 'set_alpha', 'solve_pcg_batched']
