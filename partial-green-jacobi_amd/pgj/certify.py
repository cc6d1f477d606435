"""True-residual certification outcome shared by every resident solver form.

The rule is ``exp/asg_gpu_resident_20261103/resident.py:196-208``: a load is
certified when its true residual |f - K u| is finite, at most
max(rtol |f|, 1e-13/sqrt(n)), and its recurrence never latched ``bad``. A
group that is not certified restarts from its certified iterate, at most twice,
unless a load latched ``bad`` or stopped at the iteration cap short of the
target; then the solve fails with ``RuntimeError(('true residual certification
failed', info))``. ``CertificationError`` is that ``RuntimeError``: callers
that go on after a failed solve catch it and nothing else.
"""
import numpy as np

CAP = 4000          # iteration cap of the reference (resident.py:50)
MAX_RESTARTS = 2    # resident.py:200


class CertificationError(RuntimeError):
    """A solve whose loads did not pass true-residual certification; ``args[0]`` is (message, info)."""

    @property
    def info(self):
        return self.args[0][1]


def outcome(diag, cap=CAP):
    """(ok per load, stop) from rows (rhs norm, |r|, threshold, iterations, bad) of one certification.

    ``stop`` is True when no restart can help: a load latched ``bad``, or one reached ``cap`` uncertified."""
    diag = np.asarray(diag, dtype=np.float64)
    ok = np.isfinite(diag[:, 1]) & (diag[:, 1] <= diag[:, 2]) & (diag[:, 4] == 0)
    stop = bool(np.any(diag[:, 4]) or np.any((~ok) & (diag[:, 3] >= cap)))
    return ok, stop


def infos(diag, ok, restarts):
    """Per-load info dicts with the reference's keys (resident.py:206-207)."""
    return [dict(rhs_norm=float(v[0]), absolute_residual=float(v[1]), threshold=float(v[2]), iterations=int(v[3]),
                 relative_residual=float(v[1] / max(v[0], 1e-300)), converged=bool(ok[j]), restarts=restarts)
            for j, v in enumerate(np.asarray(diag, dtype=np.float64))]


def fail(info):
    raise CertificationError(('true residual certification failed', info))
