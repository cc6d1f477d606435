"""Mixed-precision form: FP32 inner PCG inside FP64 iterative refinement.

Each outer step certifies the FP64 iterate (u <- P u, r = f - K u with the
FP64 reference operator, |r| per load), stops the loads whose |r| meets
max(rtol |f|, 1e-13/sqrt(n)), and solves K d = r / |r| for the others with the
resident PCG in FP32 (FP32 row-march HEX8 apply, FP32 Stockham FFT and Green
multiply, FP64 reductions and scalars) to relative tolerance ``inner_rtol``;
then u += |r| d in FP64. The residual, the correction accumulation, the
certification dot products, the geometric-mean diagonal normalisation of the
scale and C^H stay FP64. Every returned solution passes the same FP64
true-residual certification as the plain path (``resident.py:186-209``
semantics; up to ``max_outer`` refinement steps instead of two restarts).
"""
import numpy as np
import torch

from . import _lib
from ._chunks import MaskPipeline, StreamSplit
from .certify import CAP, CertificationError
from .operator import LAMBDA_STUDY2, MU_STUDY2, Operator

LOADS6 = np.eye(6)


class MixedResident:
    """D designs x G loads of one cubic n^3 grid solved by FP32-inner / FP64-outer refinement."""

    def __init__(self, n, loads=LOADS6, designs=1, inner_rtol=1e-4, chunk=4, max_outer=30, lam=LAMBDA_STUDY2,
                 mu=MU_STUDY2, device='cuda', cap=CAP):
        if chunk < 2 or chunk % 2:
            raise ValueError('graph chunk must be even and >= 2')
        if not 0 < inner_rtol < 1:
            raise ValueError('inner_rtol must be in (0, 1)')
        self.loads = np.atleast_2d(np.asarray(loads, dtype=np.float64))
        self.n, self.D, self.G = int(n), int(designs), len(self.loads)
        self.L = self.D * self.G
        if int(cap) < 1:
            raise ValueError('iteration cap must be at least 1')
        self.inner_rtol, self.chunk, self.max_outer, self.cap = float(inner_rtol), chunk, int(max_outer), int(cap)
        shape = (self.n,) * 3
        self.op = Operator(torch.ones(shape, dtype=torch.float64, device=device), None, lam, mu, device).enable_green()
        g = self.op.grid
        lib = _lib.lib()
        N = self.n ** 3
        nf = self.n * self.n * (self.n // 2 + 1)
        self.N, L = N, self.L
        dev = torch.device(device)
        f64 = lambda *s: torch.zeros(s, dtype=torch.float64, device=dev)
        f32 = lambda *s: torch.zeros(s, dtype=torch.float32, device=dev)
        # FP64: fields, scale, outer vectors (SoA), certification state
        self.a = torch.ones((self.D, N), dtype=torch.float64, device=dev)
        self.scale = torch.ones((self.D, 3 * N), dtype=torch.float64, device=dev)
        self.scale_hat = torch.zeros((self.D, 3 * nf), dtype=torch.complex128, device=dev)
        self.u, self.r, self.rhs = f64(L, 3 * N), f64(L, 3 * N), f64(L, 3 * N)
        self.state = torch.zeros(lib.pgj_resident_state_bytes(L), dtype=torch.uint8, device=dev)
        self.ws = torch.empty(lib.pgj_resident_ws_bytes(g, L), dtype=torch.uint8, device=dev)
        self.scale_ws = torch.empty(lib.pgj_resident_ws_bytes(g, 1), dtype=torch.uint8, device=dev)
        # FP32: plan, fields, inner vectors, inner state
        self.plan32 = torch.empty(lib.pgj_plan32_bytes(g), dtype=torch.uint8, device=dev)
        _lib.call('pgj_plan32_init', g, _lib.ptr(self.op._plan), _lib.ptr(self.plan32), _lib.stream())
        self.a32 = torch.ones((self.D, N), dtype=torch.float32, device=dev)
        self.scale32 = torch.ones((self.D, 3 * N), dtype=torch.float32, device=dev)
        self.scale_hat32 = torch.zeros((self.D, 3 * nf), dtype=torch.complex64, device=dev)
        self.d, self.r32, self.p0, self.p1, self.q = (f32(L, 3 * N) for _ in range(5))
        self.state32 = torch.zeros(lib.pgj_resident_state_bytes(L), dtype=torch.uint8, device=dev)
        self.ws32 = torch.empty(lib.pgj_inner_ws_bytes(g, L), dtype=torch.uint8, device=dev)
        offs = (_lib.c_size_t * 7)()
        _lib.check(lib.pgj_resident_state_offsets(L, offs))
        o = list(offs)
        view = lambda st, k, dt, w: st[o[k]:o[k] + w * L].view(dt)
        self.target2 = view(self.state, 1, torch.float64, 8)
        self.norm = view(self.state, 2, torch.float64, 8)
        self.absolute = view(self.state, 3, torch.float64, 8)
        self.active32 = view(self.state32, 5, torch.uint8, 1)
        self.iterations32 = view(self.state32, 4, torch.int64, 8)
        self.bad32 = view(self.state32, 6, torch.uint8, 1)
        self._pipeline = None
        # two streams pay off for the FP32 passes from 96^3 (64^3: 10.0 against 9.7 ms per cold solve)
        self._halves = StreamSplit(self.L, self.G, 2, enable=self.n >= 96)
        self.ch_ws = torch.empty(lib.pgj_effective_tensor_ws_bytes(g) // 8 + 1, dtype=torch.float64, device=dev)
        self.aos = f64(self.G, 3 * N)
        self.graphs = None

    # -- inputs -----------------------------------------------------------------------
    def material(self, d, a):
        self.a[d].copy_(torch.as_tensor(a, dtype=torch.float64).reshape(-1))
        _lib.call('pgj_to_f32', self.N, _lib.ptr(self.a[d]), _lib.ptr(self.a32[d]), _lib.stream())

    def select(self, d, alpha):
        """FP64 scale (geometric-mean normalised diagonal) and spectrum, then their FP32 copies."""
        g = self.op.grid
        _lib.call('pgj_resident_scale', g, _lib.ptr(self.op._plan), _lib.ptr(self.a[d]), float(alpha),
                  _lib.ptr(self.scale[d]), _lib.ptr(self.scale_hat[d]), _lib.ptr(self.scale_ws), _lib.stream())
        _lib.call('pgj_to_f32', 3 * self.N, _lib.ptr(self.scale[d]), _lib.ptr(self.scale32[d]), _lib.stream())
        _lib.call('pgj_to_f32', 2 * self.scale_hat.shape[1], _lib.ptr(self.scale_hat[d]),
                  _lib.ptr(self.scale_hat32[d]), _lib.stream())

    # -- launches: designs [d0, d0 + nd) (default: all) --------------------------------
    def _slice(self, d0, nd):
        nd = self.D - d0 if nd is None else nd
        if d0 < 0 or nd < 1 or d0 + nd > self.D:
            raise ValueError(f'design range [{d0}, {d0 + nd}) outside 0..{self.D}')
        return d0 * self.G, nd * self.G

    def setup(self, rtol, d0=0, nd=None):
        l0, nl = self._slice(d0, nd)
        m = np.tile(self.loads, (nl // self.G, 1)).reshape(-1)
        _lib.call('pgj_resident_setup', self.op.grid, self.L, self.G, l0, nl, _lib.ptr(self.a), _lib.host_doubles(m),
                  float(rtol), int(self.cap), _lib.ptr(self.rhs), _lib.ptr(self.state), _lib.ptr(self.ws),
                  _lib.stream())

    def certify(self, d0=0, nd=None):
        l0, nl = self._slice(d0, nd)
        _lib.call('pgj_resident_certify', self.op.grid, self.L, self.G, l0, nl, _lib.ptr(self.a), _lib.ptr(self.u),
                  _lib.ptr(self.rhs), _lib.ptr(self.r), _lib.ptr(self.state), _lib.ptr(self.ws), _lib.stream())

    def begin(self, d0=0, nd=None):
        """Scaled FP32 residual of the last certification, d = 0, inner masks and first direction."""
        l0, nl = self._slice(d0, nd)
        _lib.call('pgj_inner_begin', self.op.grid, self.L, l0, nl, _lib.ptr(self.r), _lib.ptr(self.state),
                  self.inner_rtol, _lib.ptr(self.r32), _lib.ptr(self.d), _lib.ptr(self.state32), _lib.ptr(self.ws32),
                  _lib.stream())
        _lib.call('pgj_inner_init', self.op.grid, self.L, self.G, l0, nl, _lib.ptr(self.plan32),
                  _lib.ptr(self.scale32), _lib.ptr(self.scale_hat32), _lib.ptr(self.r32), _lib.ptr(self.p0),
                  _lib.ptr(self.state32), _lib.ptr(self.ws32), _lib.stream())

    def iteration(self):
        """``chunk`` FP32 inner iterations of every load of the set (stopped loads exit at once); two halves
        of the set on two streams when it splits into whole designs or one design's loads."""
        def launch(l0, nl):
            _lib.call('pgj_inner_iterate', self.op.grid, self.L, self.G, l0, nl, _lib.ptr(self.plan32),
                      _lib.ptr(self.a32), _lib.ptr(self.scale32), _lib.ptr(self.scale_hat32), _lib.ptr(self.d),
                      _lib.ptr(self.r32), _lib.ptr(self.p0), _lib.ptr(self.p1), _lib.ptr(self.q),
                      _lib.ptr(self.state32), _lib.ptr(self.ws32), self.chunk, _lib.stream())
        self._halves.run(launch, self.L)

    def end(self, d0=0, nd=None):
        """u += |r| d in FP64 for the designs' loads."""
        l0, nl = self._slice(d0, nd)
        _lib.call('pgj_inner_end', self.op.grid, self.L, l0, nl, _lib.ptr(self.d), _lib.ptr(self.ws32),
                  _lib.ptr(self.u), _lib.stream())

    def end_and_certify(self):
        self.end()
        self.certify()

    def status(self, d0=0, nd=None):
        """Host copy of (rhs norm, |r|, threshold) per load of designs [d0, d0 + nd) from the last certification."""
        l0, nl = self._slice(d0, nd)
        sl = slice(l0, l0 + nl)
        return torch.stack((self.norm[sl], self.absolute[sl], self.target2[sl].sqrt()), dim=1).cpu().numpy()

    def tune(self, reps=6, fft=True):
        """Pick the FP32 HEX8 row-march shape and the FP32 FFT radix plans by timing them on this set's buffers
        at its launch shape; returns the choices (``pgj.tune``). Graphs captured before are captured again."""
        from . import tune as tuner
        parts = self._halves.parts
        nl = parts[0][1] if parts else self.L
        with tuner.heterogeneous(self.a), tuner.heterogeneous(self.a32):
            out = self._tune(tuner, nl, reps, fft)
        torch.cuda.synchronize()
        self.u.zero_()
        self.active32.zero_()
        if self.graphs is not None:
            self.capture()
        return out

    def _tune(self, tuner, nl, reps, fft):
        self.setup(1e-8)
        self.u.normal_()
        self.certify()
        self.begin()

        def stage(k):
            _lib.call('pgj_inner_stage', self.op.grid, self.L, self.G, 0, nl, _lib.ptr(self.plan32),
                      _lib.ptr(self.a32), _lib.ptr(self.scale32), _lib.ptr(self.scale_hat32), _lib.ptr(self.d),
                      _lib.ptr(self.r32), _lib.ptr(self.p0), _lib.ptr(self.p1), _lib.ptr(self.q),
                      _lib.ptr(self.state32), _lib.ptr(self.ws32), k, _lib.stream())

        for k in range(4):
            stage(k)
        out = {'row_march': tuner.tune_row_march(self.op.grid, nl, 4, lambda: stage(0), reps)}
        if fft:
            out['fft'] = tuner.tune_fft(self.op.grid, 4, stage, reps)
        out['active'] = int(self.active32.sum().item())   # loads still active at the end of the timing
        return out

    def capture(self):
        self.setup(1e-8)
        self.u.normal_()
        self.certify()
        self.begin()
        self.iteration()
        self.end_and_certify()
        torch.cuda.synchronize()
        self.graphs = {}
        for name in ('begin', 'iteration', 'end_and_certify'):
            graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(graph):
                getattr(self, name)()
            self.graphs[name] = graph
        torch.cuda.synchronize()
        self.u.zero_()
        return self

    def _run(self, name):
        if self.graphs is None:
            getattr(self, name)()
        else:
            self.graphs[name].replay()

    # -- the solve --------------------------------------------------------------------
    def solve(self, rtol=1e-8):
        """Refine every load to max(rtol |f|, 1e-13/sqrt(n)) from the current u; returns (infos, stats)."""
        self.setup(rtol)
        self.certify()
        inner_total = np.zeros(self.L, dtype=np.int64)
        outer = host_checks = 0
        while True:
            v = torch.stack((self.norm, self.absolute, self.target2.sqrt()), dim=1).cpu().numpy()
            ok = np.isfinite(v[:, 1]) & (v[:, 1] <= v[:, 2])
            if ok.all():
                break
            if outer >= self.max_outer or not np.all(np.isfinite(v[:, 1])):
                raise CertificationError(('mixed-precision refinement did not certify', self._infos(v, ok, outer)))
            outer += 1
            self._run('begin')
            if self._pipeline is None:
                self._pipeline = MaskPipeline([self.active32])
            host_checks += self._pipeline.run(lambda: self._run('iteration'))
            if bool(self.bad32.any().item()):
                raise CertificationError(('FP32 inner PCG broke down (non-positive or non-finite curvature)',
                                          self._infos(v, ok, outer)))
            inner_total += self.iterations32.cpu().numpy()
            self._run('end_and_certify')
        return self._infos(v, ok, outer, inner_total), dict(outer=outer, inner=int(inner_total.sum()),
                                                           host_checks=host_checks)

    @staticmethod
    def _infos(v, ok, outer, inner=None):
        inner = np.zeros(len(v), dtype=np.int64) if inner is None else inner
        return [dict(rhs_norm=float(row[0]), absolute_residual=float(row[1]), threshold=float(row[2]),
                     iterations=int(inner[j]), relative_residual=float(row[1] / max(row[0], 1e-300)),
                     converged=bool(ok[j]), outer=outer) for j, row in enumerate(v)]

    # -- results ------------------------------------------------------------------------
    def solutions(self, d):
        _lib.call('pgj_soa_to_aos', self.op.grid, self.G, _lib.ptr(self.u[d * self.G:(d + 1) * self.G]),
                  _lib.ptr(self.aos), _lib.stream())
        return self.aos.clone()

    def effective_tensor(self, d=0):
        """C^H (6, 6) of design d from the FP64 solutions (``reproduction.py:166-176``)."""
        if self.G != 6 or not np.array_equal(self.loads, LOADS6):
            raise ValueError('C^H needs the six unit Voigt strains')
        sol = self.solutions(d)
        out = torch.empty(36, dtype=torch.float64, device=self.a.device)
        _lib.call('pgj_effective_tensor', self.op.grid, _lib.ptr(self.a[d]), _lib.ptr(sol), _lib.ptr(self.ch_ws),
                  _lib.ptr(out), _lib.stream())
        return out.reshape(6, 6)

    def energies(self, d=0):
        """FP64 element energies (G, N) of design d's solutions."""
        sol = self.solutions(d)
        out = torch.empty((self.G, self.N), dtype=torch.float64, device=sol.device)
        _lib.call('pgj_hex8_energy', self.op.grid, self.G, _lib.ptr(sol), _lib.host_doubles(self.loads.reshape(-1)),
                  _lib.ptr(out), _lib.stream())
        return out
