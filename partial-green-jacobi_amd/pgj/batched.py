"""Batched form: many designs x 6 loads of uneven grids in one captured graph.

Designs of the same grid size n form a ``DesignGroup``: D designs x G loads
are one resident launch of L = D*G loads (libpgj's ``group`` argument), each
design with its own SIMP field, scale s and spectrum s_hat, all loads with
their own PCG scalars and masks. ``BatchedDesigns`` holds one group per grid
size and captures the init / iteration-chunk / certify steps of every group
into one graph each, so one replay advances every design by ``chunk``
iterations. The host reads the active masks of all groups with one copy per
chunk. The per-load recurrence, masks, cap, targets and restarts are those
of the plain resident solver (``exp/asg_gpu_resident_20261103/resident.py:39-58``,
``:119-136``, ``:186-209``); only the launch is shared.

``tune`` picks the HEX8 row-march shape of every group by timing the
candidates on the group's own buffers (``pgj_hex8_set_shape``).
"""
import numpy as np
import torch

from . import _lib
from ._chunks import Fork, MaskPipeline
from .certify import CAP, MAX_RESTARTS, fail, infos as make_infos, outcome
from .operator import LAMBDA_STUDY2, MU_STUDY2, Operator

LOADS6 = np.eye(6)


class DesignGroup:
    """D designs of one periodic n^3 grid, each solved for the same G macro loads."""

    def __init__(self, n, designs, loads=LOADS6, chunk=4, lam=LAMBDA_STUDY2, mu=MU_STUDY2, device='cuda', cap=CAP):
        if chunk < 2 or chunk % 2:
            raise ValueError('graph chunk must be even and >= 2')
        if int(cap) < 1:
            raise ValueError('iteration cap must be at least 1')
        self.loads = np.atleast_2d(np.asarray(loads, dtype=np.float64))
        self.n, self.D, self.G, self.chunk, self.cap = int(n), int(designs), len(self.loads), chunk, int(cap)
        self.L = self.D * self.G
        shape = (self.n,) * 3
        self.shape = shape
        self.op = Operator(torch.ones(shape, dtype=torch.float64, device=device), None, lam, mu, device).enable_green()
        g = self.op.grid
        N = self.n ** 3
        self.N = N
        L = self.L
        dev = torch.device(device)
        z = lambda *s: torch.zeros(s, dtype=torch.float64, device=dev)
        self.a = torch.ones((self.D, N), dtype=torch.float64, device=dev)
        self.scale = torch.ones((self.D, 3 * N), dtype=torch.float64, device=dev)
        nf = self.n * self.n * (self.n // 2 + 1)
        self.scale_hat = torch.zeros((self.D, 3 * nf), dtype=torch.complex128, device=dev)
        self.alpha = np.zeros(self.D)
        # SoA (L, 3, N) solver vectors; u persists between solves as the warm start
        self.u, self.r, self.p0, self.p1, self.q, self.rhs = (z(L, 3 * N) for _ in range(6))
        lib = _lib.lib()
        self.state = torch.zeros(lib.pgj_resident_state_bytes(L), dtype=torch.uint8, device=dev)
        offs = (_lib.c_size_t * 7)()
        _lib.check(lib.pgj_resident_state_offsets(L, offs))
        o = list(offs)
        self.norm = self.state[o[2]:o[2] + 8 * L].view(torch.float64)
        self.absolute = self.state[o[3]:o[3] + 8 * L].view(torch.float64)
        self.target2 = self.state[o[1]:o[1] + 8 * L].view(torch.float64)
        self.iterations = self.state[o[4]:o[4] + 8 * L].view(torch.int64)
        self.active = self.state[o[5]:o[5] + L]
        self.bad = self.state[o[6]:o[6] + L]
        self.ws = torch.empty(lib.pgj_resident_ws_bytes(g, L), dtype=torch.uint8, device=dev)
        self.scale_ws = torch.empty(lib.pgj_resident_ws_bytes(g, 1), dtype=torch.uint8, device=dev)
        self.ch_ws = torch.empty(lib.pgj_effective_tensor_ws_bytes(g) // 8 + 1, dtype=torch.float64, device=dev)
        self.aos = z(self.G, 3 * N)

    # -- per-design inputs ----------------------------------------------------------
    def material(self, d, a):
        """SIMP field of design d (n, n, n)."""
        self.a[d].copy_(torch.as_tensor(a, dtype=torch.float64).reshape(-1))

    def select(self, d, alpha):
        """Partial Green-Jacobi exponent of design d: s_d and its spectrum (``resident.py:156-174``)."""
        _lib.call('pgj_resident_scale', self.op.grid, _lib.ptr(self.op._plan), _lib.ptr(self.a[d]), float(alpha),
                  _lib.ptr(self.scale[d]), _lib.ptr(self.scale_hat[d]), _lib.ptr(self.scale_ws), _lib.stream())
        self.alpha[d] = float(alpha)

    # -- launches: designs [d0, d0 + nd) (default: all) -------------------------------
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

    def initialize(self, d0=0, nd=None):
        l0, nl = self._slice(d0, nd)
        _lib.call('pgj_resident_init', self.op.grid, self.L, self.G, l0, nl, _lib.ptr(self.op._plan),
                  _lib.ptr(self.scale), _lib.ptr(self.scale_hat), _lib.ptr(self.r), _lib.ptr(self.p0),
                  _lib.ptr(self.state), _lib.ptr(self.ws), _lib.stream())

    def iteration(self, count=None, d0=0, nd=None):
        l0, nl = self._slice(d0, nd)
        _lib.call('pgj_resident_iterate', self.op.grid, self.L, self.G, l0, nl, _lib.ptr(self.op._plan),
                  _lib.ptr(self.a), _lib.ptr(self.scale), _lib.ptr(self.scale_hat), _lib.ptr(self.u), _lib.ptr(self.r),
                  _lib.ptr(self.p0), _lib.ptr(self.p1), _lib.ptr(self.q), _lib.ptr(self.state),
                  _lib.ptr(self.ws), int(count or self.chunk), _lib.stream())

    def stage(self, k, p, p_next):
        _lib.call('pgj_resident_stage', self.op.grid, self.L, self.G, 0, self.L, _lib.ptr(self.op._plan),
                  _lib.ptr(self.a), _lib.ptr(self.scale), _lib.ptr(self.scale_hat), _lib.ptr(self.u), _lib.ptr(self.r),
                  _lib.ptr(p), _lib.ptr(p_next), _lib.ptr(self.q), _lib.ptr(self.state),
                  _lib.ptr(self.ws), int(k), _lib.stream())

    # -- results ----------------------------------------------------------------------
    def solutions(self, d):
        """AoS (G, 3N) solutions of design d."""
        _lib.call('pgj_soa_to_aos', self.op.grid, self.G, _lib.ptr(self.u[d * self.G:(d + 1) * self.G]),
                  _lib.ptr(self.aos), _lib.stream())
        return self.aos.clone()

    def effective_tensor(self, d):
        """C^H (6, 6) of design d (``reproduction.py:166-176``)."""
        if self.G != 6 or not np.array_equal(self.loads, LOADS6):
            raise ValueError('C^H needs the six unit Voigt strains')
        _lib.call('pgj_soa_to_aos', self.op.grid, self.G, _lib.ptr(self.u[d * self.G:(d + 1) * self.G]),
                  _lib.ptr(self.aos), _lib.stream())
        out = torch.empty(36, dtype=torch.float64, device=self.a.device)
        _lib.call('pgj_effective_tensor', self.op.grid, _lib.ptr(self.a[d]), _lib.ptr(self.aos), _lib.ptr(self.ch_ws),
                  _lib.ptr(out), _lib.stream())
        return out.reshape(6, 6)

    def tune(self, reps=6, fft=True):
        """Row-march shape for this set's launch size and FFT radix plans, timed on its own buffers
        (``pgj.tune``); returns the choices. Call before capturing graphs of this set."""
        from . import tune as tuner
        with tuner.heterogeneous(self.a):
            self.setup(1e-8)
            self.r.normal_()
            self.certify()
            self.initialize()
            for k in range(4):
                self.stage(k, self.p0, self.p1)
            out = {'row_march': tuner.tune_row_march(self.op.grid, self.L, 8,
                                                     lambda: self.stage(0, self.p0, self.p1), reps)}
            if fft:
                out['fft'] = tuner.tune_fft(self.op.grid, 8, lambda k: self.stage(k, self.p0, self.p1), reps)
            out['active'] = int(self.active.sum().item())
        torch.cuda.synchronize()
        self.u.zero_()
        self.active.zero_()
        return out

    def shape_info(self):
        """(R, W, ZC, plane, blocks per load) of the row-march launch of this set."""
        out = (_lib.c_int * 5)()
        _lib.check(_lib.lib().pgj_hex8_get_shape(self.op.grid, self.L, 8, out))
        return tuple(out)


class BatchedDesigns:
    """Designs of several grid sizes, one ``DesignGroup`` per size, advanced together by one graph.

    ``sizes`` maps n to the number of designs of that size.
    """

    def __init__(self, sizes, loads=LOADS6, chunk=4, device='cuda', cap=CAP):
        self.groups = {int(n): DesignGroup(n, d, loads, chunk, device=device, cap=cap)
                       for n, d in sorted(sizes.items())}
        self.cap = int(cap)
        self.chunk = chunk
        self.init_graph = self.iter_graph = self.cert_graph = None
        self._pipeline = None
        self._forks = {}   # stream forks (default priority) of the concurrent launches, by method name

    def _each(self, name):
        def run():
            for g in self.groups.values():
                getattr(g, name)()
        return run

    def _concurrent(self, name):
        """The sets' launches on one stream each, forked from and joined to the current stream (the sets share
        no buffers); captured into the iteration graph, so small sets fill the CUs a large set leaves idle."""
        groups = list(self.groups.values())
        if len(groups) == 1:
            return getattr(groups[0], name)
        fork = self._forks.get(name)
        if fork is None:
            fork = self._forks[name] = Fork(len(groups) - 1)
        return lambda: fork.run([getattr(g, name) for g in groups])

    def tune(self, reps=6, candidates=None, fft=True):
        """Time row-march shapes and FFT radix plans per group on its own buffers and keep the fastest; returns
        the choices (``pgj.tune``)."""
        from . import tune as tuner
        chosen = {}
        for n, g in self.groups.items():
            if candidates is None:
                choice = g.tune(reps, fft)
            else:
                g.setup(1e-8)
                g.r.normal_()
                g.certify()
                g.initialize()
                g.stage(0, g.p0, g.p1)
                choice = {'row_march': tuner.tune_row_march(g.op.grid, g.L, 8, lambda: g.stage(0, g.p0, g.p1), reps,
                                                            candidates)}
                g.u.zero_()
            chosen[n] = dict(shape=choice['row_march']['shape'], ms=choice['row_march']['ms'])
            if 'fft' in choice:
                chosen[n]['fft'] = choice['fft']
        torch.cuda.synchronize()
        return chosen

    def capture(self):
        """Capture init / chunk / certify for every group into one graph each."""
        for g in self.groups.values():
            g.setup(1e-8)
            g.r.normal_()
        self._each('certify')()
        self._each('initialize')()
        self._each('iteration')()
        torch.cuda.synchronize()
        for name, fn in (('init_graph', self._each('initialize')), ('iter_graph', self._concurrent('iteration')),
                         ('cert_graph', self._each('certify'))):
            graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(graph):
                fn()
            setattr(self, name, graph)
        torch.cuda.synchronize()
        for g in self.groups.values():
            g.u.zero_()
        return self

    def _replay(self, graph, name):
        if graph is None:
            self._each(name)()
        else:
            graph.replay()

    def _any_active(self):
        return bool(torch.cat([g.active for g in self.groups.values()]).any().item())

    def solve(self, rtol=1e-8):
        """Solve every load of every design from the current warm starts; returns per-group info lists.

        Restarts (at most two) and the failure rule follow ``resident.py:186-209`` per load.
        """
        for g in self.groups.values():
            g.setup(rtol)
        restarts = 0
        host_checks = 0
        self._replay(self.cert_graph, 'certify')
        while True:
            self._replay(self.init_graph, 'initialize')
            if self._pipeline is None:
                self._pipeline = MaskPipeline([g.active for g in self.groups.values()])
            host_checks += self._pipeline.run(lambda: self._replay(self.iter_graph, 'iteration'))
            self._replay(self.cert_graph, 'certify')
            diag = {n: torch.stack((g.norm, g.absolute, g.target2.sqrt(), g.iterations.to(torch.float64),
                                    g.bad.to(torch.float64)), dim=1).cpu().numpy() for n, g in self.groups.items()}
            res = {n: outcome(v, self.cap) for n, v in diag.items()}
            ok = {n: r[0] for n, r in res.items()}
            all_ok = all(o.all() for o in ok.values())
            hopeless = any(r[1] for r in res.values())
            if all_ok or restarts >= MAX_RESTARTS or hopeless:
                break
            restarts += 1
        infos = {n: make_infos(rows, ok[n], restarts) for n, rows in diag.items()}
        if not all_ok:
            fail(infos)
        return infos, host_checks
