"""Continuous-batching server for design-step requests on one GPU.

A client is one topology optimization: a grid size n, Emin, alpha, a seeded
initial density and a number of OC steps. Its density x and the six
warm-start displacements stay resident in HBM for the client's lifetime: the
client holds one design slot of the ``DesignGroup`` of its grid size from its
first request to its last. A request is one design step (``bench.py``'s
step, the Study-II form of ``exp/asg_sctop_reproduction_20261105/reproduction.py:69-176``):
filter and projection of x -> material -> partial Green-Jacobi scale -> six
certified PCG solves -> C^H -> shear sensitivities -> OC update of x.

Scheduling is continuous: every tick advances all solving designs of every
grid size by one captured chunk of PCG iterations (designs that are not
solving have no active loads and their blocks exit at once), then finishes
the designs whose loads all stopped (certification, restarts, responses) and
starts the requests that arrived meanwhile, each in its own slice of the
resident set (``pgj_resident_setup/certify/init`` on that design only), so
requests join and leave while others are on the GPU. Admission is by HBM
budget: a size class gets as many slots as its share of the budget holds,
and a client waits for a free slot of its size.

Per-load arithmetic does not depend on which slot a design occupies or on
the other designs in the set, so a client's results are the same alone and
in the mix (checked in tests/test_gpu_serve.py).
"""
import heapq
import math
import time

import numpy as np
import torch

from . import _lib
from .batched import DesignGroup
from .certify import CAP, MAX_RESTARTS, CertificationError, outcome
from .density import filtered_noise
from .mixed import MixedResident
from .outer import PaperOuter

LOADS6 = np.eye(6)


def slot_bytes(n, loads=6, precision='fp64'):
    """HBM bytes one design slot of grid n holds in a resident set (vectors, spectra, fields, partials).

    ``precision='mixed'``: three FP64 outer vectors and five FP32 inner vectors, FP32 fields and spectrum copies.
    """
    N = n ** 3
    nf = n * n * (n // 2 + 1)
    partials = loads * (n * n + n * 6 + n * n * 4 + (N + 255) // 256 * 4) * 8
    fields = N * 8 + 3 * N * 8 + 3 * nf * 16
    if precision == 'mixed':
        vectors = 3 * loads * 3 * N * 8 + 5 * loads * 3 * N * 4
        spectrum = loads * 3 * nf * (16 + 8)
        fields += N * 4 + 3 * N * 4 + 3 * nf * 8
        return vectors + spectrum + fields + 2 * partials
    vectors = 6 * loads * 3 * N * 8
    spectrum = loads * 3 * nf * 16
    return vectors + spectrum + fields + partials


class Client:
    """One optimization: grid n, Emin, alpha, seed, steps; x lives on the device."""

    def __init__(self, cid, n, emin, alpha, seed, steps, radius=2, device='cuda'):
        self.cid, self.n, self.emin, self.alpha, self.seed, self.steps = cid, int(n), float(emin), float(alpha), \
            int(seed), int(steps)
        self.x = torch.tensor(filtered_noise(self.n, self.seed, radius), device=device)
        self.done = 0
        self.failed = 0     # requests whose solve failed certification (the design stays as it was)
        self.slot = None
        self.busy = False
        self.history = []   # per completed step: dict(CH, iterations, restarts, change, latency)


class Request:
    def __init__(self, client, arrival):
        self.client, self.arrival = client, arrival
        self.start = self.end = None
        self.restarts = 0
        self.failed = False


class Server:
    """Design-step server: ``sizes`` are the grid sizes served, ``budget`` the HBM bytes for resident sets."""

    def __init__(self, sizes, budget, volume=0.35, move=0.2, rmin=3.0, rtol=1e-8, chunk=4, max_slots=16,
                 device='cuda', graphs=True, precision='fp64', inner_rtol=1e-4, max_outer=30, cap=CAP, tune=False):
        if precision not in ('fp64', 'mixed'):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        self.sizes = sorted(int(n) for n in sizes)
        self.rtol, self.chunk, self.device, self.precision = rtol, chunk, device, precision
        self.volume, self.move, self.rmin = volume, move, rmin
        self.max_outer, self.cap = max_outer, int(cap)
        self.failed = []   # (client id, grid, Emin, alpha) of requests whose solve failed certification
        share = budget / len(self.sizes)
        self.groups = {}
        for n in self.sizes:
            slots = int(min(max_slots, share // slot_bytes(n, precision=precision)))
            if slots < 1:
                raise ValueError(f'HBM budget {budget / 2**30:.1f} GiB holds no slot of {n}^3')
            if precision == 'mixed':
                self.groups[n] = MixedResident(n, LOADS6, slots, inner_rtol, chunk, max_outer, device=device,
                                               cap=cap)
            else:
                self.groups[n] = DesignGroup(n, slots, LOADS6, chunk, device=device, cap=cap)
        self.free = {n: list(range(g.D))[::-1] for n, g in self.groups.items()}
        self.owner = {n: [None] * g.D for n, g in self.groups.items()}
        self.solving = {n: {} for n in self.sizes}          # slot -> Request
        self.workers = {}
        self.graphs = {}
        # row-march shapes and FFT radix plans for each size class's launch size, before any capture
        self.tuned = {n: g.tune(reps=3) for n, g in self.groups.items()} if tune else None
        if graphs:
            self._capture()

    # -- set-up -----------------------------------------------------------------------
    def _capture(self):
        for n, g in self.groups.items():
            g.setup(self.rtol)
            if self.precision == 'mixed':
                g.u.normal_()
                g.certify()
                g.begin()
            else:
                g.r.normal_()
                g.certify()
                g.initialize()
            g.iteration()
            torch.cuda.synchronize()
            graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(graph):
                g.iteration()
            self.graphs[n] = graph
            torch.cuda.synchronize()
            g.u.zero_()
            (g.active32 if self.precision == 'mixed' else g.active).zero_()
        torch.cuda.synchronize()

    def _worker(self, n, emin):
        key = (n, emin)
        w = self.workers.get(key)
        if w is None:
            w = PaperOuter(n, self.volume, rmin=self.rmin, beta=4.0, eta=0.5, emin=emin, move=self.move,
                           device=self.device).capture()
            self.workers[key] = w
        return w

    def memory(self):
        return {n: g.D * slot_bytes(n, precision=self.precision) for n, g in self.groups.items()}

    # -- request lifecycle --------------------------------------------------------------
    def _admit(self, client):
        if client.slot is None:
            if not self.free[client.n]:
                return False
            client.slot = self.free[client.n].pop()
            self.owner[client.n][client.slot] = client
            self.groups[client.n].u[6 * client.slot:6 * client.slot + 6].zero_()
        return True

    def _start(self, req):
        c = req.client
        g = self.groups[c.n]
        w = self._worker(c.n, c.emin)
        w.set_initial(c.x)
        w.evaluate()
        g.material(c.slot, w.a)
        g.select(c.slot, c.alpha)
        g.setup(self.rtol, c.slot, 1)
        g.certify(c.slot, 1)
        if self.precision == 'mixed':
            req.outer, req.inner = 0, np.zeros(6, dtype=np.int64)
            req.outer_pending = True     # the host decides after the first certification
        else:
            g.initialize(c.slot, 1)
        self.solving[c.n][c.slot] = req

    def _advance(self):
        for n, g in self.groups.items():
            if self.solving[n]:
                if n in self.graphs:
                    self.graphs[n].replay()
                else:
                    g.iteration()

    def _finished(self):
        """(n, slot) of solving designs whose six loads have all stopped (one host copy for all sizes)."""
        sizes = [n for n in self.sizes if self.solving[n]]
        if not sizes:
            return []
        flag = (lambda g: g.active32) if self.precision == 'mixed' else (lambda g: g.active)
        act = torch.cat([flag(self.groups[n]) for n in sizes]).cpu().numpy()
        out, off = [], 0
        for n in sizes:
            D = self.groups[n].D
            a = act[off:off + 6 * D].reshape(D, 6)
            off += 6 * D
            out.extend((n, s) for s in self.solving[n] if not a[s].any())
        return out

    def _complete_mixed(self, n, slot):
        """One refinement boundary of a mixed-precision design: the (6, 5) status rows once certified, None
        while it refines, False when it cannot certify (a broken-down inner solve, a non-finite residual, or
        ``max_outer`` steps spent)."""
        g = self.groups[n]
        req = self.solving[n][slot]
        if not req.outer_pending:
            sl = slice(6 * slot, 6 * slot + 6)
            if bool(g.bad32[sl].any().item()):
                return False
            req.inner += g.iterations32[sl].cpu().numpy()
            g.end(slot, 1)
            g.certify(slot, 1)
        req.outer_pending = False
        v = g.status(slot, 1)
        ok = np.isfinite(v[:, 1]) & (v[:, 1] <= v[:, 2])
        if ok.all():
            return np.column_stack((v, req.inner.astype(np.float64), np.zeros(6)))
        if req.outer >= self.max_outer or not np.all(np.isfinite(v[:, 1])):
            return False
        req.outer += 1
        g.begin(slot, 1)
        return None

    def _complete(self, n, slot, t_end):
        """Advance a design whose loads all stopped: True once its step is answered (certified or failed)."""
        g = self.groups[n]
        req = self.solving[n][slot]
        if self.precision == 'mixed':
            diag = self._complete_mixed(n, slot)
            if diag is None:
                return False
            if diag is False:
                return self._fail(n, slot, t_end)
            return self._finalize(n, slot, t_end, diag)
        g.certify(slot, 1)
        sl = slice(6 * slot, 6 * slot + 6)
        diag = torch.stack((g.norm[sl], g.absolute[sl], g.target2[sl].sqrt(), g.iterations[sl].to(torch.float64),
                            g.bad[sl].to(torch.float64)), dim=1).cpu().numpy()
        ok, stop = outcome(diag, self.cap)
        if not ok.all():
            if req.restarts >= MAX_RESTARTS or stop:
                return self._fail(n, slot, t_end)
            req.restarts += 1
            g.initialize(slot, 1)
            return False
        return self._finalize(n, slot, t_end, diag)

    def _fail(self, n, slot, t_end):
        """A step that failed certification: record it, keep the client's density (no OC update), free the slot
        if the client has no step left, and go on (as ``SequentialServer`` does). The slot keeps the failed
        solve's last iterate as the next warm start (the sequential baseline keeps the previous certified one)."""
        req = self.solving[n][slot]
        c = req.client
        req.end = t_end
        req.failed = True
        self.failed.append((c.cid, c.n, c.emin, c.alpha))
        c.failed += 1
        c.busy = False
        del self.solving[n][slot]
        self._release(n, slot, c)
        return True

    def _release(self, n, slot, c):
        if c.done + c.failed >= c.steps:
            self.owner[n][slot] = None
            self.free[n].append(slot)
            c.slot = None

    def _finalize(self, n, slot, t_end, diag):
        """C^H, energies, OC update and the response of a certified design step."""
        g = self.groups[n]
        req = self.solving[n][slot]
        c = req.client
        CH = g.effective_tensor(slot)
        u = g.solutions(slot)
        e = torch.empty((6, g.N), dtype=torch.float64, device=u.device)
        _lib.call('pgj_hex8_energy', g.op.grid, 6, _lib.ptr(u), _lib.host_doubles(LOADS6.reshape(-1)),
                  _lib.ptr(e), _lib.stream())
        local = -e[3:].sum(0)
        w = self._worker(n, c.emin)
        w.set_initial(c.x)
        w.evaluate()
        count = w.step(local)
        c.x.copy_(w.new)
        req.end = t_end
        c.history.append(dict(CH=CH.cpu().numpy(), iterations=diag[:, 3].astype(int).tolist(),
                              restarts=req.restarts, outer=getattr(req, 'outer', 0), bisections=count,
                              change=float(w.change),
                              latency=req.end - req.arrival, max_rel=float((diag[:, 1] / diag[:, 0]).max())))
        c.done += 1
        c.busy = False
        del self.solving[n][slot]
        self._release(n, slot, c)
        return True

    # -- the loop -----------------------------------------------------------------------
    def run(self, clients, arrivals, horizon=None, deadline=None, clock=None, think=None):
        """Serve ``arrivals`` (sorted (time, client) pairs, time in seconds from start) until all are done.

        ``horizon`` drops arrivals after that many seconds; after ``deadline`` seconds no further request is
        started (the ones in flight finish; the rest count in ``self.unserved``). Returns the completed requests.
        ``clock`` (a callable giving seconds) replaces the wall clock. ``think`` (a callable client -> seconds)
        makes the clients closed-loop: once a client's request is answered, it submits its next step that many
        seconds later (while it has steps left and within the horizon), as a design loop that needs each step's
        result before the next.
        """
        index = {id(c): i for i, c in enumerate(clients)}
        for c in clients:                       # OC graphs of every (grid, Emin) before the clock starts
            self._worker(c.n, c.emin)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        now = (lambda: time.perf_counter() - t0) if clock is None else clock
        pending = []       # requests waiting for a slot or for the client's previous step
        heap = list(arrivals)
        heapq.heapify(heap)
        done = []
        self.unserved = 0
        # wall seconds per phase of the loop: starts, chunk launches, the status read (waits for the chunks),
        # completions (certification, C^H, energies, OC update)
        self.phases = dict(start=0.0, advance=0.0, status=0.0, complete=0.0, ticks=0)
        while heap or pending or any(self.solving[n] for n in self.sizes):
            t = now()
            while heap and heap[0][0] <= t and (horizon is None or heap[0][0] <= horizon):
                at, cid = heapq.heappop(heap)
                pending.append(Request(clients[cid], at))
            if horizon is not None:
                heap = [h for h in heap if h[0] <= horizon]
                heapq.heapify(heap)
            started = []
            if deadline is not None and t > deadline:
                self.unserved = len(pending) + len(heap)
                pending, heap = [], []
            t_start = time.perf_counter()
            for req in pending:
                c = req.client
                if c.busy or c.done + c.failed >= c.steps or not self._admit(c):
                    continue
                c.busy = True
                req.start = now()
                self._start(req)
                started.append(req)
            self.phases['start'] += time.perf_counter() - t_start
            pending = [r for r in pending if r not in started and r.client.done + r.client.failed < r.client.steps]
            if not any(self.solving[n] for n in self.sizes):
                if heap:
                    wait = heap[0][0] - now()
                    if clock is None and wait > 0:
                        time.sleep(min(wait, 0.01))
                    continue
                # nothing in flight and nothing to come: what is still pending waits for slots that idle
                # clients hold (their remaining steps never arrived)
                self.unserved += len(pending)
                break
            t1 = time.perf_counter()
            self._advance()
            t2 = time.perf_counter()
            finished = self._finished()
            t3 = time.perf_counter()
            for n, slot in finished:
                req = self.solving[n][slot]
                if self._complete(n, slot, now()):
                    if not req.failed:
                        done.append(req)
                    c = req.client
                    if think is not None and c.done + c.failed < c.steps:
                        at = now() + float(think(c))
                        if horizon is None or at <= horizon:
                            heapq.heappush(heap, (at, index[id(c)]))
            t4 = time.perf_counter()
            self.phases['advance'] += t2 - t1
            self.phases['status'] += t3 - t2
            self.phases['complete'] += t4 - t3
            self.phases['ticks'] += 1
        torch.cuda.synchronize()
        self.elapsed = now()
        return done


class SequentialServer:
    """The plain drop-in serving the same requests one at a time (the comparison point of ``Server``).

    One FP64 ``Resident`` (parallel load mode, one group of six) per grid size; each client's six warm-start
    displacements are kept on the device and copied into the solver for its requests. No batching: a request
    starts when the previous one has finished.
    """

    def __init__(self, sizes, volume=0.35, move=0.2, rmin=3.0, rtol=1e-8, device='cuda', cap=CAP, tune=False):
        from .resident import Resident
        self.rtol, self.device = rtol, device
        self.volume, self.move, self.rmin = volume, move, rmin
        self.solvers = {int(n): Resident(int(n), loads=LOADS6, load_mode='parallel', batch=6, device=device,
                                         cap=cap) for n in sizes}
        self.tuned = {n: r.tune(reps=3) for n, r in self.solvers.items()} if tune else None
        for r in self.solvers.values():
            r.capture()
        self.workers = {}
        self.warm = {}
        self.unserved = 0
        self.failed = []   # (client id, grid, Emin, alpha) of requests whose solve failed certification
        self.failed_seconds = 0.0   # wall time spent on those requests

    def _worker(self, n, emin):
        w = self.workers.get((n, emin))
        if w is None:
            w = PaperOuter(n, self.volume, rmin=self.rmin, beta=4.0, eta=0.5, emin=emin, move=self.move,
                           device=self.device).capture()
            self.workers[(n, emin)] = w
        return w

    def _step(self, c):
        res = self.solvers[c.n]
        w = self._worker(c.n, c.emin)
        warm = self.warm.get(c.cid)
        res.solutions.copy_(warm) if warm is not None else res.solutions.zero_()
        w.set_initial(c.x)
        w.evaluate()
        res.material(w.a)
        res.select(c.alpha)
        infos, _ = res.solve(rtol=self.rtol)
        _, local, _ = res.objective_energy('shear')
        H = res.effective_tensor()
        count = w.step(local)
        c.x.copy_(w.new)
        self.warm[c.cid] = res.solutions.clone()
        return infos, H, count

    def run(self, clients, arrivals, horizon=None, deadline=None):
        for c in clients:
            self._worker(c.n, c.emin)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0
        queue = sorted(a for a in arrivals if horizon is None or a[0] <= horizon)
        done = []
        for k, (at, cid) in enumerate(queue):
            if deadline is not None and now() > deadline:
                self.unserved = len(queue) - k
                break
            wait = at - now()
            if wait > 0:
                time.sleep(wait)
            req = Request(clients[cid], at)
            req.start = now()
            try:
                infos, H, count = self._step(clients[cid])
            except CertificationError:
                # the solve did not certify (the reference raises there too): the request fails, the client
                # keeps its design, the queue goes on
                torch.cuda.synchronize()
                c = clients[cid]
                c.failed += 1
                self.failed.append((cid, c.n, c.emin, c.alpha))
                self.failed_seconds += now() - req.start
                continue
            torch.cuda.synchronize()
            req.end = now()
            clients[cid].done += 1
            clients[cid].history.append(dict(CH=H.cpu().numpy(), iterations=[i['iterations'] for i in infos],
                                             bisections=count, latency=req.end - req.arrival,
                                             max_rel=max(i['relative_residual'] for i in infos)))
            done.append(req)
        self.elapsed = now()
        return done


def poisson_arrivals(rate, horizon, clients_steps, seed):
    """Arrival times of a Poisson process of ``rate``/s over ``horizon`` s, each assigned to a client.

    ``clients_steps`` is the list of step counts per client; the k-th arrival goes to a uniformly drawn client
    that still has steps to submit (a client submits at most its step count). Returns sorted (time, client).
    """
    rng = np.random.default_rng(seed)
    remaining = list(clients_steps)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t > horizon:
            break
        live = [i for i, r in enumerate(remaining) if r > 0]
        if not live:
            break
        cid = int(live[rng.integers(len(live))])
        remaining[cid] -= 1
        out.append((t, cid))
    return out


def latency_stats(done):
    lat = np.array([(r.end - r.arrival) for r in done]) if done else np.zeros(1)
    return dict(p50=float(np.percentile(lat, 50)), p99=float(np.percentile(lat, 99)),
                mean=float(lat.mean()), count=len(done))


def make_clients(count, sizes, emins, steps, seed, device='cuda'):
    """Clients with grid, Emin and alpha drawn from a seeded generator (alpha uniform in [0, 1])."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(rng.choice(sizes))
        out.append(Client(i, n, float(rng.choice(emins)), float(rng.uniform(0, 1)), seed + 1000 + i, steps,
                          radius=2 if n < 64 else 4, device=device))
    return out


def percentile_ms(values, q):
    return float(np.percentile(np.asarray(values) * 1e3, q)) if len(values) else math.nan
