"""Host-side scheduling of the chunked resident solves: pipelined checks and two-stream halves.

A solve replays a captured chunk of PCG iterations until every load's active
mask is zero. ``MaskPipeline.run`` enqueues the next chunk before it reads the
masks left by the previous one (copied to pinned host slots behind an event),
so the device never idles while the host decides. A chunk enqueued after every
load stopped is a no-op: inactive loads exit at once and keep their state, so
results and iteration counts equal those of the one-chunk-at-a-time loop.
"""
import torch


class MaskPipeline:
    def __init__(self, masks):
        self.masks = list(masks)
        total = sum(m.numel() for m in self.masks)
        self.flags = torch.zeros((2, total), dtype=torch.uint8, pin_memory=True)
        self.events = [torch.cuda.Event(), torch.cuda.Event()]

    def _record(self, slot):
        off = 0
        for m in self.masks:
            k = m.numel()
            self.flags[slot, off:off + k].copy_(m.reshape(-1), non_blocking=True)
            off += k
        self.events[slot].record()

    def _any(self, slot):
        self.events[slot].synchronize()
        return bool(self.flags[slot].any())

    def run(self, replay):
        """Replay ``replay()`` until no mask has an active load; returns the number of host checks."""
        self._record(0)
        if not self._any(0):
            return 1
        checks = 1
        replay()
        self._record(1)
        slot = 1
        while True:
            replay()
            self._record(1 - slot)
            checks += 1
            if not self._any(slot):
                return checks
            slot = 1 - slot


def split_loads(L, group, parts):
    """``parts`` equal load slices that each cover whole designs or lie within one design, or None."""
    if parts < 2 or L % parts or L // parts < 1:
        return None
    k = L // parts
    if k % group == 0 or (L == group):
        return [(i * k, k) for i in range(parts)]
    return None


class Fork:
    """Runs a list of launches, the first on the current stream and each other one on its own side stream forked
    from and joined back to it (also inside a captured graph). The fork and join events belong to this object and
    live as long as it does, so an event recorded inside a capture outlives the capture (and the graphs made
    from it). A fork inside another fork's launches while a capture is open is refused: two serving variants
    that nested them (and per-class graph replays on concurrent streams) ended the process with a host
    segmentation fault whose cause was not found. Side streams have the default priority."""

    _depth = 0   # forks currently running their launches (one host thread drives them)

    def __init__(self, sides):
        self.side = [torch.cuda.Stream() for _ in range(sides)]
        self.fork = torch.cuda.Event()
        self.join = [torch.cuda.Event() for _ in range(sides)]

    def run(self, launches):
        if len(launches) != len(self.side) + 1:
            raise ValueError(f'{len(launches)} launches for a fork of {len(self.side) + 1} streams')
        if Fork._depth and torch.cuda.is_current_stream_capturing():
            raise RuntimeError('a stream fork nested in another one under graph capture is not supported')
        Fork._depth += 1
        try:
            main = torch.cuda.current_stream()
            self.fork.record(main)
            for st in self.side:
                st.wait_event(self.fork)
            launches[0]()
            for st, ev, launch in zip(self.side, self.join, launches[1:]):
                with torch.cuda.stream(st):
                    launch()
                ev.record(st)
                main.wait_event(ev)
        finally:
            Fork._depth -= 1


class StreamSplit:
    """Runs ``launch(l0, nl)`` for ``parts`` slices of a set through a ``Fork``, so the kernels of one slice fill
    the tails of the others'."""

    def __init__(self, L, group, parts=2, enable=True):
        self.parts = split_loads(L, group, parts) if enable else None
        self._fork = Fork(len(self.parts) - 1) if self.parts else None

    def run(self, launch, L):
        if self.parts is None:
            launch(0, L)
            return
        self._fork.run([lambda l0=l0, nl=nl: launch(l0, nl) for l0, nl in self.parts])
