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


def halves(L, group):
    """Two load slices that each cover whole designs or lie within one design, or None."""
    h = L // 2
    if L < 4 or L % 2:
        return None
    if h % group == 0 or L == group:
        return (0, h), (h, L - h)
    return None


class TwoStreams:
    """Runs ``launch(l0, nl)`` for two halves of a set, the second on a side stream forked from and joined
    back to the current stream (also inside a captured graph), so the kernels of one half fill the tails of
    the other's. The side stream has the default priority."""

    def __init__(self, L, group, enable=True):
        self.parts = halves(L, group) if enable else None
        self.side = torch.cuda.Stream() if self.parts else None

    def run(self, launch, L):
        if self.parts is None:
            launch(0, L)
            return
        (a0, an), (b0, bn) = self.parts
        main = torch.cuda.current_stream()
        fork = torch.cuda.Event()
        fork.record(main)
        self.side.wait_event(fork)
        launch(a0, an)
        with torch.cuda.stream(self.side):
            launch(b0, bn)
        join = torch.cuda.Event()
        join.record(self.side)
        main.wait_event(join)
