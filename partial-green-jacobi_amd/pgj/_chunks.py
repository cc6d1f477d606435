"""Pipelined host checks of the chunked resident solves.

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
