"""``gstex_cuda.timer.Timer``: the stage timer the reference's trainer uses
(example.py:14,221-301), with the same methods and printout.

``start(key)`` opens a stage (closing none: a stage still open is discarded),
``stop()`` closes it, ``dump()`` prints each stage's total in milliseconds --
the sum of its HIP-event intervals, recorded on torch's current stream, when
``cuda_support`` is set, else host wall time -- and the total.
"""
from __future__ import annotations

import time

import torch


class Timer:
    def __init__(self, cuda_support: bool = True, disabled: bool = False):
        self.cuda_support = cuda_support
        self.disabled = disabled
        self.laps = 0
        self.key = None
        self.times = {}    # key -> host seconds
        self.spans = {}    # key -> [(start event, stop event)]
        self._t0 = 0.0
        self._ev = None

    def start(self, key):
        if self.disabled:
            return
        self.key = key
        self.times.setdefault(key, 0.0)
        self.spans.setdefault(key, [])
        self._ev = None
        if self.cuda_support:
            self._ev = torch.cuda.Event(enable_timing=True)
            self._ev.record()
        self._t0 = time.time()

    def stop(self):
        if self.disabled or self.key is None:
            return
        self.times[self.key] += time.time() - self._t0
        if self.cuda_support:
            end = torch.cuda.Event(enable_timing=True)
            end.record()
            self.spans[self.key].append((self._ev, end))
        self.key = None

    def lap(self):
        if not self.disabled:
            self.laps += 1

    def dump_key(self, key):
        if self.disabled:
            return None
        host_ms = 1000.0 * self.times[key]
        dev_ms = sum(a.elapsed_time(b) for a, b in self.spans[key]) if self.cuda_support else 0.0
        print(f"{key}\t{dev_ms if self.cuda_support else host_ms:.3f}")
        return host_ms, dev_ms

    def dump(self):
        if self.disabled:
            return
        if self.cuda_support:
            torch.cuda.synchronize()
        total = 0.0
        for key in self.times:
            host_ms, dev_ms = self.dump_key(key)
            total += dev_ms if self.cuda_support else host_ms
        print(f"TOTAL\t{total:.3f}")
