"""Timers (``scico/util.py:245-520``) with device-event elapsed times."""

from __future__ import annotations

import time

import torch


class Timer:
    """Wall-clock timer (``util.py:245``) that can also stamp the device stream."""

    def __init__(self):
        self._t0 = None
        self._acc = 0.0

    def start(self):
        if self._t0 is None:
            self._t0 = time.perf_counter()

    def stop(self):
        if self._t0 is not None:
            self._acc += time.perf_counter() - self._t0
            self._t0 = None

    def elapsed(self) -> float:
        cur = 0.0 if self._t0 is None else time.perf_counter() - self._t0
        return self._acc + cur


class DeviceElapsed:
    """Elapsed time of the device stream at the point it is created,
    measured from a reference event; resolved lazily by ``value()``."""

    def __init__(self, ref: "torch.cuda.Event", offset: float):
        self.ref = ref
        self.offset = offset
        self.ev = torch.cuda.Event(enable_timing=True)
        self.ev.record()

    def value(self) -> float:
        self.ev.synchronize()
        return self.offset + self.ref.elapsed_time(self.ev) / 1000.0
