"""Minimal profiling hooks with the reference's call shape.

SmokeSimulation2D calls ``profiler.tick()``, ``profiler.timing.scope(name)``
and ``profiler.metrics.capture(source, stage, field)`` at the points of
src/cfd/common/simulation.py:193-265.  :class:`ProfileSession` records
per-scope GPU time with HIP events when enabled (event pairs on the current
stream, resolved lazily), and nothing when disabled.
"""
from __future__ import annotations

import contextlib
from collections import defaultdict

import torch


class _NullTiming:
    @contextlib.contextmanager
    def scope(self, name):
        yield


class _NullMetrics:
    def capture(self, source, stage, field):
        pass

    def process(self):
        pass


class _EventTiming:
    def __init__(self, enabled_scopes=None):
        self.enabled = None if enabled_scopes is None else set(enabled_scopes)
        self._pending = defaultdict(list)

    @contextlib.contextmanager
    def scope(self, name):
        if self.enabled is not None and name not in self.enabled:
            yield
            return
        start = torch.cuda.Event(enable_timing=True)
        end = torch.cuda.Event(enable_timing=True)
        start.record()
        try:
            yield
        finally:
            end.record()
            self._pending[name].append((start, end))

    def summary_ms(self) -> dict:
        torch.cuda.synchronize()
        out = {}
        for name, pairs in self._pending.items():
            times = [s.elapsed_time(e) for s, e in pairs]
            out[name] = {"count": len(times), "total_ms": sum(times),
                         "avg_ms": sum(times) / max(len(times), 1)}
        return out

    def reset(self):
        self._pending.clear()


class ProfileSession:
    def __init__(self, enabled: bool = False, timings=None):
        self.enabled = enabled
        self.timing = _EventTiming(timings) if enabled else _NullTiming()
        self.metrics = _NullMetrics()
        self.frame = 0

    @classmethod
    def disabled(cls) -> "ProfileSession":
        return cls(enabled=False)

    def tick(self):
        self.frame += 1

    def reset(self):
        self.frame = 0
        if isinstance(self.timing, _EventTiming):
            self.timing.reset()
