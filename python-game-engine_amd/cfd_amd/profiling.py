"""Profiling of the GPU step with the reference's API (src/cfd/profiling).

Same vocabulary, configuration classes, recorder methods and JSON layout as
the reference (``config.py``, ``timing.py``, ``metrics.py``, ``session.py``,
``result.py``, ``summary.py``, ``schema.py``), with the measurements taken on
the device:

* :class:`TimingProfiler` times each enabled scope with a pair of HIP events
  on the current stream (GPU time of the work enqueued inside the scope; the
  reference times host wall clock), resolved lazily so the step never waits.
* :class:`MetricsRecorder` copies captures into preallocated device buffers
  and reduces them on the device (``smk_field_stats``: the six reducers over
  the three scopes in one set of launches); results stay on the device until
  they are read, so ``process()`` adds no host synchronisation.

SmokeSimulation2D calls ``profiler.tick()``, ``profiler.timing.scope(name)``,
``profiler.metrics.capture(source, stage, field)``, ``profiler.metrics.store``
and ``profiler.metrics.process()`` at the points of
src/cfd/common/simulation.py:193-265.
"""
from __future__ import annotations

import contextlib
import json
import platform
import subprocess
from collections import defaultdict
from dataclasses import dataclass, field
from datetime import datetime, timezone

import numpy as np
import torch

SCHEMA_VERSION = 1

# config.py:223-301 (vocabulary)
SCALAR_METRICS = ["solver.iterations"]
TIMINGS = [
    "simulation.substep", "simulation.emitters", "simulation.forces",
    "simulation.advection.velocity", "simulation.advection.density",
    "simulation.advection.temperature", "simulation.cooling", "simulation.dissipation",
    "simulation.divergence", "simulation.pressure.solve", "simulation.pressure.apply",
    "profiling.capture", "profiling.metrics",
]
SOURCES = ["divergence", "pressure", "pressure_residual", "velocity_u", "velocity_v"]
SOURCE_STAGES = {
    "divergence": {"before_projection", "after_projection"},
    "pressure": {"final"},
    "pressure_residual": {"final"},
}
# order of the device reducer output (smk_field_stats)
REDUCERS = ("rms", "mean", "mean_abs", "max", "max_abs", "std")
SCOPES = ["all", "interior", "boundary"]


# ---------------------------------------------------------------- config
@dataclass
class FieldMetricConfig:
    source: str
    stages: list = field(default_factory=list)
    reducers: list = field(default_factory=list)
    scopes: list = field(default_factory=list)
    every: int = 1

    def validate_config(self) -> None:
        if self.source not in SOURCES:
            raise ValueError(f"Unknown field metric source: {self.source!r}. Must be one of {SOURCES}.")
        allowed = SOURCE_STAGES.get(self.source, set())
        bad = set(self.stages) - allowed
        if bad:
            raise ValueError(f"Invalid stages {sorted(bad)} for source {self.source!r}. "
                             f"Must be a subset of {sorted(allowed)}.")
        bad = set(self.reducers) - set(REDUCERS)
        if bad:
            raise ValueError(f"Invalid reducers {sorted(bad)}. Must be a subset of {sorted(REDUCERS)}.")
        bad = set(self.scopes) - set(SCOPES)
        if bad:
            raise ValueError(f"Invalid scopes {sorted(bad)}. Must be a subset of {SCOPES}.")
        if self.every <= 0:
            raise ValueError(f"'every' must be a positive integer, got {self.every}.")

    def to_dict(self) -> dict:
        return {"source": self.source, "stages": self.stages, "reducers": self.reducers,
                "scopes": self.scopes, "every": self.every}

    @classmethod
    def from_dict(cls, data: dict) -> "FieldMetricConfig":
        return cls(source=data["source"], stages=data.get("stages", []),
                   reducers=data.get("reducers", []), scopes=data.get("scopes", []),
                   every=data.get("every", 1))


@dataclass
class ScalarMetricConfig:
    name: str
    every: int = 1

    def validate_config(self) -> None:
        if self.name not in SCALAR_METRICS:
            raise ValueError(f"Unknown scalar metric: {self.name!r}. Must be one of {SCALAR_METRICS}.")
        if self.every <= 0:
            raise ValueError(f"'every' must be a positive integer, got {self.every}.")

    def to_dict(self) -> dict:
        return {"name": self.name, "every": self.every}

    @classmethod
    def from_dict(cls, data: dict) -> "ScalarMetricConfig":
        return cls(name=data["name"], every=data.get("every", 1))


@dataclass
class MetricsConfig:
    fields: list = field(default_factory=list)
    scalars: list = field(default_factory=list)

    def validate_config(self) -> None:
        for f in self.fields:
            f.validate_config()
        for s in self.scalars:
            s.validate_config()

    def to_dict(self) -> dict:
        return {"fields": [f.to_dict() for f in self.fields],
                "scalars": [s.to_dict() for s in self.scalars]}

    @classmethod
    def from_dict(cls, data: dict) -> "MetricsConfig":
        return cls(fields=[FieldMetricConfig(**f) for f in data.get("fields", [])],
                   scalars=[ScalarMetricConfig(**s) for s in data.get("scalars", [])])


@dataclass
class ProfileConfig:
    timings: list = field(default_factory=list)
    metrics: MetricsConfig = field(default_factory=MetricsConfig)
    snapshots: list = field(default_factory=list)

    def validate_config(self) -> None:
        bad = set(self.timings) - set(TIMINGS)
        if bad:
            raise ValueError(f"Invalid timings {sorted(bad)}. Must be a subset of {TIMINGS}.")
        if not isinstance(self.metrics, MetricsConfig):
            raise ValueError("metrics must be an instance of MetricsConfig.")
        self.metrics.validate_config()

    @classmethod
    def disabled(cls) -> "ProfileConfig":
        return cls(timings=[], metrics=MetricsConfig(fields=[], scalars=[]), snapshots=[])

    @classmethod
    def full(cls) -> "ProfileConfig":
        fields = [FieldMetricConfig(source=src, stages=sorted(stages),
                                    reducers=sorted(REDUCERS), scopes=list(SCOPES), every=1)
                  for src, stages in SOURCE_STAGES.items() if src in SOURCES]
        scalars = [ScalarMetricConfig(name=n, every=1) for n in SCALAR_METRICS]
        return cls(timings=list(TIMINGS), metrics=MetricsConfig(fields=fields, scalars=scalars),
                   snapshots=[])

    def to_dict(self) -> dict:
        return {"timings": self.timings, "metrics": self.metrics.to_dict(),
                "snapshots": self.snapshots}

    @classmethod
    def from_dict(cls, data: dict) -> "ProfileConfig":
        return cls(timings=data.get("timings", []),
                   metrics=MetricsConfig.from_dict(data.get("metrics", {})),
                   snapshots=data.get("snapshots", []))

    @classmethod
    def from_json(cls, path: str) -> "ProfileConfig":
        with open(path) as f:
            return cls.from_dict(json.load(f))


# ---------------------------------------------------------------- timing
class TimingProfiler:
    """timing.py with HIP events: each scope call records an event pair on the
    current stream; durations (ns) are resolved when read."""

    def __init__(self, enabled_timings=None):
        self.enabled_timings = set(enabled_timings or [])
        self.enabled = bool(self.enabled_timings)
        self._pending: dict = defaultdict(list)
        self._resolved: dict = defaultdict(list)

    @contextlib.contextmanager
    def scope(self, name: str):
        if not self.enabled or name not in self.enabled_timings:
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

    def _resolve(self):
        if not self._pending:
            return
        torch.cuda.synchronize()
        for name, pairs in self._pending.items():
            self._resolved[name].extend(s.elapsed_time(e) * 1e6 for s, e in pairs)
        self._pending.clear()

    @property
    def durations_ns(self) -> dict:
        self._resolve()
        return self._resolved

    def get_total_ns(self, name: str) -> float:
        return float(sum(self.durations_ns.get(name, [])))

    def get_avg_ns(self, name: str) -> float:
        v = self.durations_ns.get(name, [])
        return float(sum(v) / len(v)) if v else 0.0

    def report_lines(self) -> list:
        lines = []
        for name, values in sorted(self.durations_ns.items(), key=lambda kv: sum(kv[1]),
                                   reverse=True):
            calls = len(values)
            total_ms = sum(values) / 1_000_000
            avg_ms = total_ms / calls if calls else 0.0
            lines.append(f"{name:30s} | {calls:5d} calls | {total_ms:.3f} ms total | "
                         f"{avg_ms:.3f} ms avg")
        return lines

    def to_dict(self) -> dict:
        out = {}
        for name, values in self.durations_ns.items():
            calls = len(values)
            total_ms = sum(values) / 1_000_000
            out[name] = {"total_ms": total_ms, "calls": calls,
                         "avg_ms": (total_ms / calls) if calls else 0.0}
        return out

    def summary_ms(self) -> dict:
        return {k: {"count": v["calls"], "total_ms": v["total_ms"], "avg_ms": v["avg_ms"]}
                for k, v in self.to_dict().items()}

    def reset(self):
        self._pending.clear()
        self._resolved.clear()


# --------------------------------------------------------------- metrics
_STATS_WS: dict = {}


def field_stats(t: torch.Tensor) -> torch.Tensor:
    """The 18 reductions (3 scopes x 6 reducers) of a 2-D GPU field as a
    float64 device tensor of shape (3, 6): rows all / interior / boundary,
    columns rms / mean / mean_abs / max / max_abs / std."""
    from cfd_amd.backends.hip import _lib
    from cfd_amd.backends.hip.operators import dtype_code

    if t.dim() != 2 or not t.is_contiguous() or t.device.type != "cuda":
        raise ValueError("field_stats needs a contiguous 2-D GPU tensor")
    L = _lib.lib()
    key = (t.device.type, t.device.index)
    ws = _STATS_WS.get(key)
    if ws is None:
        ws = torch.empty(int(L.smk_field_stats_workspace_bytes()), dtype=torch.uint8,
                         device=t.device)
        _STATS_WS[key] = ws
    out = torch.empty((3, 6), dtype=torch.float64, device=t.device)
    _lib.call("smk_field_stats", t.data_ptr(), t.shape[0], t.shape[1], out.data_ptr(),
              ws.data_ptr(), ws.numel(), dtype_code(t), _lib.stream_of(t))
    return out


class MetricsRecorder:
    """metrics.py with the buffers and reductions on the device."""

    def __init__(self, config: MetricsConfig, grid_shape=(1, 1)):
        self.config = config
        self.grid_shape = tuple(grid_shape)
        self.enabled = bool(config.fields or config.scalars)
        self._frame = 0
        self.scalar_results: dict = defaultdict(list)
        self._scalar_configs = {s.name: s for s in config.scalars}
        self._field_configs: dict = {}
        for fc in config.fields:
            for stage in fc.stages:
                self._field_configs[(fc.source, stage)] = fc
        self._buffers: dict = {}
        self._captured_this_frame: set = set()
        self._pending: list = []     # ((source, stage), cfg, device (3, 6) tensor)
        self._field_results: dict = defaultdict(list)

    def tick(self, frame: int):
        self._frame = frame

    def store(self, name: str, value):
        if not self.enabled:
            return
        cfg = self._scalar_configs.get(name)
        if cfg is None or self._frame % cfg.every != 0:
            return
        self.scalar_results[name].append(float(value))

    def capture(self, source: str, stage: str, array):
        if not self.enabled:
            return
        key = (source, stage)
        cfg = self._field_configs.get(key)
        if cfg is None or self._frame % cfg.every != 0:
            return
        buf = self._buffers.get(key)
        if buf is None or buf.shape != array.shape or buf.dtype != array.dtype:
            buf = torch.empty_like(array)
            self._buffers[key] = buf
        buf.copy_(array)
        self._captured_this_frame.add(key)

    def process(self):
        """Reduce this frame's captures on the device (no host sync)."""
        if not self.enabled:
            return
        for key in sorted(self._captured_this_frame):
            self._pending.append((key, self._field_configs[key], field_stats(self._buffers[key])))
        self._captured_this_frame.clear()

    def _resolve(self):
        if not self._pending:
            return
        for (source, stage), cfg, dev in self._pending:
            vals = dev.cpu().numpy()
            for scope in cfg.scopes:
                for reducer in cfg.reducers:
                    self._field_results[(source, stage, scope, reducer)].append(
                        float(vals[SCOPES.index(scope), REDUCERS.index(reducer)]))
        self._pending.clear()

    @property
    def field_results(self) -> dict:
        self._resolve()
        return self._field_results

    def report_lines(self) -> list:
        lines = []
        for name, values in sorted(self.scalar_results.items()):
            calls = len(values)
            avg = sum(values) / calls if calls else 0.0
            lines.append(f"{name:20s} | {calls:5d} calls | {sum(values):.3f} total | {avg:.3f} avg")
        for (source, stage, scope, reducer), values in sorted(self.field_results.items()):
            avg = sum(values) / len(values) if values else 0.0
            lines.append(f"{source}.{stage}.{scope}.{reducer:10s} | {len(values):5d} samples | "
                         f"{avg:.6f} avg")
        return lines

    def to_dict(self) -> dict:
        out = {"scalars": {}, "fields": {}}
        for name, values in self.scalar_results.items():
            calls = len(values)
            out["scalars"][name] = {"total": sum(values), "calls": calls,
                                    "avg": (sum(values) / calls) if calls else 0.0}
        for (source, stage, scope, reducer), values in self.field_results.items():
            calls = len(values)
            out["fields"][f"{source}.{stage}.{scope}.{reducer}"] = {
                "calls": calls, "avg": (sum(values) / calls) if calls else 0.0,
                "last": values[-1] if values else None}
        return out

    def reset(self):
        self.scalar_results.clear()
        self._field_results.clear()
        self._pending.clear()
        self._captured_this_frame.clear()


# ------------------------------------------------------------- snapshots
class SnapshotRecorder:
    """snapshots.py: values and fields captured at given frames."""

    def __init__(self, enabled: bool = True):
        self.enabled = enabled
        self._snapshots: dict = {}
        self._field_snapshots: dict = {}
        self._values: dict = {}
        self._fields: dict = {}

    def register_snapshot(self, name, frame, getter):
        self._snapshots[name] = (frame, getter)

    def register_field_snapshot(self, name, frame, getter):
        self._field_snapshots[name] = (frame, getter)

    def get_value(self, name):
        return self._values.get(name)

    def get_field(self, name):
        return self._fields.get(name)

    def tick(self, frame: int):
        if not self.enabled:
            return
        for name, (target, getter) in self._snapshots.items():
            if frame == target:
                v = getter()
                self._values[name] = float(v) if np.isscalar(v) else v
        for name, (target, getter) in self._field_snapshots.items():
            if frame == target:
                v = getter()
                self._fields[name] = v.clone() if isinstance(v, torch.Tensor) else v

    def to_dict(self) -> dict:
        return {name: value for name, value in self._values.items()}

    def reset(self):
        self._snapshots.clear()
        self._field_snapshots.clear()
        self._values.clear()
        self._fields.clear()


# ---------------------------------------------------------------- results
@dataclass
class VariableSummary:
    """summary.py."""
    name: str
    count: int
    total: float
    mean: float
    min: float
    max: float
    std: float
    p95: float
    p99: float

    @classmethod
    def from_values(cls, name, values) -> "VariableSummary":
        if not values:
            return cls(name, 0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0)
        a = np.asarray(values, dtype=np.float64)
        return cls(name, a.size, float(a.sum()), float(a.mean()), float(a.min()),
                   float(a.max()), float(a.std()), float(np.percentile(a, 95)),
                   float(np.percentile(a, 99)))

    def to_dict(self) -> dict:
        return {"count": self.count, "total": self.total, "mean": self.mean, "min": self.min,
                "max": self.max, "std": self.std, "p95": self.p95, "p99": self.p99}


@dataclass
class ProfilerResult:
    """result.py."""
    timings: list = field(default_factory=list)
    scalars: list = field(default_factory=list)
    fields: list = field(default_factory=list)
    snapshots: dict = field(default_factory=dict)

    @classmethod
    def from_recorders(cls, timing, metrics, snapshots) -> "ProfilerResult":
        return cls(
            timings=[VariableSummary.from_values(n, [ns / 1_000_000 for ns in d])
                     for n, d in timing.durations_ns.items()],
            scalars=[VariableSummary.from_values(n, v) for n, v in metrics.scalar_results.items()],
            fields=[VariableSummary.from_values(f"{a}.{b}.{c}.{d}", v)
                    for (a, b, c, d), v in metrics.field_results.items()],
            snapshots=snapshots.to_dict())

    def to_dict(self) -> dict:
        return {"timings": {s.name: s.to_dict() for s in self.timings},
                "scalars": {s.name: s.to_dict() for s in self.scalars},
                "fields": {s.name: s.to_dict() for s in self.fields},
                "snapshots": self.snapshots}


def build_metadata(extra: dict | None = None) -> dict:
    """schema.py."""
    try:
        commit = subprocess.check_output(["git", "rev-parse", "HEAD"],
                                         stderr=subprocess.DEVNULL).decode().strip()
    except Exception:
        commit = None
    md = {"schema_version": SCHEMA_VERSION,
          "timestamp": datetime.now(timezone.utc).isoformat(),
          "python_version": platform.python_version(),
          "platform": platform.platform(), "git_commit": commit,
          "backend": "hip", "timing_clock": "hip_events"}
    if extra:
        md.update(extra)
    return md


# ---------------------------------------------------------------- session
class ProfileSession:
    """session.py: timing + metrics + snapshots with unified export."""

    def __init__(self, config: ProfileConfig | bool | None = None, grid_shape=(1, 1),
                 timings=None):
        if config is None or isinstance(config, bool):
            # ProfileSession(enabled) -- time every scope (or the given ones)
            enabled = bool(config)
            config = ProfileConfig(timings=list(timings or TIMINGS)) if enabled \
                else ProfileConfig.disabled()
        self.config = config
        self.timing = TimingProfiler(enabled_timings=config.timings)
        self.metrics = MetricsRecorder(config=config.metrics, grid_shape=grid_shape)
        self.snapshots = SnapshotRecorder(enabled=bool(config.snapshots))
        self._frame = 0

    @property
    def enabled(self) -> bool:
        return self.timing.enabled or self.metrics.enabled or self.snapshots.enabled

    @property
    def frame(self) -> int:
        return self._frame

    @classmethod
    def disabled(cls, grid_shape=(1, 1)) -> "ProfileSession":
        return cls(ProfileConfig.disabled(), grid_shape=grid_shape)

    @classmethod
    def full(cls, grid_shape=(1, 1)) -> "ProfileSession":
        return cls(ProfileConfig.full(), grid_shape=grid_shape)

    def tick(self):
        self._frame += 1
        self.metrics.tick(self._frame)
        self.snapshots.tick(self._frame)

    def report(self):
        if not self.enabled:
            print("Profiling is disabled; no data recorded.")
            return
        print(self.to_string())

    def to_string(self) -> str:
        lines = ["Timing report:"]
        lines.extend(self.timing.report_lines())
        lines.append("\nVariable report:")
        lines.extend(self.metrics.report_lines())
        return "\n".join(lines)

    def to_dict(self, extra_metadata: dict | None = None) -> dict:
        return {"metadata": build_metadata(extra_metadata),
                "checkpoints": self.timing.to_dict(),
                "variables": self.metrics.to_dict(),
                "snapshots": self.snapshots.to_dict()}

    def to_result(self) -> ProfilerResult:
        return ProfilerResult.from_recorders(self.timing, self.metrics, self.snapshots)

    def export_json(self, filepath: str, indent: int = 2, extra_metadata: dict | None = None):
        if not self.enabled:
            print("Profiling is disabled; skipping export.")
            return
        with open(filepath, "w") as f:
            json.dump(self.to_dict(extra_metadata), f, indent=indent)

    def reset(self):
        self._frame = 0
        self.timing.reset()
        self.metrics.reset()
