"""Benchmark and suite configuration (src/cfd/benchmarking/config.py).

The JSON schema, benchmark ids and ValueError messages are the reference's,
so its configs (benchmarks/configs/*.json) load here once their backends read
"hip".  Two additions:

* a precision axis -- ``simulation.dtype`` in a suite is a list of field
  dtypes ("float32", "float64"; default ["float32"], the reference's array
  type) expanded like the other list axes; a single benchmark carries one;
* operator backends are checked against the operator registry (the
  reference leaves that check commented out, config.py:319-336, because its
  config names -- "advection", "forces" ... -- are not registry keys;
  ``OPERATOR_GROUPS`` maps each config name to the registry operations the
  simulation calls for it).
"""
from __future__ import annotations

import hashlib
import json
import re
from dataclasses import asdict, dataclass, field
from itertools import product
from typing import Any

from cfd_amd.common.config import _DTYPES, Backend, OperatorBackendConfig, SimulationConfig
from cfd_amd.common.emitters import EMITTER_APPROACHES
from cfd_amd.operators.registry import get_operator
from cfd_amd.profiling import ProfileConfig
from cfd_amd.solvers.params import SOLVER_PARAM_TYPES
from cfd_amd.solvers.registry import get_solver

import cfd_amd.operators  # noqa: F401  (fills the operator registry)
import cfd_amd.solvers  # noqa: F401  (fills the solver registry)

SCHEMA_VERSION = "0.1"

DEFAULT_DTYPES = ["float32"]

# config name -> registry operations the simulation dispatches for it
# (common/simulation.py's _apply_forces, _advect_*, _compute_divergence ...).
OPERATOR_GROUPS = {
    "advection": ("advection_scalar", "advection_velocity_u", "advection_velocity_v"),
    "boundaries": ("enforce_boundary_conditions",),
    "divergence": ("compute_divergence",),
    "forces": ("apply_forces", "cool_temperature"),
    "pressure_gradient": ("apply_pressure_gradient",),
    "residual": ("compute_residual",),
}


def _num(x) -> bool:
    return isinstance(x, (int, float))


# (check, message) in the order config.py:24-47 applies them
_EMITTER_CHECKS = (
    (lambda e: "position" in e and "radius" in e,
     "Each emitter must have 'position' and 'radius' keys."),
    (lambda e: isinstance(e["position"], (list, tuple)) and len(e["position"]) == 2,
     "Emitter 'position' must be a list or tuple of length 2."),
    (lambda e: _num(e["radius"]) and e["radius"] > 0,
     "Emitter 'radius' must be a positive number."),
    (lambda e: _num(e.get("density_value", None)),
     "Emitter 'density_value' must be a number."),
    (lambda e: _num(e.get("temperature_value", None)),
     "Emitter 'temperature_value' must be a number."),
    (lambda e: isinstance(e.get("velocity_value"), (list, tuple)),
     "Emitter 'velocity_value' must be a list or tuple."),
    (lambda e: len(e["velocity_value"]) == 2 and all(_num(v) for v in e["velocity_value"]),
     "Emitter 'velocity_value' must contain only numbers and be length 2."),
    (lambda e: isinstance(e.get("approach"), str),
     "Emitter 'approach' must be a string."),
    (lambda e: e["approach"] in EMITTER_APPROACHES,
     f"Emitter 'approach' must be one of {EMITTER_APPROACHES}."),
    (lambda e: isinstance(e.get("numba_optimized"), bool),
     "Emitter 'numba_optimized' must be a boolean."),
)


def validate_emitters(emitters: list[dict]) -> None:
    """config.py:20-47: a non-empty list of complete emitter dicts."""
    if not emitters or not isinstance(emitters, list):
        raise ValueError("Emitters must be provided and must be a list of dictionaries.")
    for emitter in emitters:
        if not isinstance(emitter, dict):
            raise ValueError("Each emitter must be a dictionary.")
        for ok, message in _EMITTER_CHECKS:
            if not ok(emitter):
                raise ValueError(message)


@dataclass
class RunConfig:
    """config.py:50-55."""
    frames: int
    warmup_frames: int
    repetitions: int = 1


@dataclass
class BenchmarkConfig:
    """One benchmark: a simulation, how long to run it and what to record
    (config.py:58-157)."""
    name: str
    simulation: SimulationConfig
    run: RunConfig
    profiling: ProfileConfig = field(default_factory=ProfileConfig)
    emitters: list[dict] = field(default_factory=list)

    @classmethod
    def from_dict(cls, data: dict) -> "BenchmarkConfig":
        return cls(name=data["name"],
                   simulation=SimulationConfig.from_dict(data["simulation"]),
                   run=RunConfig(**data["run"]),
                   profiling=ProfileConfig.from_dict(data.get("profiling") or {}),
                   emitters=data.get("emitters", []))

    @classmethod
    def from_json(cls, path: str) -> "BenchmarkConfig":
        with open(path) as f:
            return cls.from_dict(json.load(f))

    def validate_config(self) -> None:
        if self.run.repetitions <= 0:
            raise ValueError("Repetitions must be a positive integer.")
        try:
            self.simulation.validate_config()
        except ValueError as e:
            raise ValueError(f"Simulation config validation failed: {e}")
        if self.run.frames <= 0 or self.run.warmup_frames < 0:
            raise ValueError(
                "Run config validation failed: frames and warmup_frames must be positive.")
        try:
            self.profiling.validate_config()
        except ValueError as e:
            raise ValueError(f"Profiling config validation failed: {e}")
        try:
            validate_emitters(self.emitters)
        except ValueError as e:
            raise ValueError(f"Emitters config validation failed: {e}")

    def to_dict(self) -> dict:
        return {"name": self.name,
                "schema_version": SCHEMA_VERSION,
                "type": "benchmark",
                "simulation": self.simulation.to_dict(),
                "emitters": self.emitters,
                "run": asdict(self.run),
                "profiling": self.profiling.to_dict()}


def _slug(value: object) -> str:
    return re.sub(r"[^a-z0-9]+", "-", str(value).lower()).strip("-")


def _digest(config: BenchmarkConfig) -> str:
    blob = json.dumps(config.to_dict(), sort_keys=True, separators=(",", ":")).encode()
    return hashlib.sha256(blob).hexdigest()[:8]


def make_benchmark_id(index: int, config: BenchmarkConfig) -> str:
    """``NNN-<method>-<backend>-<W>x<H>-<sha256[:8] of the sorted config>``
    (config.py:159-196).  The digest covers the whole config, the dtype
    included, so two precisions of one solver get two ids."""
    sim = config.simulation
    w, h = sim.grid.size
    parts = [sim.pressure_solver.method, Backend(sim.pressure_solver.backend).value, f"{w}x{h}"]
    return "-".join([f"{index:03d}", *map(_slug, parts), _digest(config)])


def _list_of(kind, values) -> bool:
    return isinstance(values, list) and bool(values) and all(isinstance(v, kind) for v in values)


@dataclass
class SuiteConfig:
    """A benchmark definition with list-valued axes (config.py:198-616)."""
    name: str
    benchmark_definitions: dict[str, Any]

    # ---------------------------------------------------------- validation
    def _validate_grid(self, grid) -> None:
        if not grid:
            raise ValueError("Grid configuration must be provided.")
        sizes = grid.get("size")
        if not sizes or any(not isinstance(s, (list, tuple)) or len(s) != 2 for s in sizes):
            raise ValueError("Grid size must be a list of 2D tuples.")
        for s in sizes:
            if any(not isinstance(d, int) or d <= 0 for d in s):
                raise ValueError("Grid dimensions must be positive integers.")
        cells = grid.get("cell_size")
        if not cells or not isinstance(cells, list):
            raise ValueError("Grid cell size must be a list of positive numbers.")
        if any(not _num(c) or c <= 0 for c in cells):
            raise ValueError("Cell sizes must be all positive number.")

    def _validate_dtypes(self, dtypes) -> None:
        if not _list_of(str, dtypes) or any(d not in _DTYPES for d in dtypes):
            raise ValueError(f"Simulation dtype must be a non-empty list drawn from {list(_DTYPES)}.")
        if len(dtypes) != len(set(dtypes)):
            raise ValueError("Simulation dtypes must be unique.")

    def _validate_solvers(self, solver) -> None:
        if not solver:
            raise ValueError("Pressure solver configuration must be provided.")
        methods, backends = solver.get("method"), solver.get("backend")
        if not methods or not isinstance(methods, dict) or \
                any(not isinstance(m, str) for m in methods):
            raise ValueError("Pressure solver methods must be a dictionary with string keys.")
        if not _list_of(str, backends):
            raise ValueError("Pressure solver backend must be a list of strings.")
        if len(backends) != len(set(backends)):
            raise ValueError("Pressure solver backends must be unique.")
        known = [b.value for b in Backend]
        for method, method_config in methods.items():
            if method not in SOLVER_PARAM_TYPES:
                raise ValueError(f"Unsupported pressure solver method: {method}. Supported methods "
                                 f"are: {list(SOLVER_PARAM_TYPES)}")
            if not isinstance(method_config, dict):
                raise ValueError(
                    f"Pressure solver method configuration for {method} must be a dictionary.")
            for backend in backends:
                if backend not in known:
                    raise ValueError(f"Unsupported pressure solver backend: {backend}. Supported "
                                     f"backends are: {known}")
                try:
                    get_solver(method, backend)
                except ValueError as e:
                    raise ValueError(f"Suite pressure solver validation failed: {e}")
            if "params" not in method_config:
                raise ValueError(f"Pressure solver params for method {method} must be provided.")
            param_sets = method_config["params"]
            if not param_sets or not isinstance(param_sets, list):
                raise ValueError(
                    f"Pressure solver params for method {method} must be a list of dictionaries.")
            ptype = SOLVER_PARAM_TYPES[method]
            for params in param_sets:
                if not isinstance(params, dict):
                    raise ValueError(
                        f"Pressure solver params for method {method} must be a dictionary.")
                try:
                    ptype(**params)
                except (TypeError, ValueError) as e:
                    raise ValueError(
                        f"Invalid parameters for pressure solver method {method}: {e}") from e

    def _validate_operators(self, operators) -> None:
        if not operators or not isinstance(operators, dict):
            raise ValueError("Operators configuration must be provided and must be a dictionary.")
        required = set(asdict(OperatorBackendConfig()))
        missing = required - set(operators)
        if missing:
            raise ValueError(f"Missing required operators: {missing}. Required operators are: "
                             f"{required}")
        known = {b.value for b in Backend}
        for name, backends in operators.items():
            if not isinstance(name, str):
                raise ValueError("Operator names must be strings.")
            if not isinstance(backends, list) or not backends:
                raise ValueError(
                    f"Backends for operator '{name}' must be a non-empty list of strings.")
            if any(not isinstance(b, str) for b in backends):
                raise ValueError(f"Backends for operator '{name}' must be strings.")
            if len(backends) != len(set(backends)):
                raise ValueError(f"Backends for operator '{name}' must be unique.")
            for backend in backends:
                if backend not in known:
                    raise ValueError(f"Unsupported backend '{backend}' for operator '{name}'. "
                                     f"Supported backends are: {sorted(known)}")
                for op in OPERATOR_GROUPS.get(name, ()):
                    try:
                        get_operator(op, backend)
                    except ValueError as e:
                        raise ValueError(f"Operator '{name}' is not available for backend "
                                         f"'{backend}': {e}") from e

    def _validate_simulation(self, simulation: dict | None) -> None:
        if not simulation:
            raise ValueError("Simulation configuration must be provided.")
        self._validate_grid(simulation.get("grid"))
        self._validate_dtypes(simulation.get("dtype", DEFAULT_DTYPES))
        self._validate_solvers(simulation.get("pressure_solver"))
        self._validate_operators(simulation.get("operators"))

    def _validate_run(self, run) -> None:
        if not run or not isinstance(run, dict):
            raise ValueError("Run configuration must be provided and must be a dictionary.")
        frames, warmup, reps = run.get("frames"), run.get("warmup_frames"), run.get("repetitions")
        if not isinstance(frames, list) or not frames:
            raise ValueError("Frames must be a non-empty list of positive integers.")
        if any(not isinstance(f, int) or f <= 0 for f in frames):
            raise ValueError("Frames must be a list of positive integers.")
        if not isinstance(warmup, list) or not warmup:
            raise ValueError("Warmup frames must be a non-empty list of positive integers.")
        if any(not isinstance(f, int) or f < 0 for f in warmup):
            raise ValueError("Warmup frames must be a list of positive integers.")
        if not isinstance(reps, list) or not reps:
            raise ValueError("Repetitions must be a non-empty list of positive integers.")
        if any(not isinstance(r, int) or r <= 0 for r in reps):
            raise ValueError("Repetitions must be a list of positive integers.")

    def validate(self) -> None:
        """config.py:339-390, plus the dtype axis and operator registry."""
        if not self.name:
            raise ValueError("Suite name must be provided.")
        if not self.benchmark_definitions:
            raise ValueError("Benchmark definitions must be provided.")
        d = self.benchmark_definitions
        self._validate_simulation(d.get("simulation"))
        validate_emitters(d.get("emitters", []))
        self._validate_run(d.get("run"))
        try:
            ProfileConfig.from_dict(d.get("profiling") or {}).validate_config()
        except ValueError as e:
            raise ValueError(f"Suite config validation failed for profiler: {e}")

    # ----------------------------------------------------------- expansion
    @staticmethod
    def _solver_variants(solver: dict) -> list[dict]:
        """method -> backend -> params, in the order config.py:392-413 nests them."""
        return [{"method": m, "backend": b, "params": p}
                for m, mc in solver["method"].items()
                for b in solver["backend"]
                for p in mc["params"]]

    @staticmethod
    def _operator_variants(operators: dict) -> list[dict[str, str]]:
        names = list(operators)
        return [dict(zip(names, combo)) for combo in product(*(operators[n] for n in names))]

    def _simulation_variants(self, simulation: dict) -> list[dict]:
        """size x cell size x solver x operators (config.py:431-470), then
        dtype innermost -- a one-dtype suite keeps the reference's order."""
        grid = simulation["grid"]
        return [{"grid": {"size": size, "cell_size": cell},
                 "pressure_solver": solver,
                 "operators": ops,
                 "dtype": dtype}
                for size, cell, solver, ops, dtype in product(
                    grid["size"], grid["cell_size"],
                    self._solver_variants(simulation["pressure_solver"]),
                    self._operator_variants(simulation["operators"]),
                    simulation.get("dtype", DEFAULT_DTYPES))]

    def expand_suite(self) -> list[BenchmarkConfig]:
        """Every combination as a BenchmarkConfig named by make_benchmark_id
        (config.py:472-516)."""
        d = self.benchmark_definitions
        run = d.get("run", {})
        out: list[BenchmarkConfig] = []
        for sim, frames, warmup, reps in product(
                self._simulation_variants(d.get("simulation", {})),
                run["frames"], run["warmup_frames"], run["repetitions"]):
            cfg = BenchmarkConfig.from_dict({
                "name": "test",
                "simulation": sim,
                "emitters": d.get("emitters", []),
                "run": {"frames": frames, "warmup_frames": warmup, "repetitions": reps},
                "profiling": d.get("profiling"),
            })
            cfg.name = make_benchmark_id(len(out), cfg)
            out.append(cfg)
        return out

    # ------------------------------------------------------------------ io
    def to_dict(self) -> dict:
        return {"name": self.name, "schema_version": SCHEMA_VERSION, "type": "suite",
                "benchmark": self.benchmark_definitions}

    @classmethod
    def from_dict(cls, data: dict) -> "SuiteConfig":
        return cls(name=data["name"], benchmark_definitions=data["benchmark"])

    @classmethod
    def from_json(cls, path: str) -> "SuiteConfig":
        """A "suite" file as is; a "benchmark" file as a one-point suite."""
        with open(path) as f:
            data = json.load(f)
        kind = data.get("type")
        if kind == "suite":
            return cls.from_dict(data)
        if kind == "benchmark":
            return cls.from_benchmark_config(BenchmarkConfig.from_dict(data))
        raise ValueError(f"Invalid config type: {kind}. Expected 'suite' or 'benchmark'.")

    @classmethod
    def from_benchmark_config(cls, config: BenchmarkConfig) -> "SuiteConfig":
        """Each scalar of the benchmark becomes a one-element list
        (config.py:567-616)."""
        sim = config.simulation
        ps = sim.pressure_solver
        simulation = {
            "pressure_solver": {"method": {ps.method: {"params": [asdict(ps.params)]}},
                                "backend": [Backend(ps.backend).value]},
            "operators": {k: [Backend(v).value] for k, v in asdict(sim.operators).items()},
            "grid": {"size": [list(sim.grid.size)], "cell_size": [sim.grid.cell_size]},
            "dtype": [sim.dtype],
        }
        run = config.run
        return cls(name=config.name, benchmark_definitions={
            "simulation": simulation,
            "emitters": config.emitters or [],
            "run": {"frames": [run.frames], "warmup_frames": [run.warmup_frames],
                    "repetitions": [run.repetitions]},
            "profiling": config.profiling.to_dict(),
        })
