"""Simulation configuration (src/cfd/common/config.py:7-163).

Same dataclasses and validation as the reference, with two changes:
``Backend`` is a ``str`` Enum (Python 3.10 has no StrEnum) gaining ``HIP``,
the default everywhere; ``SimulationConfig`` gains ``dtype`` ("float32" like
the reference's arrays, or "float64") and ``device``.
"""
from __future__ import annotations

import json
from dataclasses import dataclass, field
from enum import Enum

from cfd_amd.solvers.params import SOLVER_PARAM_TYPES, JacobiParams, SolverParams


class Backend(str, Enum):
    NUMBA = "numba"
    CPP = "cpp"
    CUDA = "cuda"
    HIP = "hip"

    def __str__(self) -> str:
        return self.value


@dataclass
class GridConfig:
    size: tuple  # (width, height), as the reference (simulation.py:45)
    cell_size: float


@dataclass
class OperatorBackendConfig:
    advection: Backend = Backend.HIP
    boundaries: Backend = Backend.HIP
    divergence: Backend = Backend.HIP
    forces: Backend = Backend.HIP
    pressure_gradient: Backend = Backend.HIP
    residual: Backend = Backend.HIP


@dataclass
class PressureSolverConfig:
    method: str = "jacobi"
    backend: Backend = Backend.HIP
    params: SolverParams = field(default_factory=JacobiParams)


_DTYPES = ("float32", "float64")


@dataclass
class SimulationConfig:
    grid: GridConfig
    pressure_solver: PressureSolverConfig = field(default_factory=PressureSolverConfig)
    operators: OperatorBackendConfig = field(default_factory=OperatorBackendConfig)
    debug: bool = False
    dtype: str = "float32"
    device: str = "cuda"

    def to_dict(self) -> dict:
        return {
            "grid": {"size": list(self.grid.size), "cell_size": self.grid.cell_size},
            "pressure_solver": {
                "method": self.pressure_solver.method,
                "backend": Backend(self.pressure_solver.backend).value,
                "params": dict(self.pressure_solver.params.__dict__),
            },
            "operators": {k: Backend(v).value for k, v in vars(self.operators).items()},
            "dtype": self.dtype,
        }

    def to_json(self, path: str) -> None:
        with open(path, "w") as f:
            json.dump(self.to_dict(), f, indent=4)

    @classmethod
    def from_dict(cls, data: dict) -> "SimulationConfig":
        g = data.get("grid", {})
        grid = GridConfig(size=tuple(g.get("size", (16, 16))), cell_size=g.get("cell_size", 1.0))
        s = data.get("pressure_solver", {})
        method = s.get("method", "jacobi")
        ptype = SOLVER_PARAM_TYPES.get(method, JacobiParams)
        return cls(
            grid=grid,
            pressure_solver=PressureSolverConfig(method=method,
                                                 backend=Backend(s.get("backend", "hip")),
                                                 params=ptype(**s.get("params", {}))),
            operators=OperatorBackendConfig(
                **{k: Backend(v) for k, v in data.get("operators", {}).items()}),
            dtype=data.get("dtype", "float32"),
        )

    @classmethod
    def from_json(cls, path: str) -> "SimulationConfig":
        with open(path) as f:
            return cls.from_dict(json.load(f))

    def validate_config(self) -> None:
        """Raises ValueError like config.py:130-163."""
        if not (isinstance(self.grid.size, tuple) and len(self.grid.size) == 2):
            raise ValueError("Grid size must be a tuple of two integers.")
        if not all(isinstance(d, int) and d > 0 for d in self.grid.size):
            raise ValueError("Grid dimensions must be positive integers.")
        if not isinstance(self.grid.cell_size, (int, float)) or self.grid.cell_size <= 0:
            raise ValueError("Cell size must be a positive number.")
        if self.pressure_solver.method not in SOLVER_PARAM_TYPES:
            raise ValueError(f"Unsupported pressure solver method: {self.pressure_solver.method}")
        if not isinstance(self.pressure_solver.backend, Backend):
            raise ValueError(f"Unsupported pressure solver backend: {self.pressure_solver.backend}")
        ptype = SOLVER_PARAM_TYPES[self.pressure_solver.method]
        if not isinstance(self.pressure_solver.params, ptype):
            raise ValueError(f"Pressure solver params must be of type {ptype.__name__} for "
                             f"method {self.pressure_solver.method}")
        if not isinstance(self.operators, OperatorBackendConfig):
            raise ValueError("Operators must be an instance of OperatorBackendConfig.")
        if not all(isinstance(getattr(self.operators, op), Backend) for op in vars(self.operators)):
            raise ValueError("All operator backends must be instances of Backend.")
        if not isinstance(self.debug, bool):
            raise ValueError("Debug flag must be a boolean.")
        if self.dtype not in _DTYPES:
            raise ValueError(f"dtype must be one of {_DTYPES}")
