"""Pressure solver call signature (src/cfd/solvers/base.py:5-19)."""
from typing import Protocol


class PressureSolver(Protocol):
    def __call__(self, pressure, divergence, face_mask_u, face_mask_v, fluid_density: float,
                 cell_size: float, dt: float, **solver_kwargs) -> None: ...
