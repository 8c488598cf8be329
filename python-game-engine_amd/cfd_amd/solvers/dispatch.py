"""solve_pressure (src/cfd/solvers/dispatch.py:5-12): looks the solver up by
(params.pressure_solver_method, params.pressure_solver_backend) and passes the
solver params dataclass as keyword arguments, plus ``profiler=``."""
from dataclasses import asdict

from .registry import get_solver


def solve_pressure(pressure, divergence, face_mask_u, face_mask_v, fluid_density, cell_size, dt,
                   params, **overrides):
    """``overrides`` (new) are extra solver keyword arguments, applied over
    the params dataclass's."""
    solver_fn = get_solver(params.pressure_solver_method, params.pressure_solver_backend)
    solver_kwargs = asdict(params.solver_params)
    solver_kwargs.update(overrides)
    solver_fn(pressure, divergence, face_mask_u, face_mask_v, fluid_density, cell_size, dt,
              profiler=getattr(params, "profiler", None), **solver_kwargs)
