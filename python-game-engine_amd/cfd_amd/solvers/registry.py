"""(method, backend) -> pressure solver registry (src/cfd/solvers/registry.py:3-33)."""
from .base import PressureSolver

SOLVER_REGISTRY: dict[tuple[str, str], PressureSolver] = {}


def register_solver(method: str, backend: str, fn) -> None:
    key = (method, str(backend))
    if key in SOLVER_REGISTRY:
        raise ValueError(f"Pressure solver already registered: method={method!r}, backend={backend!r}")
    SOLVER_REGISTRY[key] = fn


def get_solver(method: str, backend: str):
    fn = SOLVER_REGISTRY.get((method, str(backend)))
    if fn is None:
        available = [f"{m}/{b}" for m, b in SOLVER_REGISTRY]
        raise ValueError(f"Unknown pressure solver method={method!r}, backend={backend!r}. "
                         f"Available: {available}")
    return fn
