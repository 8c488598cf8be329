"""Registers the HIP solvers under the reference's method keys
(src/cfd/solvers/registrations.py:17-20), backend ``"hip"``.

The reference's serial lexicographic ``gauss_seidel`` has no GPU form and is
not registered for ``"hip"``."""
from .registry import register_solver

_initialized = False


def register_solvers() -> None:
    global _initialized
    if _initialized:
        return
    from cfd_amd.backends.hip import solvers as hip

    register_solver("jacobi", "hip", hip.jacobi_pressure_solver)
    register_solver("rb_sor_gauss_seidel", "hip", hip.red_black_sor_gauss_seidel_pressure_solver)
    register_solver("multigrid", "hip", hip.MultigridSolver())
    _initialized = True
