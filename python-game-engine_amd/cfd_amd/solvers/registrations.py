"""Registers the HIP solvers under the reference's method keys
(src/cfd/solvers/registrations.py:17-20), backend ``"hip"``: jacobi,
gauss_seidel (the lexicographic sweep as a wavefront), rb_sor_gauss_seidel
and multigrid."""
from .registry import register_solver

_initialized = False


def register_solvers() -> None:
    global _initialized
    if _initialized:
        return
    from cfd_amd.backends.hip import solvers as hip

    register_solver("jacobi", "hip", hip.jacobi_pressure_solver)
    register_solver("gauss_seidel", "hip", hip.gauss_seidel_pressure_solver)
    register_solver("rb_sor_gauss_seidel", "hip", hip.red_black_sor_gauss_seidel_pressure_solver)
    register_solver("multigrid", "hip", hip.MultigridSolver())
    _initialized = True
