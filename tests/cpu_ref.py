"""Backend ``"cpu_ref"``: the oracle registered behind the same registries as
the ``"hip"`` backend, for every operator key and every solver key including
the serial ``gauss_seidel``.  Test
infrastructure: the product package never registers or imports it."""
from __future__ import annotations

import oracle as O

BACKEND = "cpu_ref"

_OPS = {
    "advection_scalar": O.advect_scalar_field,
    "advection_velocity_u": O.advect_velocity_u,
    "advection_velocity_v": O.advect_velocity_v,
    "enforce_boundary_conditions": O.enforce_boundary_conditions,
    "compute_divergence": O.compute_divergence,
    "apply_forces": O.apply_forces,
    "cool_temperature": O.cool_temperature,
    "apply_pressure_gradient": O.apply_pressure_gradient,
    "compute_residual": O.compute_residual,
}

_SOLVERS = {
    "jacobi": O.jacobi_pressure_solver,
    "gauss_seidel": O.gauss_seidel_pressure_solver,
    "rb_sor_gauss_seidel": O.red_black_sor_gauss_seidel_pressure_solver,
    "multigrid": O.multigrid_solver,
}


def register() -> None:
    """Idempotent registration of the oracle as backend "cpu_ref"."""
    from cfd_amd.operators.registry import OPERATOR_REGISTRY, register_operator
    from cfd_amd.solvers.registry import SOLVER_REGISTRY, register_solver

    for key, fn in _OPS.items():
        if (key, BACKEND) not in OPERATOR_REGISTRY:
            register_operator(key, BACKEND, fn)
    for key, fn in _SOLVERS.items():
        if (key, BACKEND) not in SOLVER_REGISTRY:
            register_solver(key, BACKEND, fn)
