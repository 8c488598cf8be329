"""Registers the HIP operators under the reference's operation keys
(src/cfd/operators/registrations.py:21-29), backend ``"hip"``."""
from .registry import register_operator

_initialized = False

OPERATION_KEYS = (
    "advection_scalar",
    "advection_velocity_u",
    "advection_velocity_v",
    "enforce_boundary_conditions",
    "compute_divergence",
    "apply_forces",
    "cool_temperature",
    "apply_pressure_gradient",
    "compute_residual",
)


def register_operators() -> None:
    global _initialized
    if _initialized:
        return
    from cfd_amd.backends.hip import operators as hip

    register_operator("advection_scalar", "hip", hip.advect_scalar_field)
    register_operator("advection_velocity_u", "hip", hip.advect_velocity_u)
    register_operator("advection_velocity_v", "hip", hip.advect_velocity_v)
    register_operator("enforce_boundary_conditions", "hip", hip.enforce_boundary_conditions)
    register_operator("compute_divergence", "hip", hip.compute_divergence)
    register_operator("apply_forces", "hip", hip.apply_forces)
    register_operator("cool_temperature", "hip", hip.cool_temperature)
    register_operator("apply_pressure_gradient", "hip", hip.apply_pressure_gradient)
    register_operator("compute_residual", "hip", hip.compute_residual)
    _initialized = True
