"""Operator call signatures (src/cfd/operators/base.py:4-139).

Arrays are torch tensors on the GPU for the ``"hip"`` backend; the argument
order and meaning are the reference's.
"""
from typing import Any, Protocol


class AdvectionScalarOperator(Protocol):
    def __call__(self, scalar_field, scalar_destination, velocity_field_u, velocity_field_v,
                 is_fluid, dt: float, cell_size: float, **kwargs) -> None: ...


class AdvectionVelocityUOperator(Protocol):
    def __call__(self, velocity_field_u, velocity_destination_u, velocity_field_v, face_mask_u,
                 dt: float, cell_size: float, **kwargs) -> None: ...


class AdvectionVelocityVOperator(Protocol):
    def __call__(self, velocity_field_v, velocity_destination_v, velocity_field_u, face_mask_v,
                 dt: float, cell_size: float, **kwargs) -> None: ...


class EnforceBoundaryConditionsOperator(Protocol):
    def __call__(self, velocity_field_u, velocity_field_v, solid_mask, **kwargs) -> None: ...


class ComputeDivergenceOperator(Protocol):
    def __call__(self, velocity_field_u, velocity_field_v, divergence, face_mask_u, face_mask_v,
                 cell_size: float, **kwargs) -> None: ...


class ApplyForcesOperator(Protocol):
    def __call__(self, velocity_field_v, density_field, temperature_field, dt: float,
                 gravity: float, buoyancy_coefficient: float, density_coefficient: float,
                 ambient_temperature: float, **kwargs) -> None: ...


class CoolTemperatureOperator(Protocol):
    def __call__(self, temperature_field, ambient_temperature: float, cooling_rate: float,
                 dt: float, **kwargs) -> None: ...


class ApplyPressureGradientOperator(Protocol):
    def __call__(self, velocity_field_u, velocity_field_v, pressure_field, face_mask_u,
                 face_mask_v, fluid_density: float, cell_size: float, dt: float,
                 **kwargs) -> None: ...


class ComputeResidualOperator(Protocol):
    def __call__(self, pressure_field, divergence_field, is_fluid, face_mask_u, face_mask_v,
                 alpha: float, reciprocal_beta: float, residual, **kwargs) -> Any: ...
