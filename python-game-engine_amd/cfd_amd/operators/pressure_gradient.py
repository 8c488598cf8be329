"""Pressure-gradient dispatcher (src/cfd/operators/pressure_gradient.py:9-36)."""
from .registry import get_operator


def apply_pressure_gradient(velocity_field_u, velocity_field_v, pressure_field, face_mask_u,
                            face_mask_v, fluid_density, cell_size, dt, *, backend: str = "hip"):
    return get_operator("apply_pressure_gradient", backend)(
        velocity_field_u, velocity_field_v, pressure_field, face_mask_u, face_mask_v,
        fluid_density, cell_size, dt)
