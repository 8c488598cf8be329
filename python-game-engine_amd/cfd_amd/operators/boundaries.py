"""Boundary dispatcher (src/cfd/operators/boundaries.py:16-33)."""
from .registry import get_operator


def enforce_boundary_conditions(velocity_field_u, velocity_field_v, solid_mask, *,
                                backend: str = "hip"):
    return get_operator("enforce_boundary_conditions", backend)(
        velocity_field_u, velocity_field_v, solid_mask)
