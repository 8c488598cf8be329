"""Divergence dispatcher (src/cfd/operators/divergence.py:9-32)."""
from .registry import get_operator


def compute_divergence(velocity_field_u, velocity_field_v, divergence, face_mask_u, face_mask_v,
                       cell_size, *, backend: str = "hip"):
    return get_operator("compute_divergence", backend)(
        velocity_field_u, velocity_field_v, divergence, face_mask_u, face_mask_v, cell_size)
