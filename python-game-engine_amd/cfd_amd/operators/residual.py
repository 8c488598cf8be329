"""Residual dispatcher (src/cfd/operators/residual.py:9-36)."""
from .registry import get_operator


def compute_residual(pressure_field, divergence_field, is_fluid, face_mask_u, face_mask_v, alpha,
                     reciprocal_beta, residual, *, backend: str = "hip"):
    return get_operator("compute_residual", backend)(
        pressure_field, divergence_field, is_fluid, face_mask_u, face_mask_v, alpha,
        reciprocal_beta, residual)
