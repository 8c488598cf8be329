"""Advection dispatchers (src/cfd/operators/advection.py:13-88)."""
from .registry import get_operator


def advect_scalar_field(field_src, field_dst, velocity_u, velocity_v, is_fluid, dt, cell_size,
                        *, backend: str = "hip"):
    return get_operator("advection_scalar", backend)(
        field_src, field_dst, velocity_u, velocity_v, is_fluid, dt, cell_size)


def advect_velocity_u(velocity_u_src, velocity_u_dst, velocity_v, face_mask_u, dt, cell_size,
                      *, backend: str = "hip"):
    return get_operator("advection_velocity_u", backend)(
        velocity_u_src, velocity_u_dst, velocity_v, face_mask_u, dt, cell_size)


def advect_velocity_v(velocity_v_src, velocity_v_dst, velocity_u, face_mask_v, dt, cell_size,
                      *, backend: str = "hip"):
    return get_operator("advection_velocity_v", backend)(
        velocity_v_src, velocity_v_dst, velocity_u, face_mask_v, dt, cell_size)
