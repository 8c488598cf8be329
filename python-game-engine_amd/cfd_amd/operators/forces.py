"""Force dispatchers (src/cfd/operators/forces.py:12-60)."""
from .registry import get_operator


def apply_forces(velocity_field_v, density_field, temperature_field, dt, gravity,
                 buoyancy_coefficient, density_coefficient, ambient_temperature, *,
                 backend: str = "hip"):
    return get_operator("apply_forces", backend)(
        velocity_field_v, density_field, temperature_field, dt, gravity, buoyancy_coefficient,
        density_coefficient, ambient_temperature)


def cool_temperature(temperature_field, ambient_temperature, cooling_rate, dt, *,
                     backend: str = "hip"):
    return get_operator("cool_temperature", backend)(
        temperature_field, ambient_temperature, cooling_rate, dt)
