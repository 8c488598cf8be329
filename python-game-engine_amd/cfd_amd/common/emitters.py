"""Sources: the reference's circular Emitter2D (src/cfd/common/emitters.py:101-313)
and a box source the plume workloads use.

``Emitter2D.emit`` adds ``value * falloff(distance) * dt`` to density,
temperature and the u / v face arrays (as plain arrays, at cell-index
coordinates without the +0.5 offset -- emitters.py:171-178, :202-221).
"""
from __future__ import annotations

from cfd_amd.backends.hip import operators as hip

EMITTER_APPROACHES = ["linear", "gaussian", "quadratic", "hard_cutoff"]


class Emitter2D:
    def __init__(self, position, radius, density_value=1.0, temperature_value=0.0,
                 velocity_value=(0.0, 0.0), approach="linear", numba_optimized=False):
        if approach not in EMITTER_APPROACHES:
            raise ValueError(f"Unknown emission approach: {approach}")
        self.position = position
        self.radius = radius
        self.approach = approach
        self.density_value = density_value
        self.temperature_value = temperature_value
        self.velocity_value = velocity_value
        self.enabled = True
        self.numba_optimized = numba_optimized  # kept for config compatibility

    def _apply_scalar_field(self, field, value, dt):
        hip.emit_scalar(field, value, dt, self.position[0], self.position[1], self.radius,
                        self.approach)

    def emit(self, density_field, velocity_u_field, velocity_v_field, temperature_field, dt):
        """emitters.py:202-221."""
        if not self.enabled:
            return
        if density_field is not None and self.density_value != 0.0:
            self._apply_scalar_field(density_field, self.density_value, dt)
        if temperature_field is not None and self.temperature_value != 0.0:
            self._apply_scalar_field(temperature_field, self.temperature_value, dt)
        if velocity_u_field is not None and self.velocity_value[0] != 0.0:
            self._apply_scalar_field(velocity_u_field, self.velocity_value[0], dt)
        if velocity_v_field is not None and self.velocity_value[1] != 0.0:
            self._apply_scalar_field(velocity_v_field, self.velocity_value[1], dt)

    @property
    def writes_temperature(self) -> bool:
        return self.enabled and self.temperature_value != 0.0

    def graph_state(self):
        """Everything emit()'s launches depend on besides the fields and dt
        (a captured step stage is replayed only while this is unchanged)."""
        return ("emitter2d", self.enabled, tuple(map(float, self.position)), float(self.radius),
                self.approach, float(self.density_value), float(self.temperature_value),
                tuple(map(float, self.velocity_value)))

    def set_enabled(self, enabled):
        self.enabled = enabled

    def set_position(self, position):
        self.position = position

    def set_radius(self, radius):
        self.radius = radius

    def set_density_value(self, v):
        self.density_value = v

    def set_temperature_value(self, v):
        self.temperature_value = v

    def set_velocity_value(self, v):
        self.velocity_value = v

    def to_dict(self):
        return {"position": self.position, "radius": self.radius,
                "density_value": self.density_value, "temperature_value": self.temperature_value,
                "velocity_value": self.velocity_value, "approach": self.approach,
                "numba_optimized": self.numba_optimized}

    @classmethod
    def from_dict(cls, data):
        return cls(position=data.get("position", (0, 0)), radius=data.get("radius", 5.0),
                   density_value=data.get("density_value", 1.0),
                   temperature_value=data.get("temperature_value", 0.0),
                   velocity_value=data.get("velocity_value", (0.0, 0.0)),
                   approach=data.get("approach", "linear"),
                   numba_optimized=data.get("numba_optimized", False))


class BoxEmitter:
    """density[y0:y1, x0:x1] = value ("set") or += value ("add") each step."""

    writes_temperature = False

    def __init__(self, x0, y0, x1, y1, value=1.0, mode="set"):
        self.x0, self.y0, self.x1, self.y1 = int(x0), int(y0), int(x1), int(y1)
        self.value = float(value)
        self.mode = mode
        self.enabled = True

    @classmethod
    def from_box(cls, box, mode="set"):
        return cls(box.x0, box.y0, box.x1, box.y1, box.value, mode)

    def graph_state(self):
        """What emit()'s launch depends on besides the field (Emitter2D.graph_state)."""
        return ("box", self.enabled, self.x0, self.y0, self.x1, self.y1, self.value, self.mode)

    def emit(self, density_field, velocity_u_field, velocity_v_field, temperature_field, dt):
        if self.enabled and density_field is not None:
            hip.box_source(density_field, self.x0, self.y0, self.x1, self.y1, self.value,
                           self.mode)
