"""SmokeSimulation2D on the GPU (src/cfd/common/simulation.py:19-335).

Same state, constants, step order and call sites as the reference; fields are
torch tensors on the GPU in ``config.dtype``.  Differences, all deliberate:

* every operator goes through its configured backend (the reference ignores
  ``OperatorBackendConfig``, simulation.py:52, :275-321);
* when all operators are ``"hip"`` the step uses fused launches that compute
  the same values: u and v advection in one kernel, density dissipation
  folded into the density advection, boundary enforcement folded into the
  pressure-gradient kernel;
* temperature advection and cooling are skipped while the temperature field
  is known to be identically zero and the ambient temperature is zero (both
  then map zero to zero exactly).  The field stops being known zero, for the
  rest of the run, once the ambient temperature is non-zero, an emitter
  writes temperature (an emitter without a ``writes_temperature`` attribute
  is taken to write it), or the tensor is written through torch or replaced
  (its version counter or identity changes).  Writes that bypass torch (a
  raw device pointer) are not seen: set ``track_temperature = True`` for
  those; it forces the temperature passes on.

Reference quirks kept for parity: face masks are built before the boundary
ring is marked solid and the mask is not flagged dirty (simulation.py:102-113),
so the pressure solve sees every cell as fluid while boundary enforcement uses
the ring; the solve uses ``self.dt`` while the gradient uses the ``dt``
argument (simulation.py:311-317); the metrics stage recomputes the divergence
and residual every step (simulation.py:251-265).
"""
from __future__ import annotations

import numpy as np
import torch

from cfd_amd.backends.hip import operators as hip
from cfd_amd.backends.hip import solvers as hip_solvers
from cfd_amd.common.config import Backend, SimulationConfig
from cfd_amd.common.emitters import BoxEmitter
from cfd_amd.operators.advection import (advect_scalar_field, advect_velocity_u,
                                         advect_velocity_v)
from cfd_amd.operators.boundaries import enforce_boundary_conditions
from cfd_amd.operators.divergence import compute_divergence
from cfd_amd.operators.forces import apply_forces, cool_temperature
from cfd_amd.operators.pressure_gradient import apply_pressure_gradient
from cfd_amd.operators.residual import compute_residual
from cfd_amd.profiling import ProfileSession
from cfd_amd.solvers.dispatch import solve_pressure
from cfd_amd.solvers.params import DEFAULT_SOLVER_PARAMS, SolverParams

_TORCH_DTYPES = {"float32": torch.float32, "float64": torch.float64}


def calculate_alpha(cell_size: float, density: float, dt: float) -> float:
    """solvers/common.py:6-18."""
    return (-(cell_size * cell_size) * density) / dt


def calculate_reciprocal_beta() -> float:
    """solvers/common.py:20-28."""
    return 0.25


class SmokeSimulation2D:
    solver_params: SolverParams

    def __init__(self, config: SimulationConfig, profiler: ProfileSession | None = None) -> None:
        self.config = config
        self._initalise_from_config(config, profiler)

    # ------------------------------------------------------------- set-up
    def _initalise_from_config(self, config: SimulationConfig, profiler):
        self.width, self.height = config.grid.size
        self.cell_size = config.grid.cell_size
        self.pressure_solver_method = config.pressure_solver.method
        self.pressure_solver_backend = config.pressure_solver.backend
        self.solver_params = (config.pressure_solver.params
                              if config.pressure_solver.params is not None
                              else DEFAULT_SOLVER_PARAMS[self.pressure_solver_method])
        self.operators_backend = config.operators

        self.domain_width = self.width * self.cell_size
        self.domain_height = self.height * self.cell_size

        # simulation.py:59-67
        self.dt = 0.016
        self.accumulator = 0.0
        self.max_substeps = 4
        self.density_dissipation = 0.999
        self.buoyancy = 5.0
        self.gravity = -0.5
        self.fluid_density = 1.0

        self.dtype = _TORCH_DTYPES[config.dtype]
        self.device = torch.device(config.device)
        H, W = self.height, self.width
        z = dict(dtype=self.dtype, device=self.device)
        self.density = torch.zeros((H, W), **z)
        self.density_temp = torch.zeros((H, W), **z)
        self.velocity_u = torch.zeros((H, W + 1), **z)
        self.velocity_v = torch.zeros((H + 1, W), **z)
        self.velocity_u_temp = torch.zeros((H, W + 1), **z)
        self.velocity_v_temp = torch.zeros((H + 1, W), **z)
        self.pressure = torch.zeros((H, W), **z)
        self.temperature = torch.zeros((H, W), **z)
        self.temperature_temp = torch.zeros((H, W), **z)
        self.ambient_temperature = 0.0
        self.cooling_rate = 0.5
        self.divergence = torch.zeros((H, W), **z)
        self.pressure_residual = torch.zeros((H, W), **z)

        # simulation.py:102-110 -- masks from the all-fluid solid mask
        b = dict(dtype=torch.bool, device=self.device)
        self._solid_mask = torch.zeros((H, W), **b)
        self._face_mask_u = torch.zeros((H, W + 1), **b)
        self._face_mask_v = torch.zeros((H + 1, W), **b)
        self.is_fluid = torch.zeros((H, W), **b)
        hip.face_masks_from_solid(self._solid_mask, self._face_mask_u, self._face_mask_v)
        hip.is_fluid_from_masks(self._face_mask_u, self._face_mask_v, self.is_fluid)
        self._mask_dirty = False

        self._init_boundaries()  # simulation.py:113 (not flagged dirty)

        self.emitters = []
        self.debug = config.debug
        self.profiler = profiler if profiler is not None else ProfileSession.disabled()

        self.track_temperature = None  # None: automatic (see module docstring)
        self._temperature_live = False
        self._watch_temperature()
        self.compute_metrics = True
        self.last_solve_info: dict = {}
        # HIP graphs of the step's stages before and after the pressure solve
        # (see _run_stage); off by default
        self.use_graphs = False
        self._graphs: dict = {}

    def _init_boundaries(self):
        """simulation.py:126-131."""
        self._solid_mask[0, :] = True
        self._solid_mask[-1, :] = True
        self._solid_mask[:, 0] = True
        self._solid_mask[:, -1] = True

    def set_solid_mask(self, solid_mask):
        """simulation.py:133-143."""
        if tuple(solid_mask.shape) != (self.height, self.width):
            raise ValueError(f"Solid mask must have shape ({self.height}, {self.width})")
        if isinstance(solid_mask, np.ndarray):
            solid_mask = torch.from_numpy(solid_mask.astype(bool))
        # a copy, as the reference's astype(bool): later edits of the caller's
        # tensor do not reach the simulation unless set again
        self._solid_mask = solid_mask.to(device=self.device, dtype=torch.bool).clone(
            memory_format=torch.contiguous_format)
        self._mask_dirty = True

    def get_solid_mask(self):
        return self._solid_mask.clone()

    def add_emitter(self, emitter):
        self.emitters.append(emitter)

    @property
    def face_mask_u(self):
        return self._face_mask_u

    @property
    def face_mask_v(self):
        return self._face_mask_v

    def _all_hip(self) -> bool:
        ob = self.operators_backend
        return all(Backend(getattr(ob, k)) == Backend.HIP for k in vars(ob))

    def _watch_temperature(self):
        """Remember the temperature tensor and its version while it is known
        to be zero."""
        self._temperature_ref = self.temperature
        self._temperature_version = self.temperature._version

    def _temperature_active(self) -> bool:
        if self.track_temperature is not None:
            return bool(self.track_temperature)
        if not self._temperature_live and (
                self.ambient_temperature != 0.0
                or any(getattr(e, "writes_temperature", True) for e in self.emitters)
                or self.temperature is not self._temperature_ref
                or self.temperature._version != self._temperature_version):
            self._temperature_live = True
        return self._temperature_live

    # --------------------------------------------------------------- step
    def update(self, frame_dt):
        """simulation.py:166-184 fixed-dt accumulator."""
        self.accumulator += frame_dt
        substeps = 0
        while self.accumulator >= self.dt and substeps < self.max_substeps:
            self.substep(self.dt)
            self.accumulator -= self.dt
            substeps += 1
        if substeps >= self.max_substeps:
            self.accumulator = 0.0

    def substep(self, dt):
        """simulation.py:186-265."""
        prof = self.profiler
        prof.tick()
        if self._mask_dirty:
            hip.face_masks_from_solid(self._solid_mask, self._face_mask_u, self._face_mask_v)
            hip.is_fluid_from_masks(self._face_mask_u, self._face_mask_v, self.is_fluid)
            self._mask_dirty = False

        t = prof.timing
        graphs = self._graphs_usable()
        with t.scope("simulation.substep"):
            if graphs:
                self._run_stage("pre", dt, self._pre_solve)
            else:
                self._pre_solve(dt)
            with t.scope("simulation.pressure.solve"):
                self._solve_pressure()
            if not graphs:
                self._project(dt)
        if graphs:  # projection and metrics pass in one graph (profiling is off)
            self._run_stage("post", dt, self._post_solve)
        else:
            self._metrics(dt)

    def _pre_solve(self, dt):
        """simulation.py:205-237: sources, forces, transport, divergence."""
        prof = self.profiler
        t = prof.timing
        with t.scope("simulation.emitters"):
            self._apply_emitters(dt)
        with t.scope("simulation.forces"):
            self._apply_forces(dt)
        with t.scope("simulation.advection.velocity"):
            self._advect_velocity(dt)
        fused = self._all_hip()
        with t.scope("simulation.advection.density"):
            self._advect_density(dt, fuse_dissipation=fused)
        temp = self._temperature_active()
        if temp:
            with t.scope("simulation.advection.temperature"):
                self._advect_temperature(dt)
            with t.scope("simulation.cooling"):
                self._cool_temperature(dt)
        if not fused:
            with t.scope("simulation.dissipation"):
                hip.scale_field(self.density, self.density_dissipation)
        with t.scope("simulation.divergence"):
            self._compute_divergence()
        with t.scope("profiling.capture"):
            prof.metrics.capture("divergence", "before_projection", self.divergence)

    def _post_solve(self, dt):
        """simulation.py:242-265: projection, walls, the metrics pass."""
        self._project(dt)
        self._metrics(dt)

    def _project(self, dt):
        """simulation.py:242-247."""
        t = self.profiler.timing
        fused = self._all_hip()
        with t.scope("simulation.pressure.apply"):
            self._apply_pressure_gradient(dt, fuse_boundaries=fused)
        if not fused:
            with t.scope("simulation.boundaries"):
                self._enforce_boundaries()

    def _metrics(self, dt):
        """simulation.py:250-265 (outside the substep's timing scope)."""
        prof = self.profiler
        t = prof.timing
        fused = self._all_hip()
        if self.compute_metrics:
            with t.scope("profiling.metrics"):
                alpha = calculate_alpha(self.cell_size, self.fluid_density, dt)
                if fused:  # divergence and residual in one launch, same results
                    hip.divergence_residual(self.velocity_u, self.velocity_v, self.pressure,
                                            self.is_fluid, self._face_mask_u, self._face_mask_v,
                                            self.cell_size, alpha, calculate_reciprocal_beta(),
                                            self.divergence, self.pressure_residual)
                    prof.metrics.capture("divergence", "after_projection", self.divergence)
                    prof.metrics.capture("pressure", "final", self.pressure)
                else:
                    self._compute_divergence()
                    prof.metrics.capture("divergence", "after_projection", self.divergence)
                    prof.metrics.capture("pressure", "final", self.pressure)
                    compute_residual(self.pressure, self.divergence, self.is_fluid,
                                     self._face_mask_u, self._face_mask_v, alpha,
                                     calculate_reciprocal_beta(), self.pressure_residual,
                                     backend=self.operators_backend.residual)
                prof.metrics.capture("pressure_residual", "final", self.pressure_residual)
                prof.metrics.process()

    # ------------------------------------------------------------- graphs
    _BUFFERS = ("density", "density_temp", "velocity_u", "velocity_u_temp", "velocity_v",
                "velocity_v_temp", "temperature", "temperature_temp")

    def _graphs_usable(self) -> bool:
        """Graphs replace the eager stages when asked for, every operator
        runs on the hip backend (a fixed sequence of library kernels), no
        profiling is on (its scopes record events and read back) and every
        emitter can state what its launch depends on (graph_state)."""
        return (self.use_graphs and self.device.type == "cuda" and self._all_hip()
                and not self.profiler.enabled
                and all(hasattr(e, "graph_state") for e in self.emitters))

    def invalidate_graphs(self):
        """Drop the captured stages (they are captured again on next use).
        Changes the stage keys cover need no call: dt, the buffer
        assignment, the temperature passes, the emitters' states, the
        solid-mask tensor and the metrics switch."""
        self._graphs.clear()

    def _run_stage(self, stage, dt, fn):
        """Run a stage from a HIP graph of its kernels.  A stage's kernels
        and their arguments are fixed by the key below; the ping-pong swaps
        of the transported fields (simulation.py:285-298) change the buffer
        assignment every step, so each assignment has its graph, captured on
        first use (capture runs the stage's Python, including its swaps, and
        records the kernels without running them, so the graph is replayed
        once at once) and afterwards replayed with the swaps applied in
        Python."""
        key = (stage, float(dt), self._temperature_active(), self.compute_metrics,
               self._solid_mask.data_ptr(), tuple(e.graph_state() for e in self.emitters),
               tuple(getattr(self, b).data_ptr() for b in self._BUFFERS))
        entry = self._graphs.get(key)
        if entry is None:
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                fn(dt)
            entry = (g, {b: getattr(self, b) for b in self._BUFFERS})
            self._graphs[key] = entry
        g, after = entry
        g.replay()
        for b, tensor in after.items():
            setattr(self, b, tensor)

    # --------------------------------------------------------- components
    def _apply_emitters(self, dt):
        temp = self.temperature if self._temperature_active() else None
        run = []  # consecutive enabled box emitters: their boxes in one call

        def flush():
            if len(run) > 1:
                hip.box_sources(self.density, [(b.x0, b.y0, b.x1, b.y1, b.value, b.mode)
                                               for b in run])
            elif run:
                run[0].emit(self.density, self.velocity_u, self.velocity_v, temp, dt)
            run.clear()

        for e in self.emitters:
            if type(e) is BoxEmitter and e.enabled and self.density is not None:
                run.append(e)
                continue
            flush()
            e.emit(self.density, self.velocity_u, self.velocity_v, temp, dt)
        flush()

    def _apply_forces(self, dt):
        """simulation.py:273-277 (density_coefficient fixed at 1.0 there)."""
        temp = self.temperature if self._temperature_active() else None
        apply_forces(self.velocity_v, self.density, temp, dt, self.gravity, self.buoyancy, 1.0,
                     self.ambient_temperature, backend=self.operators_backend.forces)

    def _advect_velocity(self, dt):
        b = self.operators_backend.advection
        if Backend(b) == Backend.HIP:
            hip.advect_velocity(self.velocity_u, self.velocity_u_temp, self.velocity_v,
                                self.velocity_v_temp, self._face_mask_u, self._face_mask_v, dt,
                                self.cell_size)
        else:
            advect_velocity_u(self.velocity_u, self.velocity_u_temp, self.velocity_v,
                              self._face_mask_u, dt, self.cell_size, backend=b)
            advect_velocity_v(self.velocity_v, self.velocity_v_temp, self.velocity_u,
                              self._face_mask_v, dt, self.cell_size, backend=b)
        self.velocity_u, self.velocity_u_temp = self.velocity_u_temp, self.velocity_u
        self.velocity_v, self.velocity_v_temp = self.velocity_v_temp, self.velocity_v

    def _advect_density(self, dt, fuse_dissipation=False):
        if fuse_dissipation:
            hip.advect_scalar_field(self.density, self.density_temp, self.velocity_u,
                                    self.velocity_v, self.is_fluid, dt, self.cell_size,
                                    scale=self.density_dissipation)
        else:
            advect_scalar_field(self.density, self.density_temp, self.velocity_u, self.velocity_v,
                                self.is_fluid, dt, self.cell_size,
                                backend=self.operators_backend.advection)
        self.density, self.density_temp = self.density_temp, self.density

    def _advect_temperature(self, dt):
        advect_scalar_field(self.temperature, self.temperature_temp, self.velocity_u,
                            self.velocity_v, self.is_fluid, dt, self.cell_size,
                            backend=self.operators_backend.advection)
        self.temperature, self.temperature_temp = self.temperature_temp, self.temperature

    def _compute_divergence(self):
        compute_divergence(self.velocity_u, self.velocity_v, self.divergence, self._face_mask_u,
                           self._face_mask_v, self.cell_size,
                           backend=self.operators_backend.divergence)

    def _cool_temperature(self, dt):
        cool_temperature(self.temperature, self.ambient_temperature, self.cooling_rate, dt,
                         backend=self.operators_backend.forces)

    def _solve_pressure(self):
        """simulation.py:308-312 (p zeroed, solved with self.dt).  The hip
        multigrid solver takes the zeroing over (initial_guess="zero": it
        reads no p, writes every cell and reuses the rhs norm as the initial
        residual), with the same result."""
        extra = {}
        if (self.pressure_solver_method == "multigrid"
                and Backend(self.pressure_solver_backend) == Backend.HIP):
            extra["initial_guess"] = "zero"
        else:
            self.pressure.zero_()
        solve_pressure(self.pressure, self.divergence, self._face_mask_u, self._face_mask_v,
                       self.fluid_density, self.cell_size, self.dt, self, **extra)
        self.last_solve_info = dict(hip_solvers.LAST_SOLVE_INFO)

    def _apply_pressure_gradient(self, dt, fuse_boundaries=False):
        if fuse_boundaries:
            hip.apply_pressure_gradient(self.velocity_u, self.velocity_v, self.pressure,
                                        self._face_mask_u, self._face_mask_v, self.fluid_density,
                                        self.cell_size, dt, solid_mask=self._solid_mask)
        else:
            apply_pressure_gradient(self.velocity_u, self.velocity_v, self.pressure,
                                    self._face_mask_u, self._face_mask_v, self.fluid_density,
                                    self.cell_size, dt,
                                    backend=self.operators_backend.pressure_gradient)

    def _enforce_boundaries(self):
        enforce_boundary_conditions(self.velocity_u, self.velocity_v, self._solid_mask,
                                    backend=self.operators_backend.boundaries)

    # -------------------------------------------------------------- access
    def get_density_field(self):
        return self.density

    def reset(self):
        for f in (self.density, self.velocity_u, self.velocity_v, self.temperature,
                  self.pressure, self.divergence):
            f.zero_()
        self.accumulator = 0.0
        self._temperature_live = False
        self._watch_temperature()

