"""Synthetic smoke-plume workloads (host-side numpy; identical arrays feed every path).

Each builder returns a :class:`Scene`: grid, time step, initial MAC velocity,
box sources and solver settings.  Generation is seeded with
``numpy.random.default_rng(seed)`` (PCG64) and uses a fixed draw order, so a
scene is reproducible on any host.

Workloads (sizes as the benchmark suite names them):

* ``plume_128_jacobi``   128x128, zero velocity, one 16x8 box at (0.5, 0.1),
  Jacobi 200 iterations, 100 steps.
* ``plume_512_rbsor``    512x512, 4 boxes of 12x12 with strengths U(0.5,1.5)
  in the lower third, velocity = curl of 32 Gaussian blobs, RB-SOR
  omega=1.9 to 1e-6 relative residual, 300 steps.
* ``plume_2048_mg``      2048x2048, same generator scaled (48x48 boxes),
  multigrid V(2,2) RB-GS, 8 levels to 16x16, coarsest 50 sweeps, 1e-8.
* ``uneven_scenes``      96 scenes, N log-uniform in [64,1024] rounded to 16,
  aspect 1:1 or 2:1, one 8..32 box each, multigrid to 1e-6, 20 steps.
* ``diff_rollout_256``   256x256 differentiable rollout, 64 steps.

Physical constants follow the reference defaults (``simulation.py:59-67``)
overridden by the workload: buoyancy f_y = 1.0*density is
``density_coefficient=1, gravity=-1``; dissipation 1 (no fade); no
temperature.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


@dataclass
class BoxSource:
    """density := value on rows [y0, y1) and columns [x0, x1) every step."""

    x0: int
    y0: int
    x1: int
    y1: int
    value: float = 1.0


@dataclass
class Scene:
    name: str
    H: int
    W: int
    cell_size: float
    dt: float
    steps: int
    u0: np.ndarray
    v0: np.ndarray
    boxes: list = field(default_factory=list)
    solver: str = "multigrid"
    solver_params: dict = field(default_factory=dict)
    density_coefficient: float = 1.0
    gravity: float = -1.0
    buoyancy: float = 5.0
    dissipation: float = 1.0
    seed: int = 0

    @property
    def cells(self) -> int:
        return self.H * self.W


def box_from_centre(cx: float, cy: float, w_cells: int, h_cells: int, h: float,
                    H: int, W: int, value: float = 1.0) -> BoxSource:
    """Box of w_cells x h_cells cells centred at physical (cx, cy).

    lower-left = floor(centre/h - size/2 + 0.5), clipped to the grid."""
    x0 = int(math.floor(cx / h - w_cells / 2.0 + 0.5))
    y0 = int(math.floor(cy / h - h_cells / 2.0 + 0.5))
    x0 = min(max(x0, 0), W - w_cells)
    y0 = min(max(y0, 0), H - h_cells)
    return BoxSource(x0, y0, x0 + w_cells, y0 + h_cells, float(value))


def outer_wall_face_masks(H: int, W: int):
    """Face masks of the reference default simulation (simulation.py:102-108).

    The masks are built from an all-fluid solid mask before the boundary ring
    is marked solid (simulation.py:113, :126-131, mask not flagged dirty), so
    every interior face is open and only the domain's outer faces are closed.
    """
    fmu = np.zeros((H, W + 1), dtype=bool)
    fmv = np.zeros((H + 1, W), dtype=bool)
    fmu[:, 1:-1] = True
    fmv[1:-1, :] = True
    return fmu, fmv


def streamfunction_velocity(H: int, W: int, h: float, centres: np.ndarray,
                            amps: np.ndarray, radii: np.ndarray):
    """MAC velocity = discrete curl of psi = sum_k a_k exp(-|x-c_k|^2/(2 r_k^2)).

    psi lives on cell corners (H+1, W+1) at (i*h, j*h);
    u[j,i] = (psi[j+1,i]-psi[j,i])/h, v[j,i] = -(psi[j,i+1]-psi[j,i])/h.
    The result is exactly divergence-free up to rounding; outer faces are
    then zeroed (closed walls)."""
    xs = np.arange(W + 1, dtype=np.float64) * h
    ys = np.arange(H + 1, dtype=np.float64) * h
    psi = np.zeros((H + 1, W + 1), dtype=np.float64)
    for (cx, cy), a, r in zip(centres, amps, radii):
        gx = np.exp(-((xs - cx) ** 2) / (2.0 * r * r))
        gy = np.exp(-((ys - cy) ** 2) / (2.0 * r * r))
        psi += a * np.outer(gy, gx)
    u = (psi[1:, :] - psi[:-1, :]) / h
    v = -(psi[:, 1:] - psi[:, :-1]) / h
    u = np.ascontiguousarray(u)
    v = np.ascontiguousarray(v)
    fmu, fmv = outer_wall_face_masks(H, W)
    u[~fmu] = 0.0
    v[~fmv] = 0.0
    return u, v


def _blob_scene(name, N_h, N_w, seed, box_cells, n_boxes, steps, solver,
                solver_params, dt=0.005, n_blobs=32):
    rng = np.random.default_rng(seed)
    h = 1.0 / N_h
    Lx, Ly = N_w * h, N_h * h
    # draw order: centres, amplitudes, radii, source positions, strengths
    centres = rng.uniform(0.0, 1.0, size=(n_blobs, 2)) * np.array([Lx, Ly])
    amps = rng.normal(0.0, 1.0, size=n_blobs)
    radii = rng.uniform(0.02, 0.1, size=n_blobs)
    src = rng.uniform([0.05, 0.05], [0.95, 1.0 / 3.0], size=(n_boxes, 2)) * np.array([Lx, Ly])
    strengths = rng.uniform(0.5, 1.5, size=n_boxes)
    u0, v0 = streamfunction_velocity(N_h, N_w, h, centres, amps, radii)
    boxes = [box_from_centre(cx, cy, box_cells, box_cells, h, N_h, N_w, s)
             for (cx, cy), s in zip(src, strengths)]
    return Scene(name=name, H=N_h, W=N_w, cell_size=h, dt=dt, steps=steps, u0=u0,
                 v0=v0, boxes=boxes, solver=solver, solver_params=solver_params,
                 seed=seed)


def plume_128_jacobi(N: int = 128, steps: int = 100) -> Scene:
    h = 1.0 / N
    box = box_from_centre(0.5, 0.1, 16 * N // 128, 8 * N // 128, h, N, N, 1.0)
    return Scene(name=f"plume_{N}_jacobi", H=N, W=N, cell_size=h, dt=0.005,
                 steps=steps, u0=np.zeros((N, N + 1)), v0=np.zeros((N + 1, N)),
                 boxes=[box], solver="jacobi", solver_params={"iterations": 200})


def plume_512_rbsor(N: int = 512, steps: int = 300, seed: int = 2) -> Scene:
    box = max(2, int(round(12 * N / 512)))
    return _blob_scene(f"plume_{N}_rbsor", N, N, seed, box, 4, steps,
                       "rb_sor_gauss_seidel",
                       {"omega": 1.9, "tolerance": 1e-6, "check_every": 20,
                        "iterations": 200000})


def mg_levels_to(N: int, coarsest: int = 16) -> int:
    levels, n = 1, N
    while n > coarsest and n % 2 == 0:
        n //= 2
        levels += 1
    return levels


def plume_2048_mg(N: int = 2048, steps: int = 50, seed: int = 3) -> Scene:
    box = max(2, int(round(48 * N / 2048)))
    return _blob_scene(f"plume_{N}_mg", N, N, seed, box, 4, steps, "multigrid",
                       {"smoother_type": "rb_sor_gauss_seidel", "smoother_omega": 1.0,
                        "num_levels": mg_levels_to(N), "smoother_iterations": 2,
                        "coarsest_iterations": 50, "v_cycles": 400,
                        "coarse_scale": "rediscretized", "tolerance": 1e-8})


def uneven_scenes(count: int = 96, steps: int = 20, base_seed: int = 4000,
                  n_min: int = 64, n_max: int = 1024) -> list:
    scenes = []
    for s in range(count):
        rng = np.random.default_rng(base_seed + s)
        N = int(16 * round(math.exp(rng.uniform(math.log(n_min), math.log(n_max))) / 16))
        N = max(N, 16)
        wide = rng.uniform() < 0.5
        H, W = (N, 2 * N) if wide else (N, N)
        side = int(rng.integers(8, 33))
        sc = _blob_scene(f"scene_{s}", H, W, int(rng.integers(0, 2**31 - 1)), side, 1,
                         steps, "multigrid",
                         {"smoother_type": "rb_sor_gauss_seidel", "smoother_omega": 1.0,
                          "num_levels": mg_levels_to(min(H, W)), "smoother_iterations": 2,
                          "coarsest_iterations": 50, "v_cycles": 400,
                          "coarse_scale": "rediscretized", "tolerance": 1e-6,
                          "coarsest_omega": -1.0})
        scenes.append(sc)
    return scenes


def diff_rollout_256(N: int = 256, steps: int = 64, seed: int = 5) -> Scene:
    h = 1.0 / N
    return Scene(name=f"diff_{N}", H=N, W=N, cell_size=h, dt=0.005, steps=steps,
                 u0=np.zeros((N, N + 1)), v0=np.zeros((N + 1, N)),
                 boxes=[box_from_centre(0.5, 0.2, max(2, N // 16), max(2, N // 16), h, N, N, 1.0)],
                 solver="multigrid",
                 solver_params={"smoother_type": "rb_sor_gauss_seidel", "smoother_omega": 1.0,
                                "num_levels": mg_levels_to(N), "smoother_iterations": 2,
                                "coarsest_iterations": 50, "v_cycles": 400,
                                "coarse_scale": "rediscretized", "tolerance": 1e-8},
                 seed=seed)


def gaussian_target(H: int, W: int, h: float, cx: float = 0.7, cy: float = 0.7,
                    sigma: float = 0.08, peak: float = 1.0) -> np.ndarray:
    x = (np.arange(W) + 0.5) * h
    y = (np.arange(H) + 0.5) * h
    gx = np.exp(-((x - cx) ** 2) / (2 * sigma * sigma))
    gy = np.exp(-((y - cy) ** 2) / (2 * sigma * sigma))
    return peak * np.outer(gy, gx)


WORKLOADS = {
    "plume_128_jacobi": plume_128_jacobi,
    "plume_512_rbsor": plume_512_rbsor,
    "plume_2048_mg": plume_2048_mg,
}


def make_simulation(scene: Scene, dtype: str = "float64", device: str = "cuda",
                    solver_overrides: dict | None = None):
    """SmokeSimulation2D for a scene: grid, constants (buoyancy f_y = 1.0*rho,
    no dissipation), initial velocity and box sources."""
    import torch

    from cfd_amd.common.config import (Backend, GridConfig, PressureSolverConfig,
                                       SimulationConfig)
    from cfd_amd.common.emitters import BoxEmitter
    from cfd_amd.common.simulation import SmokeSimulation2D
    from cfd_amd.solvers.params import SOLVER_PARAM_TYPES

    params = dict(scene.solver_params)
    if solver_overrides:
        params.update(solver_overrides)
    ptype = SOLVER_PARAM_TYPES[scene.solver]
    names = set(ptype.__dataclass_fields__)
    cfg = SimulationConfig(
        grid=GridConfig(size=(scene.W, scene.H), cell_size=scene.cell_size),
        pressure_solver=PressureSolverConfig(
            method=scene.solver, backend=Backend.HIP,
            params=ptype(**{k: v for k, v in params.items() if k in names})),
        dtype=dtype, device=device)
    sim = SmokeSimulation2D(cfg)
    sim.dt = scene.dt
    sim.gravity = scene.gravity
    sim.density_dissipation = scene.dissipation
    tdt = sim.dtype
    sim.velocity_u.copy_(torch.from_numpy(scene.u0).to(device=sim.device, dtype=tdt))
    sim.velocity_v.copy_(torch.from_numpy(scene.v0).to(device=sim.device, dtype=tdt))
    for b in scene.boxes:
        sim.add_emitter(BoxEmitter.from_box(b))
    return sim
