"""Differentiable smoke rollouts on the GPU.

The step differentiated is SmokeSimulation2D.substep
(``src/cfd/common/simulation.py:186-249``) of the plume workloads, with
density ``rho``, face velocities ``u`` (H x (W+1)) and ``v`` ((H+1) x W):

1. box source      rho0 = rho;  rho0[box] = s[n]          (strength s[n] is a control)
2. buoyancy        v1 = v + dt * (-dc * g) * face_avg(rho0)
3. advect u, v     (u2, v2) = A_vel(u, v1)                (both by the old velocity (u, v1))
4. advect rho      rho2 = A_rho(rho0; u2, v2)             (by the new, unprojected velocity)
5. project         div = D(u2, v2); p = S(div); (u3, v3) = B M ((u2, v2) - c G p)

with the reference's masks: face masks M from the all-fluid solid mask
(outer faces closed), boundary enforcement B on the solid ring, and the
multigrid solve S run to a relative-residual tolerance from p = 0.  The one
departure from the simulation is the buoyancy force, kept in the field dtype
rather than rounded through the reference's float32 scratch
(``forces.py:33``), whose rounding is a step function with zero derivative.

Every step is a ``torch.autograd.Function`` whose backward runs the HIP
adjoint kernels (transposed semi-Lagrangian gather, buoyancy, box) and the
projection's adjoint, which is the same divergence / solve / gradient applied
to B u_bar (the masked Poisson operator is symmetric).  Rollouts are
checkpointed every ``checkpoint_every`` steps with
``torch.utils.checkpoint``: only segment inputs are kept and each segment's
forward is recomputed during the backward pass.

Precision: float64 runs everything in FP64; float32 keeps fields, transport
and adjoints in FP32 while the Poisson solves run FP32 V-cycles inside FP64
defect correction on an FP64 copy of the divergence, so that the same
relative tolerance is reachable.
"""
from __future__ import annotations

from collections import deque

import numpy as np
import torch
from torch.utils.checkpoint import checkpoint

from cfd_amd.backends.hip import adjoint as adj
from cfd_amd.backends.hip import operators as hip
from cfd_amd.backends.hip.solvers import LAST_SOLVE_INFO, MultigridSolver


class DifferentiableSmoke:
    """One scene's differentiable step (see the module docstring).

    ``box`` = (x0, y0, x1, y1) in cells.  ``adjoint_mode`` is "gather"
    (deterministic) or "scatter" (atomics).  ``warm_start_adjoint`` starts
    each adjoint Poisson solve from the previous one's solution (the solve
    is run to the same tolerance either way).
    """

    def __init__(self, H: int, W: int, cell_size: float, dt: float, box, *,
                 dtype=torch.float64, device="cuda", gravity: float = -1.0,
                 density_coefficient: float = 1.0, fluid_density: float = 1.0,
                 num_levels: int | None = None, tolerance: float = 1e-8, v_cycles: int = 400,
                 smoother_iterations: int = 2, coarsest_iterations: int = 50,
                 adjoint_mode: str = "gather", warm_start_adjoint: bool = True):
        if dtype not in (torch.float64, torch.float32):
            raise ValueError(f"dtype must be torch.float64 or torch.float32, got {dtype}")
        adj._mode(adjoint_mode)
        self.H, self.W = int(H), int(W)
        self.cell_size = float(cell_size)
        self.dt = float(dt)
        self.box = tuple(int(b) for b in box)
        self.dtype = dtype
        self.device = torch.device(device)
        self.gravity = float(gravity)
        self.density_coefficient = float(density_coefficient)
        self.fluid_density = float(fluid_density)
        self.adjoint_mode = adjoint_mode
        self.warm_start_adjoint = warm_start_adjoint
        if num_levels is None:
            num_levels, n = 1, min(self.H, self.W)
            while n > 16 and n % 2 == 0:
                n //= 2
                num_levels += 1
        self.solver_kwargs = dict(v_cycles=int(v_cycles), smoother_type="rb_sor_gauss_seidel",
                                  num_levels=int(num_levels),
                                  smoother_iterations=int(smoother_iterations),
                                  coarsest_iterations=int(coarsest_iterations),
                                  smoother_omega=1.0, coarse_scale="rediscretized",
                                  tolerance=float(tolerance),
                                  precision="native" if dtype == torch.float64 else "mixed")
        Hh, Ww = self.H, self.W
        b = dict(dtype=torch.bool, device=self.device)
        # simulation.py:102-113: masks from the all-fluid solid mask, then the ring
        solid0 = torch.zeros((Hh, Ww), **b)
        self.face_mask_u = torch.zeros((Hh, Ww + 1), **b)
        self.face_mask_v = torch.zeros((Hh + 1, Ww), **b)
        self.is_fluid = torch.zeros((Hh, Ww), **b)
        hip.face_masks_from_solid(solid0, self.face_mask_u, self.face_mask_v)
        hip.is_fluid_from_masks(self.face_mask_u, self.face_mask_v, self.is_fluid)
        self.solid = torch.zeros((Hh, Ww), **b)
        self.solid[0, :] = True
        self.solid[-1, :] = True
        self.solid[:, 0] = True
        self.solid[:, -1] = True
        f64 = dict(dtype=torch.float64, device=self.device)
        self._p64 = torch.zeros((Hh, Ww), **f64)
        self._div64 = torch.zeros((Hh, Ww), **f64) if dtype == torch.float32 else None
        self._z64 = torch.zeros((Hh, Ww), **f64)
        self._div = torch.zeros((Hh, Ww), dtype=dtype, device=self.device)
        self._mg = MultigridSolver()
        self.solve_log: deque = deque(maxlen=4096)  # (tag, V-cycles, rel. residual)

    # ------------------------------------------------------------ pieces
    def zeros_state(self):
        z = dict(dtype=self.dtype, device=self.device)
        return (torch.zeros((self.H, self.W), **z), torch.zeros((self.H, self.W + 1), **z),
                torch.zeros((self.H + 1, self.W), **z))

    def _solve(self, out64, div, warm: bool, tag: str):
        """out64 = S(div) in FP64 (div in the field dtype)."""
        if not warm:
            out64.zero_()
        if self.dtype == torch.float32:
            self._div64.copy_(div)
            rhs = self._div64
        else:
            rhs = div
        self._mg(out64, rhs, self.face_mask_u, self.face_mask_v, self.fluid_density,
                 self.cell_size, self.dt, **self.solver_kwargs)
        self.solve_log.append((tag, LAST_SOLVE_INFO.get("iterations", 0),
                               LAST_SOLVE_INFO.get("rel_residual")))
        return out64 if self.dtype == torch.float64 else out64.to(self.dtype)

    def project_(self, u, v):
        """(u, v) <- B M ((u, v) - c G S D (u, v)) in place."""
        hip.compute_divergence(u, v, self._div, self.face_mask_u, self.face_mask_v,
                               self.cell_size)
        p = self._solve(self._p64, self._div, False, "forward")
        hip.apply_pressure_gradient(u, v, p, self.face_mask_u, self.face_mask_v,
                                    self.fluid_density, self.cell_size, self.dt,
                                    solid_mask=self.solid)

    def project_adjoint_(self, u_bar, v_bar):
        """(u_bar, v_bar) <- M (I - c G S D) B (u_bar, v_bar) in place: the
        transpose of project_."""
        hip.enforce_boundary_conditions(u_bar, v_bar, self.solid)
        hip.compute_divergence(u_bar, v_bar, self._div, self.face_mask_u, self.face_mask_v,
                               self.cell_size)
        z = self._solve(self._z64, self._div, self.warm_start_adjoint, "adjoint")
        hip.apply_pressure_gradient(u_bar, v_bar, z, self.face_mask_u, self.face_mask_v,
                                    self.fluid_density, self.cell_size, self.dt)

    def step_forward(self, rho, u, v, strengths, n: int):
        """One step; returns (rho2, u3, v3) and the tensors the adjoint needs."""
        h, dt = self.cell_size, self.dt
        rho0 = rho.clone()
        adj.box_set_indexed(rho0, self.box, strengths, n)
        v1 = v.clone()
        adj.apply_buoyancy(v1, rho0, dt, self.gravity, self.density_coefficient)
        u2 = torch.empty_like(u)
        v2 = torch.empty_like(v1)
        hip.advect_velocity(u, u2, v1, v2, self.face_mask_u, self.face_mask_v, dt, h)
        rho2 = torch.empty_like(rho0)
        hip.advect_scalar_field(rho0, rho2, u2, v2, self.is_fluid, dt, h)
        u3 = u2.clone()
        v3 = v2.clone()
        self.project_(u3, v3)
        return rho2, u3, v3, (rho0, u, v1, u2, v2)

    def step_backward(self, saved, rho_bar, u_bar, v_bar, strengths_bar, n: int):
        """Adjoint of step_forward: returns (rho_bar, u_bar, v_bar) of the
        step's inputs and adds the strength gradient to strengths_bar[n]."""
        rho0, u, v1, u2, v2 = saved
        h, dt, mode = self.cell_size, self.dt, self.adjoint_mode
        wu = u_bar.clone()
        wv = v_bar.clone()
        self.project_adjoint_(wu, wv)                                    # 5
        rho0_bar = torch.empty_like(rho0)
        adj.advect_scalar_adjoint(rho_bar, rho0, u2, v2, self.is_fluid, dt, h, rho0_bar,
                                  wu, wv, mode=mode)                     # 4
        u_in_bar = torch.empty_like(u)
        v1_bar = torch.empty_like(v1)
        adj.advect_velocity_adjoint(wu, wv, u, v1, self.face_mask_u, self.face_mask_v, dt, h,
                                    u_in_bar, v1_bar, mode=mode)         # 3
        adj.apply_buoyancy_adjoint(v1_bar, rho0_bar, dt, self.gravity,
                                   self.density_coefficient)             # 2
        adj.box_set_adjoint(rho0_bar, self.box, strengths_bar, n)        # 1
        return rho0_bar, u_in_bar, v1_bar

    # ---------------------------------------------------------- rollouts
    def rollout(self, u0, v0, strengths, steps: int, rho0=None, checkpoint_every: int = 8):
        """Differentiable rollout of ``steps`` steps from (rho0 or 0, u0, v0);
        returns (rho, u, v) at the end."""
        if strengths.numel() < steps:
            raise ValueError(f"need {steps} strengths, got {strengths.numel()}")
        rho = rho0 if rho0 is not None else torch.zeros((self.H, self.W), dtype=self.dtype,
                                                        device=self.device)
        u, v = u0, v0
        k = int(checkpoint_every)
        for a in range(0, steps, k if k > 0 else max(steps, 1)):
            b = min(a + k, steps) if k > 0 else steps
            if k > 0 and torch.is_grad_enabled():
                rho, u, v = checkpoint(self._segment, a, b, rho, u, v, strengths,
                                       use_reentrant=False)
            else:
                rho, u, v = self._segment(a, b, rho, u, v, strengths)
        return rho, u, v

    def _segment(self, a, b, rho, u, v, strengths):
        for n in range(a, b):
            rho, u, v = SmokeStepFunction.apply(self, n, rho, u, v, strengths)
        return rho, u, v

    def forward_only(self, u0, v0, strengths, steps: int, rho0=None):
        """The same rollout without autograd bookkeeping (timing baseline)."""
        with torch.no_grad():
            rho = rho0 if rho0 is not None else torch.zeros((self.H, self.W), dtype=self.dtype,
                                                            device=self.device)
            u, v = u0, v0
            for n in range(steps):
                rho, u, v, _ = self.step_forward(rho, u, v, strengths, n)
        return rho, u, v


class SmokeStepFunction(torch.autograd.Function):
    """One differentiable step: (rho, u, v, strengths) -> (rho', u', v')."""

    @staticmethod
    def forward(ctx, model: DifferentiableSmoke, n: int, rho, u, v, strengths):
        rho2, u3, v3, saved = model.step_forward(rho, u, v, strengths, n)
        ctx.model = model
        ctx.n = n
        ctx.save_for_backward(*saved)
        ctx.nstr = strengths.numel()
        return rho2, u3, v3

    @staticmethod
    def backward(ctx, rho_bar, u_bar, v_bar):
        model = ctx.model
        saved = ctx.saved_tensors
        s_bar = torch.zeros(ctx.nstr, dtype=model.dtype, device=model.device)
        r, uu, vv = model.step_backward(saved, rho_bar.contiguous(), u_bar.contiguous(),
                                        v_bar.contiguous(), s_bar, ctx.n)
        return None, None, r, uu, vv, s_bar


def density_loss(rho, target):
    """0.5 * ||rho - target||^2."""
    d = rho - target
    return 0.5 * (d * d).sum()


def gaussian_target(H: int, W: int, h: float, cx: float = 0.7, cy: float = 0.7,
                    sigma: float = 0.08, peak: float = 1.0) -> np.ndarray:
    x = (np.arange(W) + 0.5) * h
    y = (np.arange(H) + 0.5) * h
    gx = np.exp(-((x - cx) ** 2) / (2 * sigma * sigma))
    gy = np.exp(-((y - cy) ** 2) / (2 * sigma * sigma))
    return peak * np.outer(gy, gx)


def seeded_targets(H: int, W: int, h: float, seeds, sigma: float = 0.08) -> list:
    """One Gaussian target per seed, centre ~ U(0.2, 0.8)^2."""
    out = []
    for s in seeds:
        cx, cy = np.random.default_rng(int(s)).uniform(0.2, 0.8, size=2)
        out.append(gaussian_target(H, W, h, float(cx), float(cy), sigma))
    return out


def from_scene(scene, dtype=torch.float64, device="cuda", **kw) -> DifferentiableSmoke:
    """DifferentiableSmoke for a Scene with one box source (see scenes.py)."""
    b = scene.boxes[0]
    sp = scene.solver_params
    kw.setdefault("tolerance", sp.get("tolerance", 1e-8))
    kw.setdefault("num_levels", sp.get("num_levels"))
    return DifferentiableSmoke(scene.H, scene.W, scene.cell_size, scene.dt,
                               (b.x0, b.y0, b.x1, b.y1), dtype=dtype, device=device,
                               gravity=scene.gravity,
                               density_coefficient=scene.density_coefficient, **kw)


def directional_fd(model: DifferentiableSmoke, u0, v0, strengths, steps, target, du, dv, ds,
                   eps: float) -> float:
    """Central difference of the loss along (du, dv, ds)."""
    def loss_at(sign):
        rho, _, _ = model.forward_only(u0 + sign * eps * du, v0 + sign * eps * dv,
                                       strengths + sign * eps * ds, steps)
        return float(density_loss(rho, target))
    return (loss_at(+1.0) - loss_at(-1.0)) / (2.0 * eps)


__all__ = ["DifferentiableSmoke", "SmokeStepFunction", "density_loss", "gaussian_target",
           "seeded_targets", "from_scene", "directional_fd"]
