"""HIP implementations of the smoke-step operators (backend key ``"hip"``).

Each function keeps the exact positional signature of the Numba kernel it
replaces (``src/cfd/backends/numba/operators/*.py`` in the reference) and of
the Protocols in ``src/cfd/operators/base.py``.  Arguments are torch tensors
on the GPU: float64 or float32 fields (one dtype per call), ``torch.bool`` (or
uint8) masks, all C-contiguous.  Like the Numba kernels they write into the
caller's buffers and return None (``compute_residual`` returns its output).
Work is enqueued on the current torch stream.

Shape and dtype errors raise ValueError (the Numba kernels do no checks; a
wrong shape there reads out of bounds).
"""
from __future__ import annotations

import ctypes
import math

import torch

from ._lib import SMK_F32, SMK_F64, call, stream_of

_DT = {torch.float64: SMK_F64, torch.float32: SMK_F32}


def _field(t, name: str, dtype=None):
    if not isinstance(t, torch.Tensor):
        raise ValueError(f"{name} must be a torch.Tensor on the GPU, got {type(t).__name__}")
    if t.device.type != "cuda":
        raise ValueError(f"{name} must live on the GPU (got device {t.device})")
    if t.dtype not in _DT:
        raise ValueError(f"{name} must be float64 or float32 (got {t.dtype})")
    if dtype is not None and t.dtype != dtype:
        raise ValueError(f"{name} has dtype {t.dtype}, expected {dtype} like the other fields")
    if t.dim() != 2 or not t.is_contiguous():
        raise ValueError(f"{name} must be a contiguous 2-D tensor (got shape {tuple(t.shape)})")
    return t


def _mask(t, name: str, shape=None):
    if not isinstance(t, torch.Tensor) or t.device.type != "cuda":
        raise ValueError(f"{name} must be a torch.Tensor on the GPU")
    if t.dtype not in (torch.bool, torch.uint8):
        raise ValueError(f"{name} must be torch.bool or torch.uint8 (got {t.dtype})")
    if t.dim() != 2 or not t.is_contiguous():
        raise ValueError(f"{name} must be a contiguous 2-D tensor")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name} has shape {tuple(t.shape)}, expected {tuple(shape)}")
    return t


def _shape(t, shape, name):
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name} has shape {tuple(t.shape)}, expected {tuple(shape)}")


def _p(t):
    return t.data_ptr()


def dtype_code(t: torch.Tensor) -> int:
    return _DT[t.dtype]


# -------------------------------------------------------------- advection
def advect_scalar_field(field_source, field_dest, velocity_u, velocity_v, is_fluid, dt,
                        cell_size, *, scale: float = 1.0):
    """advection.py:112-128.  ``scale`` multiplies the result (density
    dissipation fused into the advection; 1.0 = the reference operator)."""
    _field(field_source, "field_source")
    dt_ = field_source.dtype
    _field(field_dest, "field_dest", dt_)
    _field(velocity_u, "velocity_u", dt_)
    _field(velocity_v, "velocity_v", dt_)
    H, W = field_source.shape
    _shape(field_dest, (H, W), "field_dest")
    _mask(is_fluid, "is_fluid", (H, W))
    call("smk_advect_scalar", _p(field_source), _p(field_dest), H, W, _p(velocity_u),
         velocity_u.shape[0], velocity_u.shape[1], _p(velocity_v), velocity_v.shape[0],
         velocity_v.shape[1], _p(is_fluid), float(dt), float(cell_size), float(scale),
         dtype_code(field_source), stream_of(field_source))


def advect_velocity_u(velocity_u_source, velocity_u_dest, velocity_v, face_mask_u, dt,
                      cell_size):
    """advection.py:130-154."""
    _field(velocity_u_source, "velocity_u_source")
    dt_ = velocity_u_source.dtype
    _field(velocity_u_dest, "velocity_u_dest", dt_)
    _field(velocity_v, "velocity_v", dt_)
    uh, uw = velocity_u_source.shape
    _shape(velocity_u_dest, (uh, uw), "velocity_u_dest")
    _mask(face_mask_u, "face_mask_u", (uh, uw))
    call("smk_advect_u", _p(velocity_u_source), _p(velocity_u_dest), uh, uw, _p(velocity_v),
         velocity_v.shape[0], velocity_v.shape[1], _p(face_mask_u), float(dt), float(cell_size),
         dtype_code(velocity_u_source), stream_of(velocity_u_source))


def advect_velocity_v(velocity_v_source, velocity_v_dest, velocity_u, face_mask_v, dt,
                      cell_size):
    """advection.py:156-180."""
    _field(velocity_v_source, "velocity_v_source")
    dt_ = velocity_v_source.dtype
    _field(velocity_v_dest, "velocity_v_dest", dt_)
    _field(velocity_u, "velocity_u", dt_)
    vh, vw = velocity_v_source.shape
    _shape(velocity_v_dest, (vh, vw), "velocity_v_dest")
    _mask(face_mask_v, "face_mask_v", (vh, vw))
    call("smk_advect_v", _p(velocity_v_source), _p(velocity_v_dest), vh, vw, _p(velocity_u),
         velocity_u.shape[0], velocity_u.shape[1], _p(face_mask_v), float(dt), float(cell_size),
         dtype_code(velocity_v_source), stream_of(velocity_v_source))


def advect_velocity(u_src, u_dst, v_src, v_dst, face_mask_u, face_mask_v, dt, cell_size):
    """advect_velocity_u + advect_velocity_v from the same old velocity in one
    launch (simulation.py:280-286)."""
    _field(u_src, "u_src")
    dt_ = u_src.dtype
    H, W1 = u_src.shape
    W = W1 - 1
    for t, n, s in ((u_dst, "u_dst", (H, W + 1)), (v_src, "v_src", (H + 1, W)),
                    (v_dst, "v_dst", (H + 1, W))):
        _field(t, n, dt_)
        _shape(t, s, n)
    _mask(face_mask_u, "face_mask_u", (H, W + 1))
    _mask(face_mask_v, "face_mask_v", (H + 1, W))
    call("smk_advect_velocity", _p(u_src), _p(u_dst), _p(v_src), _p(v_dst), _p(face_mask_u),
         _p(face_mask_v), H, W, float(dt), float(cell_size), dtype_code(u_src), stream_of(u_src))


# ----------------------------------------------------------------- forces
def apply_forces(velocity_v, density, temperature, dt, gravity, buoyancy_coefficient,
                 density_coefficient, ambient_temperature):
    """forces.py:78-105 (float32 force scratch reproduced).  ``temperature``
    may be None for an all-zero temperature field."""
    _field(velocity_v, "velocity_v")
    dt_ = velocity_v.dtype
    _field(density, "density", dt_)
    H, W = density.shape
    if temperature is None:
        tp, Ht, Wt = None, H, W
    else:
        _field(temperature, "temperature", dt_)
        tp, (Ht, Wt) = _p(temperature), temperature.shape
    call("smk_apply_forces", _p(velocity_v), velocity_v.shape[0], velocity_v.shape[1],
         _p(density), H, W, tp, Ht, Wt, float(dt), float(gravity), float(buoyancy_coefficient),
         float(density_coefficient), float(ambient_temperature), dtype_code(velocity_v),
         stream_of(velocity_v))


def cool_temperature(temperature, ambient_temperature, cooling_rate, dt):
    """forces.py:108-123 (decay = exp(-rate*dt) evaluated once on the host,
    as forces.py:120 does)."""
    _field(temperature, "temperature")
    H, W = temperature.shape
    decay = math.exp(-cooling_rate * dt)
    call("smk_cool_temperature", _p(temperature), H, W, float(ambient_temperature), decay,
         dtype_code(temperature), stream_of(temperature))


def scale_field(field, factor):
    """simulation.py:232 ``density *= density_dissipation``."""
    if not isinstance(field, torch.Tensor) or not field.is_contiguous():
        raise ValueError("field must be a contiguous GPU tensor")
    _field(field.view(1, -1) if field.dim() != 2 else field, "field")
    call("smk_scale", _p(field), field.numel(), float(factor), dtype_code(field), stream_of(field))


# ------------------------------------------------------------- projection
def compute_divergence(velocity_u, velocity_v, divergence, mask_u, mask_v, cell_size):
    """divergence.py:4-25."""
    _field(divergence, "divergence")
    dt_ = divergence.dtype
    H, W = divergence.shape
    _field(velocity_u, "velocity_u", dt_)
    _field(velocity_v, "velocity_v", dt_)
    _shape(velocity_u, (H, W + 1), "velocity_u")
    _shape(velocity_v, (H + 1, W), "velocity_v")
    _mask(mask_u, "mask_u", (H, W + 1))
    _mask(mask_v, "mask_v", (H + 1, W))
    call("smk_divergence", _p(velocity_u), _p(velocity_v), _p(divergence), _p(mask_u), _p(mask_v),
         H, W, float(cell_size), dtype_code(divergence), stream_of(divergence))


def apply_pressure_gradient(velocity_u, velocity_v, pressure, face_mask_u, face_mask_v,
                            fluid_density, cell_size, dt, *, solid_mask=None):
    """pressure_gradient.py:4-37.  With ``solid_mask`` the boundary
    enforcement of boundaries.py:7-39 is fused into the same launch."""
    _field(pressure, "pressure")
    dt_ = pressure.dtype
    H, W = pressure.shape
    _field(velocity_u, "velocity_u", dt_)
    _field(velocity_v, "velocity_v", dt_)
    _shape(velocity_u, (H, W + 1), "velocity_u")
    _shape(velocity_v, (H + 1, W), "velocity_v")
    _mask(face_mask_u, "face_mask_u", (H, W + 1))
    _mask(face_mask_v, "face_mask_v", (H + 1, W))
    sp = None
    if solid_mask is not None:
        _mask(solid_mask, "solid_mask", (H, W))
        sp = _p(solid_mask)
    call("smk_pressure_gradient", _p(velocity_u), _p(velocity_v), _p(pressure), _p(face_mask_u),
         _p(face_mask_v), sp, H, W, float(fluid_density), float(cell_size), float(dt),
         dtype_code(pressure), stream_of(pressure))


def enforce_boundary_conditions(velocity_u, velocity_v, solid_mask):
    """boundaries.py:7-39."""
    _mask(solid_mask, "solid_mask")
    H, W = solid_mask.shape
    _field(velocity_u, "velocity_u")
    _field(velocity_v, "velocity_v", velocity_u.dtype)
    _shape(velocity_u, (H, W + 1), "velocity_u")
    _shape(velocity_v, (H + 1, W), "velocity_v")
    call("smk_enforce_boundaries", _p(velocity_u), _p(velocity_v), _p(solid_mask), H, W,
         dtype_code(velocity_u), stream_of(velocity_u))


def compute_residual(pressure, divergence, is_fluid, face_mask_u, face_mask_v, alpha,
                     reciprocal_beta, residual_out):
    """residual.py:6-42; returns residual_out."""
    _field(pressure, "pressure")
    dt_ = pressure.dtype
    H, W = pressure.shape
    _field(divergence, "divergence", dt_)
    _field(residual_out, "residual_out", dt_)
    _shape(divergence, (H, W), "divergence")
    _shape(residual_out, (H, W), "residual_out")
    _mask(is_fluid, "is_fluid", (H, W))
    _mask(face_mask_u, "face_mask_u", (H, W + 1))
    _mask(face_mask_v, "face_mask_v", (H + 1, W))
    if reciprocal_beta == 0:
        raise ValueError("reciprocal_beta must be non-zero")
    call("smk_residual", _p(pressure), _p(divergence), _p(is_fluid), _p(face_mask_u),
         _p(face_mask_v), H, W, float(alpha), float(reciprocal_beta), _p(residual_out),
         dtype_code(pressure), stream_of(pressure))
    return residual_out


def divergence_residual(velocity_u, velocity_v, pressure, is_fluid, face_mask_u, face_mask_v,
                        cell_size, alpha, reciprocal_beta, divergence, residual_out):
    """The step's metrics pass (simulation.py:251-265) in one launch:
    compute_divergence(velocity_u, velocity_v, divergence, ...) then
    compute_residual(pressure, divergence, ...) into residual_out, with the
    same results as the two calls."""
    _field(pressure, "pressure")
    dt_ = pressure.dtype
    H, W = pressure.shape
    for t, name, shape in ((velocity_u, "velocity_u", (H, W + 1)),
                           (velocity_v, "velocity_v", (H + 1, W)),
                           (divergence, "divergence", (H, W)),
                           (residual_out, "residual_out", (H, W))):
        _field(t, name, dt_)
        _shape(t, shape, name)
    _mask(is_fluid, "is_fluid", (H, W))
    _mask(face_mask_u, "face_mask_u", (H, W + 1))
    _mask(face_mask_v, "face_mask_v", (H + 1, W))
    if reciprocal_beta == 0:
        raise ValueError("reciprocal_beta must be non-zero")
    call("smk_divergence_residual", _p(velocity_u), _p(velocity_v), _p(pressure), _p(is_fluid),
         _p(face_mask_u), _p(face_mask_v), H, W, float(cell_size), float(alpha),
         float(reciprocal_beta), _p(divergence), _p(residual_out), dtype_code(pressure),
         stream_of(pressure))
    return residual_out


# ------------------------------------------------------------------ masks
def _written(*tensors):
    """Bump the version counters of mask tensors a kernel wrote in place
    (writes through a raw pointer do not), so that holders of those tensors
    -- the multigrid solver's level-0 cache -- see the change."""
    for t in tensors:
        torch.autograd.graph.increment_version(t)


def face_masks_from_solid(solid_mask, face_mask_u, face_mask_v):
    """simulation.py:200-202 / multigrid/masks.py:30-49 (outer faces closed)."""
    _mask(solid_mask, "solid_mask")
    H, W = solid_mask.shape
    _mask(face_mask_u, "face_mask_u", (H, W + 1))
    _mask(face_mask_v, "face_mask_v", (H + 1, W))
    call("smk_face_masks_from_solid", _p(solid_mask), H, W, _p(face_mask_u), _p(face_mask_v),
         stream_of(solid_mask))
    _written(face_mask_u, face_mask_v)


def is_fluid_from_masks(face_mask_u, face_mask_v, out):
    """multigrid/masks.py:52-74."""
    _mask(face_mask_u, "face_mask_u")
    H, W = face_mask_u.shape[0], face_mask_u.shape[1] - 1
    _mask(face_mask_v, "face_mask_v", (H + 1, W))
    _mask(out, "is_fluid", (H, W))
    call("smk_is_fluid", _p(face_mask_u), _p(face_mask_v), H, W, _p(out), stream_of(out))
    _written(out)
    return out


def coarsen_solid_mask(solid_fine, solid_coarse_out):
    """multigrid/masks.py:3-27."""
    _mask(solid_fine, "solid_fine")
    H, W = solid_fine.shape
    _mask(solid_coarse_out, "solid_coarse_out", (H // 2, W // 2))
    call("smk_coarsen_solid", _p(solid_fine), H, W, _p(solid_coarse_out), stream_of(solid_fine))
    _written(solid_coarse_out)
    return solid_coarse_out


# --------------------------------------------------------------- emitters
APPROACHES = {"hard_cutoff": 0, "linear": 1, "quadratic": 2, "gaussian": 3}


def emit_scalar(field, value, dt, pos_x, pos_y, radius, approach):
    """emitters.py:157-178 + falloffs emitters.py:339-411."""
    _field(field, "field")
    if approach not in APPROACHES:
        raise ValueError(f"Unknown emission approach: {approach}")
    H, W = field.shape
    call("smk_emit_scalar", _p(field), H, W, float(value), float(dt), float(pos_x), float(pos_y),
         float(radius), APPROACHES[approach], dtype_code(field), stream_of(field))


def box_source(field, x0, y0, x1, y1, value, mode="set"):
    """field[y0:y1, x0:x1] = value (mode "set") or += value ("add")."""
    _field(field, "field")
    if mode not in ("set", "add"):
        raise ValueError("mode must be 'set' or 'add'")
    H, W = field.shape
    call("smk_box_source", _p(field), H, W, int(x0), int(y0), int(x1), int(y1), float(value),
         0 if mode == "set" else 1, dtype_code(field), stream_of(field))


def box_sources(field, boxes):
    """box_source for each (x0, y0, x1, y1, value, mode) of ``boxes`` in list
    order; pairwise-disjoint boxes (after clamping to the field) run in one
    launch."""
    _field(field, "field")
    n = len(boxes)
    for b in boxes:
        if b[5] not in ("set", "add"):
            raise ValueError("mode must be 'set' or 'add'")
    H, W = field.shape
    rects = (ctypes.c_int * max(4 * n, 1))(*[int(v) for b in boxes for v in b[:4]])
    values = (ctypes.c_double * max(n, 1))(*[float(b[4]) for b in boxes])
    modes = (ctypes.c_int * max(n, 1))(*[0 if b[5] == "set" else 1 for b in boxes])
    call("smk_box_sources", _p(field), H, W, n, ctypes.addressof(rects),
         ctypes.addressof(values), ctypes.addressof(modes), dtype_code(field), stream_of(field))

