"""Adjoint (transposed-Jacobian) operators of the smoke step's transport.

Thin validated wrappers over the ``smk_*_adjoint`` entry points of
libsmoke_hip.so and the two forward pieces whose controls are device data
(a box strength read from a tensor, buoyancy kept in the field dtype).  The
forward operators they transpose are the reference's Numba kernels
(``src/cfd/backends/numba/operators/advection.py:112-180``,
``forces.py:19-46``) and the box source of the plume workloads.

``mode`` selects how the transposed bilinear gather is summed: ``"gather"``
(deterministic, each source visits the targets that can reach it) or
``"scatter"`` (atomic adds).
"""
from __future__ import annotations

import torch

from ._lib import SMK_ADJ_GATHER, SMK_ADJ_SCATTER, call, lib, stream_of
from .operators import _field, _mask, _p, _shape, dtype_code

ADJOINT_MODES = {"gather": SMK_ADJ_GATHER, "scatter": SMK_ADJ_SCATTER}

_WS: dict = {}


def workspace(H: int, W: int, like: torch.Tensor) -> torch.Tensor:
    """Grow-only scratch per device for the advection adjoints (used in
    stream order on the current stream)."""
    nbytes = int(lib().smk_adjoint_workspace_bytes(H, W, dtype_code(like)))
    key = (like.device.type, like.device.index)
    ws = _WS.get(key)
    if ws is None or ws.numel() < nbytes:
        ws = torch.empty(nbytes, dtype=torch.uint8, device=like.device)
        _WS[key] = ws
    return ws


def _mode(mode: str) -> int:
    if mode not in ADJOINT_MODES:
        raise ValueError(f"adjoint mode must be one of {sorted(ADJOINT_MODES)}, got {mode!r}")
    return ADJOINT_MODES[mode]


def _values(t, like, name):
    if not isinstance(t, torch.Tensor) or t.device != like.device or t.dtype != like.dtype:
        raise ValueError(f"{name} must be a {like.dtype} tensor on {like.device}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    return t


def box_set_indexed(field, box, values, index: int):
    """field[y0:y1, x0:x1] = values[index] (box = (x0, y0, x1, y1))."""
    _field(field, "field")
    _values(values, field, "values")
    if not 0 <= int(index) < values.numel():
        raise ValueError(f"index {index} out of range for {values.numel()} values")
    H, W = field.shape
    x0, y0, x1, y1 = (int(b) for b in box)
    call("smk_box_set_indexed", _p(field), H, W, x0, y0, x1, y1, _p(values), int(index),
         dtype_code(field), stream_of(field))


def box_set_adjoint(field_bar, box, values_bar, index: int):
    """values_bar[index] += field_bar[box].sum(); field_bar[box] = 0."""
    _field(field_bar, "field_bar")
    _values(values_bar, field_bar, "values_bar")
    if not 0 <= int(index) < values_bar.numel():
        raise ValueError(f"index {index} out of range for {values_bar.numel()} values")
    H, W = field_bar.shape
    x0, y0, x1, y1 = (int(b) for b in box)
    call("smk_box_set_adjoint", _p(field_bar), H, W, x0, y0, x1, y1, _p(values_bar), int(index),
         dtype_code(field_bar), stream_of(field_bar))


def apply_buoyancy(velocity_v, density, dt, gravity, density_coefficient=1.0):
    """forces.py:19-46 buoyancy with the force in the field dtype."""
    _field(density, "density")
    _field(velocity_v, "velocity_v", density.dtype)
    H, W = density.shape
    _shape(velocity_v, (H + 1, W), "velocity_v")
    call("smk_apply_buoyancy", _p(velocity_v), _p(density), H, W, float(dt), float(gravity),
         float(density_coefficient), dtype_code(density), stream_of(density))


def apply_buoyancy_adjoint(velocity_v_bar, density_bar, dt, gravity, density_coefficient=1.0):
    """density_bar += (d v / d density)^T velocity_v_bar."""
    _field(density_bar, "density_bar")
    _field(velocity_v_bar, "velocity_v_bar", density_bar.dtype)
    H, W = density_bar.shape
    _shape(velocity_v_bar, (H + 1, W), "velocity_v_bar")
    call("smk_apply_buoyancy_adjoint", _p(velocity_v_bar), _p(density_bar), H, W, float(dt),
         float(gravity), float(density_coefficient), dtype_code(density_bar),
         stream_of(density_bar))


def advect_scalar_adjoint(dest_bar, field_source, velocity_u, velocity_v, is_fluid, dt,
                          cell_size, source_bar, velocity_u_bar=None, velocity_v_bar=None,
                          mode: str = "gather"):
    """Adjoint of advect_scalar_field: source_bar = (d dest / d source)^T
    dest_bar; the backtrace sensitivity is added to velocity_u_bar /
    velocity_v_bar when given."""
    _field(field_source, "field_source")
    dt_ = field_source.dtype
    H, W = field_source.shape
    for t, n, shp in ((dest_bar, "dest_bar", (H, W)), (source_bar, "source_bar", (H, W)),
                      (velocity_u, "velocity_u", (H, W + 1)),
                      (velocity_v, "velocity_v", (H + 1, W))):
        _field(t, n, dt_)
        _shape(t, shp, n)
    _mask(is_fluid, "is_fluid", (H, W))
    if (velocity_u_bar is None) != (velocity_v_bar is None):
        raise ValueError("velocity_u_bar and velocity_v_bar must both be given or both be None")
    if velocity_u_bar is not None:
        _field(velocity_u_bar, "velocity_u_bar", dt_)
        _shape(velocity_u_bar, (H, W + 1), "velocity_u_bar")
        _field(velocity_v_bar, "velocity_v_bar", dt_)
        _shape(velocity_v_bar, (H + 1, W), "velocity_v_bar")
    ws = workspace(H, W, field_source)
    call("smk_advect_scalar_adjoint", _p(dest_bar), _p(field_source), _p(velocity_u),
         _p(velocity_v), _p(is_fluid), H, W, float(dt), float(cell_size), _p(source_bar),
         _p(velocity_u_bar) if velocity_u_bar is not None else None,
         _p(velocity_v_bar) if velocity_v_bar is not None else None, _mode(mode), _p(ws),
         ws.numel(), dtype_code(field_source), stream_of(field_source))


def advect_velocity_adjoint(u_dest_bar, v_dest_bar, u_source, v_source, face_mask_u, face_mask_v,
                            dt, cell_size, u_source_bar, v_source_bar, mode: str = "gather"):
    """Adjoint of the u and v advection (both by the old velocity):
    (u_source_bar, v_source_bar) = J^T (u_dest_bar, v_dest_bar), overwritten."""
    _field(u_source, "u_source")
    dt_ = u_source.dtype
    H, W1 = u_source.shape
    W = W1 - 1
    for t, n, shp in ((u_dest_bar, "u_dest_bar", (H, W + 1)),
                      (u_source_bar, "u_source_bar", (H, W + 1)),
                      (v_source, "v_source", (H + 1, W)), (v_dest_bar, "v_dest_bar", (H + 1, W)),
                      (v_source_bar, "v_source_bar", (H + 1, W))):
        _field(t, n, dt_)
        _shape(t, shp, n)
    _mask(face_mask_u, "face_mask_u", (H, W + 1))
    _mask(face_mask_v, "face_mask_v", (H + 1, W))
    ws = workspace(H, W, u_source)
    call("smk_advect_velocity_adjoint", _p(u_dest_bar), _p(v_dest_bar), _p(u_source),
         _p(v_source), _p(face_mask_u), _p(face_mask_v), H, W, float(dt), float(cell_size),
         _p(u_source_bar), _p(v_source_bar), _mode(mode), _p(ws), ws.numel(),
         dtype_code(u_source), stream_of(u_source))
