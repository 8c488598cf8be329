"""HIP pressure solvers (backend key ``"hip"``).

Signatures follow the reference's PressureSolver Protocol
(``src/cfd/solvers/base.py:5-19``): ``fn(pressure, divergence, face_mask_u,
face_mask_v, fluid_density, cell_size, dt, **solver_kwargs)`` updating
``pressure`` in place.  Extra keyword arguments the reference does not have
(``tolerance``, ``check_every``, ``smoother_omega``, ``coarse_scale``,
``precision``, ``inner_cycles``) default to the reference behaviour.

The statistics of the last solve (sweeps or V-cycles, relative residual) are
kept in :data:`LAST_SOLVE_INFO`.
"""
from __future__ import annotations

import ctypes

import torch

from ._lib import (SMK_COARSE_REDISCRETIZED, SMK_COARSE_REFERENCE, SMK_COARSEST_EXACT,
                   SMK_COARSEST_EXACT_1024,
                   SMK_COARSEST_OPERATOR, SMK_COARSEST_SWEEPS, SMK_INNER_F32, SMK_INNER_F64,
                   SMK_PREC_MIXED,
                   SMK_PREC_NATIVE, SMK_SMOOTH_GS, SMK_SMOOTH_JACOBI, SMK_SMOOTH_RBSOR, MgParams,
                   SolveInfo,
                   call, lib, stream_of)
from .operators import _field, _mask, _shape, dtype_code

LAST_SOLVE_INFO: dict = {}

_WS: dict = {}


def _workspace(device: torch.device, nbytes: int) -> torch.Tensor:
    """Grow-only device scratch shared by the single-grid solvers of one
    device; stream-ordered use on the current stream makes reuse safe."""
    key = (device.type, device.index)
    ws = _WS.get(key)
    if ws is None or ws.numel() < nbytes:
        ws = torch.empty(int(nbytes), dtype=torch.uint8, device=device)
        _WS[key] = ws
    return ws


def _check(pressure, divergence, face_mask_u, face_mask_v):
    _field(pressure, "pressure")
    H, W = pressure.shape
    _field(divergence, "divergence", pressure.dtype)
    _shape(divergence, (H, W), "divergence")
    _mask(face_mask_u, "face_mask_u", (H, W + 1))
    _mask(face_mask_v, "face_mask_v", (H + 1, W))
    return H, W


def jacobi_pressure_solver(pressure, divergence, face_mask_u, face_mask_v, fluid_density,
                           cell_size, dt, iterations, **_ignored):
    """jacobi.py:27-52 (iterations sweeps, LDS-tiled, several per launch)."""
    H, W = _check(pressure, divergence, face_mask_u, face_mask_v)
    code = dtype_code(pressure)
    nb = lib().smk_solver_workspace_bytes(H, W, code)
    ws = _workspace(pressure.device, nb)
    call("smk_jacobi_solve", pressure.data_ptr(), divergence.data_ptr(), face_mask_u.data_ptr(),
         face_mask_v.data_ptr(), H, W, float(fluid_density), float(cell_size), float(dt),
         int(iterations), ws.data_ptr(), ws.numel(), code, stream_of(pressure))
    LAST_SOLVE_INFO.clear()
    LAST_SOLVE_INFO.update(method="jacobi", iterations=int(iterations))


def red_black_sor_gauss_seidel_pressure_solver(pressure, divergence, face_mask_u, face_mask_v,
                                               fluid_density, cell_size, dt, iterations,
                                               omega=0.0, tolerance=0.0, check_every=20,
                                               precision="native", **_ignored):
    """red_black_sor_gauss_seidel.py:30-42.

    omega <= 0 or > 2 selects calculateOptimalOmega(max(H, W)).  With
    tolerance > 0 the sweeps stop once ||r||/||alpha*div|| <= tolerance
    (checked every ``check_every`` sweeps; ``iterations`` is then the cap).
    ``precision="mixed"`` (float64 fields only) runs the sweeps in FP32 inside
    FP64 defect correction, one correction per ``check_every`` sweeps."""
    H, W = _check(pressure, divergence, face_mask_u, face_mask_v)
    code = dtype_code(pressure)
    if precision not in _PREC:
        raise ValueError(f"precision must be one of {sorted(_PREC)}, got {precision!r}")
    nb = lib().smk_solver_workspace_bytes(H, W, code)
    ws = _workspace(pressure.device, nb)
    info = SolveInfo()
    if precision == "mixed":
        if pressure.dtype != torch.float64:
            raise ValueError("precision='mixed' needs float64 pressure and divergence")
        call("smk_rbsor_solve_mixed", pressure.data_ptr(), divergence.data_ptr(),
             face_mask_u.data_ptr(), face_mask_v.data_ptr(), H, W, float(fluid_density),
             float(cell_size), float(dt), int(iterations), float(omega), float(tolerance),
             int(check_every), ws.data_ptr(), ws.numel(), ctypes.byref(info),
             stream_of(pressure))
    else:
        call("smk_rbsor_solve", pressure.data_ptr(), divergence.data_ptr(),
             face_mask_u.data_ptr(), face_mask_v.data_ptr(), H, W, float(fluid_density),
             float(cell_size), float(dt), int(iterations), float(omega), float(tolerance),
             int(check_every), ws.data_ptr(), ws.numel(), ctypes.byref(info), code,
             stream_of(pressure))
    LAST_SOLVE_INFO.clear()
    LAST_SOLVE_INFO.update(method="rb_sor_gauss_seidel", **info.as_dict())


def gauss_seidel_pressure_solver(pressure, divergence, face_mask_u, face_mask_v, fluid_density,
                                 cell_size, dt, iterations, **_ignored):
    """gauss_seidel.py:29-48: ``iterations`` lexicographic sweeps, run as a
    wavefront (every cell of one time step i + j + 2*sweep at once), which
    reproduces the serial sweep exactly."""
    H, W = _check(pressure, divergence, face_mask_u, face_mask_v)
    code = dtype_code(pressure)
    ws = _workspace(pressure.device, lib().smk_solver_workspace_bytes(H, W, code))
    call("smk_gs_solve", pressure.data_ptr(), divergence.data_ptr(), face_mask_u.data_ptr(),
         face_mask_v.data_ptr(), H, W, float(fluid_density), float(cell_size), float(dt),
         int(iterations), ws.data_ptr(), ws.numel(), code, stream_of(pressure))
    LAST_SOLVE_INFO.clear()
    LAST_SOLVE_INFO.update(method="gauss_seidel", iterations=int(iterations))


def residual_norm2(pressure, divergence, face_mask_u, face_mask_v, alpha) -> float:
    """||r||^2 over fluid cells (r of residual.py:6-42 with 1/beta = 1/4),
    reduced on the device; synchronises."""
    H, W = _check(pressure, divergence, face_mask_u, face_mask_v)
    code = dtype_code(pressure)
    ws = _workspace(pressure.device, lib().smk_solver_workspace_bytes(H, W, code))
    out = ctypes.c_double(0.0)
    call("smk_residual_norm2", pressure.data_ptr(), divergence.data_ptr(), face_mask_u.data_ptr(),
         face_mask_v.data_ptr(), H, W, float(alpha), ws.data_ptr(), ws.numel(), ctypes.byref(out),
         code, stream_of(pressure))
    return out.value


_SMOOTHERS = {"jacobi": SMK_SMOOTH_JACOBI, "gauss_seidel": SMK_SMOOTH_GS,
              "rb_sor_gauss_seidel": SMK_SMOOTH_RBSOR}
_COARSE = {"reference": SMK_COARSE_REFERENCE, "rediscretized": SMK_COARSE_REDISCRETIZED}
_PREC = {"native": SMK_PREC_NATIVE, "mixed": SMK_PREC_MIXED}
_COARSEST = {"sweeps": SMK_COARSEST_SWEEPS, "operator": SMK_COARSEST_OPERATOR,
             "exact": SMK_COARSEST_EXACT, "exact_1024": SMK_COARSEST_EXACT_1024}
_INNER = {"float64": SMK_INNER_F64, "float32": SMK_INNER_F32}
_GUESS = {"given": 0, "zero": 1}


class _MgHandle:
    def __init__(self, face_mask_u, face_mask_v, num_levels, dtype_code_):
        H, W = face_mask_u.shape[0], face_mask_u.shape[1] - 1
        h = ctypes.c_void_p()
        call("smk_mg_create", face_mask_u.data_ptr(), face_mask_v.data_ptr(), H, W,
             int(num_levels), int(dtype_code_), stream_of(face_mask_u), ctypes.byref(h))
        self.ptr = h
        self.levels = lib().smk_mg_num_levels(h)

    def close(self):
        if self.ptr is not None and self.ptr.value:
            lib().smk_mg_destroy(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MultigridSolver:
    """multigrid/solver.py:16-177 on the GPU.

    Level masks and buffers are cached per instance under the reference's
    cache rule (solver.py:87-92): the masks of every level are derived again
    from the face masks passed in when ``rebuild_masks`` is set or when the
    hierarchy has fewer levels than ``num_levels`` (the reference then
    rebuilds on every call); otherwise they stay as built, except that level
    0 keeps the face-mask *tensors* of the build, as the reference keeps the
    arrays (solver.py:42-47): when their content changes in place (seen
    through torch's version counter; the simulation's mask kernels bump it),
    the level-0 smoother follows them while the level-0 residual,
    restriction and prolongation keep the is_fluid of the build, as the
    reference's do.  Buffers are reallocated only when the grid shape, level
    count, dtype or device changes; mask refreshes are in place.

    Smoothing is fused into the V-cycle kernels, so ``smoother_backend`` is
    accepted for signature compatibility and the smoother always runs on the
    GPU.  ``smoother_type`` is "jacobi", "gauss_seidel" (the lexicographic
    sweep as an exact wavefront, slow: one workgroup per level) or
    "rb_sor_gauss_seidel" -- the three solvers the reference's registry
    holds (solver.py:164 looks the smoother up there).

    ``initial_guess`` (new): "given" (the reference: ``pressure`` is the
    initial guess) or "zero": its content is ignored and taken as zero, and
    every cell is written -- the caller's ``pressure.zero_()`` before the
    solve (simulation.py:310) folded in, with the initial residual taken from
    the right-hand side's norm (the same numbers, bit for bit).
    """

    def __init__(self):
        self._key = None
        self._handle = None
        self._level0 = None    # (face_mask_u, face_mask_v) tensors level 0 holds
        self._versions = None  # their versions when the masks were last derived
        self._fresh = False    # every level derived from them (no stale level-0 is_fluid)

    @property
    def num_levels_built(self) -> int:
        return self._handle.levels if self._handle is not None else 0

    def _remember(self, face_mask_u, face_mask_v, fresh):
        self._level0 = (face_mask_u, face_mask_v)
        self._versions = (face_mask_u._version, face_mask_v._version)
        self._fresh = fresh

    def _ensure(self, face_mask_u, face_mask_v, num_levels, pressure, rebuild):
        key = (tuple(face_mask_u.shape), int(num_levels), pressure.dtype, pressure.device)
        if self._handle is None or self._key != key:
            if self._handle is not None:
                torch.cuda.current_stream(pressure.device).synchronize()
                self._handle.close()
            self._handle = _MgHandle(face_mask_u, face_mask_v, num_levels, dtype_code(pressure))
            self._key = key
            self._remember(face_mask_u, face_mask_v, True)
            return
        if rebuild or self._handle.levels != int(num_levels):
            # solver.py:87-90 rebuilds from the masks passed in; skipped when
            # that would reproduce the hierarchy already held
            f0, v0 = self._level0
            same = (f0 is face_mask_u and v0 is face_mask_v
                    and self._versions == (face_mask_u._version, face_mask_v._version))
            if not (same and self._fresh):
                call("smk_mg_refresh_masks", self._handle.ptr, face_mask_u.data_ptr(),
                     face_mask_v.data_ptr(), 1, stream_of(pressure))
                self._remember(face_mask_u, face_mask_v, True)
            return
        f0, v0 = self._level0
        if self._versions != (f0._version, v0._version):
            call("smk_mg_refresh_masks", self._handle.ptr, f0.data_ptr(), v0.data_ptr(), 0,
                 stream_of(pressure))
            self._remember(f0, v0, False)

    def __call__(self, pressure, divergence, face_mask_u, face_mask_v, fluid_density, cell_size,
                 dt, v_cycles=2, smoother_type="jacobi", smoother_backend="hip", num_levels=4,
                 smoother_iterations=2, coarsest_iterations=50, rebuild_masks=False,
                 profiler=None, smoother_omega=0.0, coarse_scale="reference", tolerance=0.0,
                 precision="native", inner_cycles=1, coarsest_omega=0.0,
                 coarsest_solve="sweeps", inner_arithmetic="float64", initial_guess="given",
                 **solver_kwargs):
        _check(pressure, divergence, face_mask_u, face_mask_v)
        if smoother_type not in _SMOOTHERS:
            raise ValueError(
                f"unknown multigrid smoother {smoother_type!r}; use one of {list(_SMOOTHERS)}")
        if coarse_scale not in _COARSE:
            raise ValueError(f"coarse_scale must be one of {list(_COARSE)}")
        if precision not in _PREC:
            raise ValueError(f"precision must be one of {list(_PREC)}")
        if precision == "mixed" and pressure.dtype != torch.float64:
            raise ValueError("mixed precision needs float64 pressure/divergence")
        if coarsest_solve not in _COARSEST:
            raise ValueError(f"coarsest_solve must be one of {list(_COARSEST)}")
        if inner_arithmetic not in _INNER:
            raise ValueError(f"inner_arithmetic must be one of {list(_INNER)}")
        if initial_guess not in _GUESS:
            raise ValueError(f"initial_guess must be one of {list(_GUESS)}")
        self._ensure(face_mask_u, face_mask_v, num_levels, pressure, rebuild_masks)
        prm = MgParams(v_cycles=int(v_cycles), smoother=_SMOOTHERS[smoother_type],
                       smoother_iterations=int(smoother_iterations),
                       coarsest_iterations=int(coarsest_iterations),
                       smoother_omega=float(smoother_omega), coarse_scale=_COARSE[coarse_scale],
                       tolerance=float(tolerance), precision=_PREC[precision],
                       inner_cycles=int(inner_cycles), coarsest_omega=float(coarsest_omega),
                       coarsest_solve=_COARSEST[coarsest_solve],
                       inner_arithmetic=_INNER[inner_arithmetic],
                       zero_guess=_GUESS[initial_guess])
        info = SolveInfo()
        call("smk_mg_solve", self._handle.ptr, pressure.data_ptr(), divergence.data_ptr(),
             float(fluid_density), float(cell_size), float(dt), ctypes.byref(prm),
             ctypes.byref(info), stream_of(pressure))
        LAST_SOLVE_INFO.clear()
        LAST_SOLVE_INFO.update(method="multigrid", levels=self._handle.levels, **info.as_dict())
