"""Oracle: CPU restatement of the reference smoke step (test infrastructure).

This package is the checker for the HIP path and nothing else.  Only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it; the product package never does.

The arithmetic lives in ``smoke_oracle.c`` (built into ``libsmoke_oracle.so``
by ``make -C oracle``), a line-by-line C restatement of
``src/cfd/backends/numba/{operators,solvers,solvers/multigrid}`` of the
reference.  The functions here are thin numpy wrappers over that library; each
cites the reference file:line its C body follows.  ``oracle.sim`` restates the
step order of ``SmokeSimulation2D.substep`` (``src/cfd/common/simulation.py``).

Parity pins: ``tests/test_oracle_*.py`` check every wrapper against the
reference's own known-answer tests (restated) and against dense linear algebra
or closed forms.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libsmoke_oracle.so")

_dp = ctypes.POINTER(ctypes.c_double)
_u8 = ctypes.POINTER(ctypes.c_uint8)
_i = ctypes.c_int
_d = ctypes.c_double

_SIGS = {
    "orc_sample": (_d, [_dp, _i, _i, _d, _d]),
    "orc_advect_scalar": (None, [_dp, _dp, _i, _i, _dp, _i, _i, _dp, _i, _i, _u8, _d, _d]),
    "orc_advect_u": (None, [_dp, _dp, _i, _i, _dp, _i, _i, _u8, _d, _d]),
    "orc_advect_v": (None, [_dp, _dp, _i, _i, _dp, _i, _i, _u8, _d, _d]),
    "orc_apply_forces": (None, [_dp, _i, _i, _dp, _i, _i, _dp, _i, _i, _d, _d, _d, _d, _d]),
    "orc_cool_temperature": (None, [_dp, _i, _i, _d, _d, _d]),
    "orc_divergence": (None, [_dp, _dp, _dp, _u8, _u8, _i, _i, _d]),
    "orc_pressure_gradient": (None, [_dp, _dp, _dp, _u8, _u8, _i, _i, _d, _d, _d]),
    "orc_enforce_boundaries": (None, [_dp, _dp, _u8, _i, _i]),
    "orc_residual": (None, [_dp, _dp, _u8, _u8, _u8, _i, _i, _d, _d, _dp]),
    "orc_alpha": (_d, [_d, _d, _d]),
    "orc_optimal_omega": (_d, [_i]),
    "orc_is_fluid": (None, [_u8, _u8, _i, _i, _u8]),
    "orc_jacobi_sweep": (None, [_dp, _dp, _dp, _u8, _u8, _u8, _i, _i, _d, _d]),
    "orc_jacobi_solve": (None, [_dp, _dp, _u8, _u8, _i, _i, _d, _d, _d, _i]),
    "orc_rbsor_sweeps": (None, [_dp, _dp, _u8, _u8, _u8, _i, _i, _d, _d, _d, _i]),
    "orc_rbsor_solve": (None, [_dp, _dp, _u8, _u8, _i, _i, _d, _d, _d, _i, _d]),
    "orc_gs_solve": (None, [_dp, _dp, _u8, _u8, _i, _i, _d, _d, _d, _i]),
    "orc_norm2_fluid": (_d, [_dp, _u8, _i, _i]),
    "orc_rbsor_solve_tol": (_i, [_dp, _dp, _u8, _u8, _i, _i, _d, _d, _d, _d, _d, _i, _i, _dp]),
    "orc_coarsen_solid": (None, [_u8, _i, _i, _u8]),
    "orc_build_face_masks": (None, [_u8, _i, _i, _u8, _u8]),
    "orc_restrict": (None, [_dp, _u8, _i, _i, _dp]),
    "orc_prolong": (None, [_dp, _i, _i, _u8, _i, _i, _dp]),
    "orc_mg_solve": (_i, [_dp, _dp, _u8, _u8, _i, _i, _d, _d, _d, _i, _i, _i, _i, _i, _d, _i, _d, _dp]),
    "orc_mg_solve_cached": (_i, [_dp, _dp, _u8, _u8, _u8, _i, _i, _d, _d, _d, _i, _i, _i, _i, _i,
                                 _d, _d, _i, _d, _dp]),
    "orc_falloff": (_d, [_d, _d, _i]),
    "orc_emit_scalar": (None, [_dp, _i, _i, _d, _d, _d, _d, _d, _i]),
    "orc_num_threads": (_i, []),
}

_lib = None


def build(force: bool = False) -> str:
    """Compile libsmoke_oracle.so with the recipe in oracle/Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "all"], check=True,
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _f(a: np.ndarray):
    if a.dtype != np.float64 or not a.flags.c_contiguous:
        raise ValueError("oracle arrays must be C-contiguous float64")
    return a.ctypes.data_as(_dp)


def _m(a: np.ndarray):
    if a.dtype not in (np.bool_, np.uint8) or not a.flags.c_contiguous:
        raise ValueError("oracle masks must be C-contiguous bool/uint8")
    return a.ctypes.data_as(_u8)


def _mask(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a).astype(np.uint8))


# ---------------------------------------------------------------- advection
def sample(field, x, y):
    """advection.py:21-49 sample_scalar_field."""
    f = np.ascontiguousarray(field, dtype=np.float64)
    return lib().orc_sample(_f(f), f.shape[0], f.shape[1], float(x), float(y))


def advect_scalar_field(src, dst, u, v, is_fluid, dt, cell_size):
    """advection.py:112-128."""
    fl = _mask(is_fluid)
    H, W = src.shape
    lib().orc_advect_scalar(_f(src), _f(dst), H, W, _f(u), u.shape[0], u.shape[1],
                            _f(v), v.shape[0], v.shape[1], _m(fl), dt, cell_size)


def advect_velocity_u(u_src, u_dst, v, face_mask_u, dt, cell_size):
    """advection.py:130-154."""
    m = _mask(face_mask_u)
    lib().orc_advect_u(_f(u_src), _f(u_dst), u_src.shape[0], u_src.shape[1],
                       _f(v), v.shape[0], v.shape[1], _m(m), dt, cell_size)


def advect_velocity_v(v_src, v_dst, u, face_mask_v, dt, cell_size):
    """advection.py:156-180."""
    m = _mask(face_mask_v)
    lib().orc_advect_v(_f(v_src), _f(v_dst), v_src.shape[0], v_src.shape[1],
                       _f(u), u.shape[0], u.shape[1], _m(m), dt, cell_size)


# ------------------------------------------------------------------- forces
def apply_forces(v, density, temperature, dt, gravity, buoyancy_coefficient,
                 density_coefficient, ambient_temperature):
    """forces.py:78-105 (float32 force scratch, forces.py:33,62,105)."""
    H, W = density.shape
    if temperature is None:
        tp, Ht, Wt = None, H, W
    else:
        tp, (Ht, Wt) = _f(temperature), temperature.shape
    lib().orc_apply_forces(_f(v), v.shape[0], v.shape[1], _f(density), H, W, tp, Ht, Wt,
                           dt, gravity, buoyancy_coefficient, density_coefficient,
                           ambient_temperature)


def cool_temperature(T, ambient_temperature, cooling_rate, dt):
    """forces.py:108-123."""
    lib().orc_cool_temperature(_f(T), T.shape[0], T.shape[1], ambient_temperature,
                               cooling_rate, dt)


# --------------------------------------------------------------- projection
def compute_divergence(u, v, div, face_mask_u, face_mask_v, cell_size):
    """divergence.py:4-25."""
    H, W = div.shape
    lib().orc_divergence(_f(u), _f(v), _f(div), _m(_mask(face_mask_u)),
                         _m(_mask(face_mask_v)), H, W, cell_size)


def apply_pressure_gradient(u, v, p, face_mask_u, face_mask_v, fluid_density,
                            cell_size, dt):
    """pressure_gradient.py:4-37."""
    H, W = p.shape
    lib().orc_pressure_gradient(_f(u), _f(v), _f(p), _m(_mask(face_mask_u)),
                                _m(_mask(face_mask_v)), H, W, fluid_density,
                                cell_size, dt)


def enforce_boundary_conditions(u, v, solid_mask):
    """boundaries.py:7-39."""
    H, W = solid_mask.shape
    lib().orc_enforce_boundaries(_f(u), _f(v), _m(_mask(solid_mask)), H, W)


def compute_residual(p, div, is_fluid, face_mask_u, face_mask_v, alpha,
                     reciprocal_beta, residual_out):
    """residual.py:6-42."""
    H, W = p.shape
    lib().orc_residual(_f(p), _f(div), _m(_mask(is_fluid)), _m(_mask(face_mask_u)),
                       _m(_mask(face_mask_v)), H, W, alpha, reciprocal_beta,
                       _f(residual_out))
    return residual_out


# ------------------------------------------------------------------ solvers
def calculate_alpha(cell_size, density, dt):
    """solvers/common.py:6-18."""
    return lib().orc_alpha(cell_size, density, dt)


def calculate_optimal_omega(size):
    """solvers/common.py:30-39."""
    return lib().orc_optimal_omega(int(size))


def is_fluid_from_masks(face_mask_u, face_mask_v):
    """masks.py:52-74 / jacobi.py:46."""
    H, W = face_mask_u.shape[0], face_mask_u.shape[1] - 1
    out = np.zeros((H, W), dtype=np.uint8)
    lib().orc_is_fluid(_m(_mask(face_mask_u)), _m(_mask(face_mask_v)), H, W, _m(out))
    return out.astype(bool)


def jacobi_sweep(p, p_out, div, is_fluid, face_mask_u, face_mask_v, alpha,
                 reciprocal_beta):
    """jacobi.py:8-24 (_jacobi_kernel)."""
    H, W = p.shape
    lib().orc_jacobi_sweep(_f(p), _f(p_out), _f(div), _m(_mask(is_fluid)),
                           _m(_mask(face_mask_u)), _m(_mask(face_mask_v)), H, W,
                           alpha, reciprocal_beta)


def jacobi_pressure_solver(p, div, face_mask_u, face_mask_v, fluid_density,
                           cell_size, dt, iterations, **_ignored):
    """jacobi.py:27-52."""
    H, W = p.shape
    lib().orc_jacobi_solve(_f(p), _f(div), _m(_mask(face_mask_u)),
                           _m(_mask(face_mask_v)), H, W, fluid_density, cell_size,
                           dt, int(iterations))


def rbsor_sweeps(p, div, is_fluid, face_mask_u, face_mask_v, alpha,
                 reciprocal_beta, omega, iterations):
    """red_black_sor_gauss_seidel.py:8-27 (kernel)."""
    H, W = p.shape
    lib().orc_rbsor_sweeps(_f(p), _f(div), _m(_mask(is_fluid)), _m(_mask(face_mask_u)),
                           _m(_mask(face_mask_v)), H, W, alpha, reciprocal_beta,
                           omega, int(iterations))


def red_black_sor_gauss_seidel_pressure_solver(p, div, face_mask_u, face_mask_v,
                                               fluid_density, cell_size, dt,
                                               iterations, omega=0.0, **_ignored):
    """red_black_sor_gauss_seidel.py:30-42."""
    H, W = p.shape
    lib().orc_rbsor_solve(_f(p), _f(div), _m(_mask(face_mask_u)), _m(_mask(face_mask_v)),
                          H, W, fluid_density, cell_size, dt, int(iterations), float(omega))


def gauss_seidel_pressure_solver(p, div, face_mask_u, face_mask_v, fluid_density,
                                 cell_size, dt, iterations, **_ignored):
    """gauss_seidel.py:29-48."""
    H, W = p.shape
    lib().orc_gs_solve(_f(p), _f(div), _m(_mask(face_mask_u)), _m(_mask(face_mask_v)),
                       H, W, fluid_density, cell_size, dt, int(iterations))


def rbsor_solve_tol(p, div, face_mask_u, face_mask_v, fluid_density, cell_size,
                    dt, omega, tol, check_every, max_sweeps):
    """Tolerance-driven RB-SOR (new; the reference runs fixed sweeps).

    Returns (sweeps, relative residual)."""
    H, W = p.shape
    rel = ctypes.c_double(0.0)
    n = lib().orc_rbsor_solve_tol(_f(p), _f(div), _m(_mask(face_mask_u)),
                                  _m(_mask(face_mask_v)), H, W, fluid_density,
                                  cell_size, dt, float(omega), float(tol),
                                  int(check_every), int(max_sweeps), ctypes.byref(rel))
    return n, rel.value


def residual_norm2(r, is_fluid):
    H, W = r.shape
    return lib().orc_norm2_fluid(_f(r), _m(_mask(is_fluid)), H, W)


# ---------------------------------------------------------------- multigrid
def coarsen_solid_mask(solid_fine):
    """masks.py:3-27."""
    H, W = solid_fine.shape
    out = np.zeros((H // 2, W // 2), dtype=np.uint8)
    lib().orc_coarsen_solid(_m(_mask(solid_fine)), H, W, _m(out))
    return out.astype(bool)


def build_face_masks(solid):
    """masks.py:30-49."""
    H, W = solid.shape
    fmu = np.zeros((H, W + 1), dtype=np.uint8)
    fmv = np.zeros((H + 1, W), dtype=np.uint8)
    lib().orc_build_face_masks(_m(_mask(solid)), H, W, _m(fmu), _m(fmv))
    return fmu.astype(bool), fmv.astype(bool)


def restrict(fine, is_fluid_fine):
    """transfer.py:5-36."""
    H, W = fine.shape
    out = np.zeros((H // 2, W // 2), dtype=np.float64)
    lib().orc_restrict(_f(fine), _m(_mask(is_fluid_fine)), H, W, _f(out))
    return out


def prolong(coarse, is_fluid_fine):
    """transfer.py:38-72."""
    H, W = is_fluid_fine.shape
    out = np.zeros((H, W), dtype=np.float64)
    lib().orc_prolong(_f(coarse), coarse.shape[0], coarse.shape[1],
                      _m(_mask(is_fluid_fine)), H, W, _f(out))
    return out


SMOOTHERS = {"jacobi": 0, "gauss_seidel": 1, "rb_sor_gauss_seidel": 2}


def multigrid_solver(p, div, face_mask_u, face_mask_v, fluid_density, cell_size,
                     dt, v_cycles=2, smoother_type="jacobi", num_levels=4,
                     smoother_iterations=2, coarsest_iterations=50,
                     smoother_omega=0.0, coarse_scale="reference", tolerance=0.0,
                     level_is_fluid=None, coarsest_omega=0.0, **_ignored):
    """multigrid/solver.py:141-176 (fresh cache per call).

    ``coarse_scale="reference"`` keeps solver.py:124-125 (restrict(r)/alpha(2h));
    ``"rediscretized"`` divides by alpha(h) instead (new option).

    ``level_is_fluid``: level 0's is_fluid as an earlier call's cache built
    it (solver.py:87-92 without a rebuild), while the face masks passed in
    are the arrays that cache holds, changed in place since: the smoothers
    see the masks, the level-0 residual, restriction and prolongation and
    the coarse levels the cached is_fluid.  None: a fresh cache.

    ``coarsest_omega`` (new): the red-black omega of the coarsest level only,
    > 0 as given, < 0 the optimal omega of that level's size, 0 (the
    reference) the other levels' omega.

    Returns (v-cycles run, relative residual or -1 when no tolerance)."""
    H, W = p.shape
    rel = ctypes.c_double(0.0)
    lf = None if level_is_fluid is None else _mask(level_is_fluid)
    if lf is not None and lf.shape != (H, W):
        raise ValueError("level_is_fluid must have the shape of p")
    n = lib().orc_mg_solve_cached(_f(p), _f(div), _m(_mask(face_mask_u)), _m(_mask(face_mask_v)),
                                  None if lf is None else _m(lf), H, W, fluid_density,
                                  cell_size, dt, int(v_cycles), SMOOTHERS[smoother_type],
                                  int(num_levels), int(smoother_iterations),
                                  int(coarsest_iterations), float(smoother_omega),
                                  float(coarsest_omega),
                                  1 if coarse_scale == "rediscretized" else 0, float(tolerance),
                                  ctypes.byref(rel))
    return n, rel.value


# ----------------------------------------------------------------- emitters
APPROACHES = {"hard_cutoff": 0, "linear": 1, "quadratic": 2, "gaussian": 3}


def falloff(distance, radius, approach):
    """emitters.py:6-35 / :339-411."""
    return lib().orc_falloff(float(distance), float(radius), APPROACHES[approach])


def emit_scalar(field, value, dt, pos_x, pos_y, radius, approach):
    """emitters.py:157-178 (python path, every cell)."""
    H, W = field.shape
    lib().orc_emit_scalar(_f(field), H, W, float(value), float(dt), float(pos_x),
                          float(pos_y), float(radius), APPROACHES[approach])


def num_threads() -> int:
    return lib().orc_num_threads()


__all__ = [n for n in dir() if not n.startswith("_")]
