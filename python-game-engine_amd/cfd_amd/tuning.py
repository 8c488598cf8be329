"""Autotuner of the single-scene pressure solve.

The multigrid and RB-SOR solves run their smoothing passes through tile
kernels whose shape is a launch choice: per level class (the finest level,
coarse levels above 512x512 cells, coarse levels up to 512x512 cells) a
64-column region of 16 to 64 rows with p, rhs and flags in LDS, or p in LDS
and rhs and flags in registers, or p staged by LDS-DMA; and the workgroup of
the single-workgroup kernel that takes the levels from 32x32 down.  Every
choice gives the same results bit for bit (they reorder nothing), but their
speed depends on the grid and dtype.  :func:`autotune` times whole steps of
a simulation for each candidate, one level class at a time (coordinate
descent), restores the simulation's state after each trial, keeps the
fastest and remembers it per grid shape, dtype and solver configuration.

With ``sweeps`` given it also tries smoothing sweep counts per level
(``smoother_iterations``, the V(s,s) of multigrid/solver.py:107,137) on
the time per step to the solver's tolerance: those change the iterates, so
they are only tried when asked for.  :class:`cfd_amd.batch.BatchedSmoke`
has the batch's own autotuner.
"""
from __future__ import annotations

import torch

from cfd_amd.backends.hip import _lib

# candidates per level class (names of _lib.SMOOTHER_VARIANTS / SMALL_VARIANTS)
CANDIDATES = {
    "fine": ("lds64r", "dma64", "lds48r", "lds32r"),
    "coarse": ("lds32r", "dma32", "lds32", "lds48r", "lds16"),
    "coarse_small": ("lds16", "dma16", "lds32", "lds32r"),
    "small": ("nt1024", "nt512", "nt256", "nt512w", "nt1024c"),
}

_CACHE: dict = {}
# the variants the process started with, per (dtype code, level class)
_DEFAULTS: dict = {}


def _remember_defaults() -> None:
    if _DEFAULTS:
        return
    for code in (_lib.SMK_F64, _lib.SMK_F32):
        for which in CANDIDATES:
            _DEFAULTS[(code, which)] = _current(code, which)


def restore_defaults() -> None:
    """Put back the tile variants the process started with.  The variants are
    process-wide per dtype, so a choice :func:`autotune` made for one grid
    otherwise stays in force for every later solve of that dtype."""
    for (code, which), name in _DEFAULTS.items():
        _lib.set_smoother_variant(code, which, name)


def _names(which):
    return _lib.SMALL_VARIANTS if which == "small" else _lib.SMOOTHER_VARIANTS


def _current(code, which):
    v = _lib.lib().smk_get_smoother_variant(code, _lib._WHICH[which])
    return next(k for k, x in _names(which).items() if x == v)


def _solve_key(sim):
    p = sim.solver_params
    fields = ("precision", "inner_cycles", "coarsest_solve", "inner_arithmetic", "tolerance",
              "num_levels", "smoother_type", "coarse_scale", "smoother_omega")
    return (sim.pressure_solver_method,) + tuple(getattr(p, f, None) for f in fields)


def _variant_dtype(sim) -> int:
    """The dtype whose tile variants the solve runs: the FP32 V-cycles of a
    mixed-precision solve, else the fields' dtype."""
    if getattr(sim.solver_params, "precision", "native") == "mixed":
        return _lib.SMK_F32
    return _lib.SMK_F64 if sim.dtype == torch.float64 else _lib.SMK_F32


class _Saved:
    """The simulation's state: field contents and which buffer holds which
    field (the transport swaps), restored after every trial."""

    def __init__(self, sim):
        self.sim = sim
        names = sim._BUFFERS + ("pressure", "divergence", "pressure_residual")
        self.refs = {n: getattr(sim, n) for n in names}
        self.data = {n: t.clone() for n, t in self.refs.items()}
        self.flags = (sim._temperature_live, sim.accumulator)
        self.params = sim.solver_params

    def restore(self):
        for n, t in self.refs.items():
            setattr(self.sim, n, t)
            t.copy_(self.data[n])
        self.sim._temperature_live, self.sim.accumulator = self.flags
        if not self.sim._temperature_live:  # the copies above are no write of the user's
            self.sim._watch_temperature()
        self.sim.solver_params = self.params


def _time_steps(sim, dt, steps, repeats):
    best = float("inf")
    for _ in range(repeats):
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(steps):
            sim.substep(dt)
        b.record()
        torch.cuda.synchronize()
        best = min(best, a.elapsed_time(b) / steps)
    return best


def autotune(sim, dt=None, *, steps: int = 2, repeats: int = 2, classes=tuple(CANDIDATES),
             sweeps=None, use_cache: bool = True) -> dict:
    """Pick the tile variant of each level class (and, if ``sweeps`` is a
    sequence of counts, the smoothing sweeps per level) for ``sim`` by
    timing ``steps`` substeps per candidate (best of ``repeats``) from the
    simulation's current state, which is restored afterwards.  The choice is
    applied (variants are process-wide per dtype) and returned:
    ``{"chosen": {...}, "ms_per_step": t, "candidates": [(setting, ms), ...],
    "cached": bool}``."""
    from dataclasses import replace

    dt = sim.dt if dt is None else dt
    _remember_defaults()
    code = _variant_dtype(sim)
    key = (sim.height, sim.width, str(sim.dtype), _solve_key(sim), tuple(classes),
           tuple(sweeps) if sweeps else None)
    if use_cache and key in _CACHE:
        res = _CACHE[key]
        _apply(sim, code, res["chosen"])
        return dict(res, cached=True)
    saved = _Saved(sim)
    chosen = {w: _current(code, w) for w in classes}
    trials = []
    try:
        # one untimed step per trial configuration warms its kernels
        for which in classes:
            best = None
            for cand in CANDIDATES[which]:
                _lib.set_smoother_variant(code, which, cand)
                sim.substep(dt)
                saved.restore()
                ms = _time_steps(sim, dt, steps, repeats)
                saved.restore()
                trials.append(({which: cand}, ms))
                if best is None or ms < best[1]:
                    best = (cand, ms)
            chosen[which] = best[0]
            _lib.set_smoother_variant(code, which, best[0])
        if sweeps:
            best = None
            for s in sweeps:
                sim.solver_params = replace(saved.params, smoother_iterations=int(s))
                sim.substep(dt)
                saved.restore()
                sim.solver_params = replace(saved.params, smoother_iterations=int(s))
                ms = _time_steps(sim, dt, steps, repeats)
                saved.restore()
                trials.append(({"smoother_iterations": int(s)}, ms))
                if best is None or ms < best[1]:
                    best = (int(s), ms)
            chosen["smoother_iterations"] = best[0]
    finally:
        saved.restore()
    _apply(sim, code, chosen)
    ms = min(t for _, t in trials) if trials else None
    res = {"chosen": chosen, "ms_per_step": ms, "candidates": trials}
    _CACHE[key] = res
    return dict(res, cached=False)


def _apply(sim, code, chosen):
    from dataclasses import replace

    for which, cand in chosen.items():
        if which == "smoother_iterations":
            sim.solver_params = replace(sim.solver_params, smoother_iterations=int(cand))
        else:
            _lib.set_smoother_variant(code, which, cand)
