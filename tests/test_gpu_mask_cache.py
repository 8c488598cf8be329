"""State the simulation keeps between steps, against the oracle rollout.

* The multigrid solver's level cache (solver.py:87-92) across an obstacle
  set mid-run without ``rebuild_masks``: the reference's level 0 keeps the
  simulation's face-mask arrays, which the substep updates in place
  (simulation.py:198-204), so its level-0 smoother sees the obstacle while
  the level-0 residual, restriction and prolongation keep the is_fluid of
  the first solve and the coarse levels stay as built.  The oracle caches
  the same way (``OracleSmokeSim.mg_level_fluid``).  Two grids: one whose
  level 0 is smoothed by the tiled kernels, one small enough for the
  single-workgroup kernel to own every level.  Each shape is used by no
  other test, because the solver instance is shared by the whole process,
  as the reference's registered instance is (solvers/registrations.py:20).
* Temperature written into ``sim.temperature`` directly, and by an emitter
  with no ``writes_temperature`` attribute, reaches forces, advection and
  cooling (the reference always transports the array,
  simulation.py:222-227).
"""
import numpy as np
import pytest
import torch

import oracle.sim
from cfd_amd.common.config import Backend, GridConfig, PressureSolverConfig, SimulationConfig
from cfd_amd.common.emitters import Emitter2D
from cfd_amd.common.simulation import SmokeSimulation2D
from cfd_amd.solvers.params import SOLVER_PARAM_TYPES

pytestmark = pytest.mark.gpu

FIELDS = ("density", "temperature", "velocity_u", "velocity_v", "pressure", "divergence",
          "pressure_residual")


def make_pair(W, H, method, params):
    cfg = SimulationConfig(
        grid=GridConfig(size=(W, H), cell_size=1.0),
        pressure_solver=PressureSolverConfig(method=method, backend=Backend.HIP,
                                             params=SOLVER_PARAM_TYPES[method](**params)),
        dtype="float64", device="cuda")
    sim = SmokeSimulation2D(cfg)
    ref = oracle.sim.OracleSmokeSim(H, W, 1.0, solver=method, solver_params=params)
    return sim, ref


def add_emitter(sim, ref, **e):
    sim.add_emitter(Emitter2D(e["position"], e["radius"], e["density_value"],
                              e["temperature_value"], e["velocity_value"], e["approach"]))
    ref.emitters.append(dict(e))


def obstacle(W, H):
    y, x = np.mgrid[0:H, 0:W]
    m = (x - 0.55 * W) ** 2 + (y - 0.5 * H) ** 2 <= (0.12 * min(W, H)) ** 2
    m |= (x >= W // 5) & (x < W // 5 + 3) & (y >= H // 3) & (y < 2 * H // 3)
    m[0, :] = m[-1, :] = True
    m[:, 0] = m[:, -1] = True
    return m


def compare(sim, ref):
    torch.cuda.synchronize()
    bad = {}
    for f in FIELDS:
        got, want = getattr(sim, f).cpu().numpy(), getattr(ref, f)
        if not np.array_equal(got, want):
            bad[f] = (int(np.count_nonzero(got != want)), float(np.abs(got - want).max()))
    return bad


@pytest.mark.parametrize("W,H", [(264, 200), (40, 24)])
@pytest.mark.parametrize("smoother", ["jacobi", "rb_sor_gauss_seidel"])
def test_multigrid_level_cache_across_obstacle(W, H, smoother):
    params = {"v_cycles": 2, "smoother_type": smoother, "num_levels": 4,
              "smoother_iterations": 2, "coarsest_iterations": 30}
    sim, ref = make_pair(W, H, "multigrid", params)
    add_emitter(sim, ref, position=(0.3 * W, 0.2 * H), radius=0.1 * H, density_value=1.0,
                temperature_value=1.5, velocity_value=(0.5, 3.0), approach="linear")
    m = obstacle(W, H)
    for step in range(8):
        if step == 3:
            sim.set_solid_mask(m)
            ref.set_solid_mask(m)
        sim.substep(0.016)
        ref.substep(0.016)
    bad = compare(sim, ref)
    assert not bad, bad
    # the cache state matters: a rebuilt hierarchy gives other pressures
    fresh = oracle.sim.OracleSmokeSim(H, W, 1.0, solver="multigrid",
                                      solver_params=dict(params, rebuild_masks=True))
    fresh.emitters = [dict(e) for e in ref.emitters]
    for step in range(8):
        if step == 3:
            fresh.set_solid_mask(m)
        fresh.substep(0.016)
    assert not np.array_equal(fresh.pressure, ref.pressure)


def test_temperature_written_directly_is_transported():
    W, H = 72, 56
    sim, ref = make_pair(W, H, "jacobi", {"iterations": 30})
    add_emitter(sim, ref, position=(36.0, 10.0), radius=5.0, density_value=1.0,
                temperature_value=0.0, velocity_value=(0.0, 0.0), approach="linear")
    for step in range(6):
        if step == 2:
            sim.temperature[20:30, 30:42] = 2.0
            ref.temperature[20:30, 30:42] = 2.0
        sim.substep(0.016)
        ref.substep(0.016)
    bad = compare(sim, ref)
    assert not bad, bad
    assert np.abs(ref.temperature).max() > 0.5


class _HeatPatch:
    """An emitter in the reference's duck-typed sense (emit(...) only): it
    adds heat to a patch through torch."""

    def emit(self, density, velocity_u, velocity_v, temperature, dt):
        temperature[8:14, 20:30] += 3.0 * dt


def test_duck_typed_emitter_receives_temperature():
    W, H = 64, 48
    sim, ref = make_pair(W, H, "jacobi", {"iterations": 30})
    sim.add_emitter(_HeatPatch())
    for _ in range(5):
        sim.substep(0.016)
        ref.temperature[8:14, 20:30] += 3.0 * 0.016  # the emitter stage opens the step
        ref.substep(0.016)
    bad = compare(sim, ref)
    assert not bad, bad
    assert np.abs(ref.temperature).max() > 0.1
