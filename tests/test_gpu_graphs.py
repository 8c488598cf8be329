"""SmokeSimulation2D with the stages before and after the pressure solve
replayed from HIP graphs (use_graphs) against the same simulation run
eagerly: every field is bit-identical after every step.  Replays alternate
with eager work (the solve, the other simulation's steps) and host
synchronisations; changes of dt, emitters, temperature and obstacles
select or capture other graphs."""
import numpy as np
import pytest
import torch

from cfd_amd import scenes
from cfd_amd.common.emitters import Emitter2D

pytestmark = pytest.mark.gpu

FIELDS = ("density", "velocity_u", "velocity_v", "pressure", "divergence", "pressure_residual",
          "temperature")


def _pair(scene, overrides=None):
    a = scenes.make_simulation(scene, dtype="float64", solver_overrides=overrides)
    b = scenes.make_simulation(scene, dtype="float64", solver_overrides=overrides)
    b.use_graphs = True
    return a, b


def _same(a, b, where):
    torch.cuda.synchronize()
    for f in FIELDS:
        x, y = getattr(a, f), getattr(b, f)
        assert torch.equal(x, y), f"{where}: {f} differs (max {float((x - y).abs().max())})"


@pytest.mark.parametrize("overrides", [None] + [
    {"precision": "mixed", "inner_cycles": ic, "coarsest_solve": cs, "inner_arithmetic": "float32",
     "smoother_omega": w}
    for cs, ic, w in (("exact", 2, 1.0), ("exact_1024", 2, 1.2))])
def test_graph_stages_match_eager_steps(overrides):
    sc = scenes.plume_2048_mg(N=256, steps=8)
    a, b = _pair(sc, overrides)
    for k in range(8):
        a.substep(sc.dt)
        b.substep(sc.dt)
        _same(a, b, f"step {k}")
        # eager work of the program's own between replays
        _ = float(b.density.sum())
    # one graph per stage and buffer assignment (two assignments alternate)
    assert len(b._graphs) == 4


def test_graph_keys_follow_emitters_temperature_dt_and_obstacles():
    sc = scenes.plume_128_jacobi(steps=10)
    a, b = _pair(sc)
    for s in (a, b):
        s.add_emitter(Emitter2D((40.0, 30.0), 6.0, 0.5, 0.0, (0.0, 2.0), "linear"))
    step = 0
    for dt in (sc.dt, sc.dt, sc.dt, 0.5 * sc.dt, 0.5 * sc.dt):
        a.substep(dt)
        b.substep(dt)
        _same(a, b, f"step {step}")
        step += 1
    n_before = len(b._graphs)
    for s in (a, b):  # the emitter now writes temperature: temperature passes start
        s.emitters[-1].set_temperature_value(1.5)
    for _ in range(3):
        a.substep(sc.dt)
        b.substep(sc.dt)
        _same(a, b, f"step {step}")
        step += 1
    assert len(b._graphs) > n_before
    m = np.zeros((sc.H, sc.W), dtype=bool)
    m[0, :] = m[-1, :] = m[:, 0] = m[:, -1] = True
    m[50:70, 60:64] = True
    for s in (a, b):
        s.set_solid_mask(m)
        s.emitters[-1].set_enabled(False)
    for _ in range(3):
        a.substep(sc.dt)
        b.substep(sc.dt)
        _same(a, b, f"step {step}")
        step += 1
    assert float(a.temperature.abs().max()) > 0.0
