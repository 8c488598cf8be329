"""The single-scene autotuner (cfd_amd.tuning): it restores the simulation's
state, its tile choices leave results bit for bit unchanged, its choice is
cached per grid and solver configuration, and a sweeps search converges."""
import pytest
import torch

from cfd_amd import scenes, tuning
from cfd_amd.backends.hip import _lib

pytestmark = pytest.mark.gpu

MIXED = {"precision": "mixed", "inner_cycles": 2, "smoother_omega": 1.2,
         "coarsest_solve": "exact_1024",
         "inner_arithmetic": "float32"}
FIELDS = ("density", "velocity_u", "velocity_v", "pressure", "divergence", "pressure_residual")


@pytest.fixture
def keep_variants():
    """Variants are process-wide: put back what the test found."""
    saved = {(c, w): _lib.lib().smk_get_smoother_variant(c, _lib._WHICH[w])
             for c in (_lib.SMK_F64, _lib.SMK_F32) for w in _lib._WHICH}
    yield
    for (c, w), v in saved.items():
        _lib.call("smk_set_smoother_variant", c, _lib._WHICH[w], v)


@pytest.mark.parametrize("overrides", [None, MIXED])
def test_autotune_keeps_state_and_results(keep_variants, overrides):
    sc = scenes.plume_2048_mg(N=256, steps=4)
    ref = scenes.make_simulation(sc, dtype="float64", solver_overrides=overrides)
    sim = scenes.make_simulation(sc, dtype="float64", solver_overrides=overrides)
    ref.substep(sc.dt)
    sim.substep(sc.dt)
    for _ in range(3):  # the reference steps run before any tuning
        ref.substep(sc.dt)
    torch.cuda.synchronize()
    want = {f: getattr(ref, f).clone() for f in FIELDS}
    before = {f: getattr(sim, f).clone() for f in FIELDS}
    res = tuning.autotune(sim, steps=1, repeats=1, use_cache=False)
    assert set(res["chosen"]) == set(tuning.CANDIDATES)
    assert len(res["candidates"]) == sum(len(c) for c in tuning.CANDIDATES.values())
    torch.cuda.synchronize()
    for f in FIELDS:
        assert torch.equal(getattr(sim, f), before[f]), f"autotune changed {f}"
    for _ in range(3):
        sim.substep(sc.dt)
    torch.cuda.synchronize()
    for f in FIELDS:
        assert torch.equal(getattr(sim, f), want[f]), f"{f} differs under the tuned variants"
    again = tuning.autotune(sim, steps=1, repeats=1)
    assert again["cached"] and again["chosen"] == res["chosen"]


def test_autotune_sweeps(keep_variants):
    sc = scenes.plume_2048_mg(N=128, steps=3)
    sim = scenes.make_simulation(sc, dtype="float64", solver_overrides=MIXED)
    sim.substep(sc.dt)
    res = tuning.autotune(sim, steps=1, repeats=1, classes=("fine",), sweeps=(1, 2, 3),
                          use_cache=False)
    assert res["chosen"]["smoother_iterations"] in (1, 2, 3)
    assert sim.solver_params.smoother_iterations == res["chosen"]["smoother_iterations"]
    sim.substep(sc.dt)
    assert sim.last_solve_info["converged"] == 1


def test_restore_defaults(keep_variants):
    """restore_defaults() puts back the variants the process started with, which
    autotune() changes process-wide."""
    def snapshot():
        return {(c, w): _lib.lib().smk_get_smoother_variant(c, _lib._WHICH[w])
                for c in (_lib.SMK_F64, _lib.SMK_F32) for w in tuning.CANDIDATES}

    start = snapshot()
    sc = scenes.plume_2048_mg(N=128, steps=2)
    sim = scenes.make_simulation(sc, dtype="float64", solver_overrides=MIXED)
    sim.substep(sc.dt)
    tuning.autotune(sim, steps=1, repeats=1, use_cache=False)
    for code in (_lib.SMK_F64, _lib.SMK_F32):
        _lib.set_smoother_variant(code, "fine", "lds32")
        _lib.set_smoother_variant(code, "small", "nt256")
    assert snapshot() != start
    tuning.restore_defaults()
    assert snapshot() == start
