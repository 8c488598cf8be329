"""The reference's profiling API (src/cfd/profiling) on the GPU step: config
vocabulary and validation (CPU), device reducers against numpy's, and the
session's JSON layout from a profiled simulation (GPU)."""
import json

import numpy as np
import pytest

from cfd_amd import profiling as P


def test_full_config_validates_and_round_trips(tmp_path):
    cfg = P.ProfileConfig.full()
    cfg.validate_config()
    assert sorted(cfg.timings) == sorted(P.TIMINGS)
    assert {f.source for f in cfg.metrics.fields} == {"divergence", "pressure",
                                                      "pressure_residual"}
    path = tmp_path / "cfg.json"
    path.write_text(json.dumps(cfg.to_dict()))
    assert P.ProfileConfig.from_json(str(path)).to_dict() == cfg.to_dict()


@pytest.mark.parametrize("bad,match", [
    (P.FieldMetricConfig(source="density"), "Unknown field metric source"),
    (P.FieldMetricConfig(source="pressure", stages=["initial"]), "Invalid stages"),
    (P.FieldMetricConfig(source="pressure", reducers=["median"]), "Invalid reducers"),
    (P.FieldMetricConfig(source="pressure", scopes=["edges"]), "Invalid scopes"),
    (P.FieldMetricConfig(source="pressure", every=0), "positive integer"),
    (P.ScalarMetricConfig(name="solver.residual"), "Unknown scalar metric"),
])
def test_validation_messages(bad, match):
    with pytest.raises(ValueError, match=match):
        bad.validate_config()


def test_disabled_session_records_nothing():
    s = P.ProfileSession.disabled()
    assert not s.enabled
    with s.timing.scope("simulation.substep"):
        pass
    s.metrics.capture("pressure", "final", None)
    s.metrics.process()
    assert s.timing.to_dict() == {} and s.metrics.to_dict() == {"scalars": {}, "fields": {}}


def test_variable_summary_matches_numpy():
    v = [3.0, 1.0, 4.0, 1.0, 5.0, 9.0, 2.0, 6.0]
    s = P.VariableSummary.from_values("x", v).to_dict()
    a = np.asarray(v)
    assert s["count"] == 8 and s["total"] == a.sum() and s["std"] == a.std()
    assert s["p95"] == np.percentile(a, 95)


def _numpy_reducers(a):
    H, W = a.shape
    interior = a[1:-1, 1:-1]
    mask = np.ones_like(a, dtype=bool)
    mask[1:-1, 1:-1] = False
    out = []
    for x in (a, interior, a[mask]):
        out.append([np.sqrt(np.mean(np.square(x))), np.mean(x), np.mean(np.abs(x)), np.max(x),
                    np.max(np.abs(x)), np.std(x)])
    return np.array(out)


@pytest.mark.gpu
@pytest.mark.parametrize("shape", [(37, 53), (64, 65), (513, 512), (3, 3)])
@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_device_reducers_match_numpy(shape, dtype):
    import torch

    rng = np.random.default_rng(shape[0])
    a = (rng.standard_normal(shape) * 3 + 0.5).astype(dtype)
    got = P.field_stats(torch.from_numpy(a).cuda()).cpu().numpy()
    exp = _numpy_reducers(a.astype(np.float64))
    np.testing.assert_allclose(got, exp, rtol=1e-12, atol=1e-15)


@pytest.mark.gpu
def test_profiled_simulation_exports_reference_layout(tmp_path):
    import torch

    from cfd_amd import scenes

    sc = scenes.plume_2048_mg(N=64, steps=3)
    sim = scenes.make_simulation(sc, dtype="float64", device="cuda")
    sim.profiler = P.ProfileSession.full(grid_shape=(64, 64))
    for _ in range(3):
        sim.substep(sc.dt)
    torch.cuda.synchronize()
    d = sim.profiler.to_dict()
    assert set(d) == {"metadata", "checkpoints", "variables", "snapshots"}
    assert d["metadata"]["schema_version"] == 1
    assert d["checkpoints"]["simulation.substep"]["calls"] == 3
    assert d["checkpoints"]["simulation.pressure.solve"]["total_ms"] > 0
    f = d["variables"]["fields"]
    key = "divergence.after_projection.all.max_abs"
    assert f[key]["calls"] == 3
    exp = float(np.abs(sim.divergence.cpu().numpy()).max())
    assert abs(f[key]["last"] - exp) <= 1e-12 * max(exp, 1e-300)
    assert "pressure_residual.final.interior.rms" in f
    sim.profiler.export_json(str(tmp_path / "p.json"))
    assert json.loads((tmp_path / "p.json").read_text())["variables"]["fields"][key]["calls"] == 3
