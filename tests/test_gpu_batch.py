"""Batched uneven scenes (cfd_amd.batch) vs the same scenes stepped one at a
time by SmokeSimulation2D with the same solver settings: the library runs
the same kernels with the same arguments per scene, so the fields must be
bit-identical, with and without HIP graphs and for every tile height."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _scenes(count=5, n_max=192, seed=4100):
    from cfd_amd import scenes
    return scenes.uneven_scenes(count=count, steps=3, base_seed=seed, n_min=48, n_max=n_max)


def _single(sc, steps, omega=None):
    from cfd_amd import scenes
    over = {"precision": "mixed", "inner_cycles": 1}
    if omega is not None:
        over["smoother_omega"] = omega
    sim = scenes.make_simulation(sc, dtype="float64", device="cuda", solver_overrides=over)
    for _ in range(steps):
        sim.substep(sc.dt)
    torch.cuda.synchronize()
    return sim


@pytest.fixture(scope="module")
def scene_list():
    return _scenes()


# omega 1.2: the smoother's SOR weight reaches every batched level as it
# does the single-scene solve's (smoother_omega)
@pytest.mark.parametrize("graphs,omega", [(True, None), (False, None), (True, 1.2)])
def test_batch_equals_single_scene_steps(scene_list, graphs, omega):
    from cfd_amd.batch import BatchedSmoke
    steps = 3
    kw = {} if omega is None else {"smoother_omega": omega}
    b = BatchedSmoke(scene_list, tolerance=1e-6, max_cycles=400, use_graphs=graphs, **kw)
    for _ in range(steps):
        info = b.step()
        assert info["max_rel_residual"] <= 1e-6
    torch.cuda.synchronize()
    for k, sc in enumerate(scene_list):
        sim = _single(sc, steps, omega)
        assert torch.equal(b.density(k), sim.density), f"scene {k} ({sc.H}x{sc.W}) density"
        u, v = b.velocity(k)
        assert torch.equal(u, sim.velocity_u), f"scene {k} u"
        assert torch.equal(v, sim.velocity_v), f"scene {k} v"
        assert torch.equal(b.pressure(k), sim.pressure), f"scene {k} pressure"
    b.close()


def test_tile_height_and_grid_size_do_not_change_results(scene_list):
    from cfd_amd.batch import BatchedSmoke
    out = []
    for rh, pb in ((48, 0), (32, 64)):
        b = BatchedSmoke(scene_list, tile_rows=rh, persistent_blocks=pb)
        b.step()
        b.step()
        torch.cuda.synchronize()
        out.append([b.density(k).clone() for k in range(len(scene_list))])
        b.close()
    for a, c in zip(*out):
        assert torch.equal(a, c)


def test_autotune_keeps_state_and_converges():
    from cfd_amd.batch import BatchedSmoke
    sl = _scenes(count=4, n_max=256, seed=4200)
    b = BatchedSmoke(sl)
    b.step()
    before = [b.density(k).clone() for k in range(len(sl))]
    res = b.autotune(tile_rows=(48, 32), smoother_iterations=(1, 2), persistent_blocks=(512,),
                     repeats=1)
    assert len(res["candidates"]) == 4
    after = [b.density(k) for k in range(len(sl))]
    for x, y in zip(before, after):
        assert torch.equal(x, y)
    info = b.step()
    assert info["max_rel_residual"] <= 1e-6
    cyc, rel = b.scene_stats()
    assert all(c >= 1 for c in cyc) and max(rel) <= 1e-6
    assert all(np.isfinite(b.density(k).cpu().numpy()).all() for k in range(len(sl)))
    b.close()


def test_large_uneven_batch_converges():
    """12 scenes up to 1024 x 2048 (the size range of the 96-scene workload)."""
    from cfd_amd import scenes
    from cfd_amd.batch import BatchedSmoke
    sl = scenes.uneven_scenes(count=12, steps=2, base_seed=4000)
    b = BatchedSmoke(sl)
    for _ in range(2):
        info = b.step()
        assert info["max_rel_residual"] <= 1e-6
    for k in range(len(sl)):
        d = b.density(k)
        assert torch.isfinite(d).all()
        assert float(d.max()) > 0.0
    b.close()


def test_batch_scenes_against_fp64_oracle():
    """The first 24 scenes of the 96-scene workload generator in one batch
    (FP32 V-cycles in FP64 refinement to 1e-6, coarsest level SOR at its
    optimal omega), 3 steps; the three smallest scenes against the all-FP64
    oracle rollout with the same solver settings: every scene converges and
    density and velocity are within 1e-6 relative L2 of the oracle's."""
    import oracle.sim
    from cfd_amd import scenes
    from cfd_amd.batch import BatchedSmoke

    sl = scenes.uneven_scenes(count=24, steps=3)
    b = BatchedSmoke(sl, tolerance=1e-6)
    for _ in range(3):
        info = b.step()
        assert info["max_rel_residual"] <= 1e-6
    torch.cuda.synchronize()
    picks = sorted(range(len(sl)), key=lambda k: sl[k].H * sl[k].W)[:3]
    for k in picks:
        sc = sl[k]
        ref = oracle.sim.from_scene(sc)
        for _ in range(3):
            ref.substep(sc.dt)
        u, v = b.velocity(k)
        errs = {"density": _rel(b.density(k), ref.density), "velocity_u": _rel(u, ref.velocity_u),
                "velocity_v": _rel(v, ref.velocity_v)}
        print(f"scene {k} ({sc.H}x{sc.W}) batch vs FP64 oracle:", errs)
        for f, e in errs.items():
            assert e <= 1e-6, (k, f, e)
    b.close()


def _rel(t, want):
    got = t.cpu().numpy()
    return float(np.linalg.norm(got - want) / max(np.linalg.norm(want), 1e-300))
