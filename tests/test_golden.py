"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py
from the C oracle): the oracle must still reproduce them bit for bit (CPU),
and the HIP library must reproduce every operator and fixed-count solver
fixture bit for bit in float64 (GPU)."""
import os
import sys

import numpy as np
import pytest

from impls import OracleImpl

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, HERE)
import make_golden  # noqa: E402

SIZES = (16, 32)
SEEDS = {16: 160, 32: 320}


def load(name):
    with np.load(os.path.join(HERE, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("N", SIZES)
def test_oracle_reproduces_operator_fixtures(N):
    ref = load(f"ops_{N}.npz")
    now = make_golden.ops_fixture(N, SEEDS[N])
    assert sorted(ref) == sorted(now)
    for k in ref:
        np.testing.assert_array_equal(now[k], ref[k], err_msg=k)


def test_oracle_reproduces_step_fixture():
    ref = load("step_64.npz")
    now = make_golden.step_fixture()
    for k in ref:
        np.testing.assert_array_equal(now[k], ref[k], err_msg=k)


def _run_ops(impl, g):
    h, dt = float(g["h"]), float(g["dt"])
    out = {}
    a = np.zeros_like(g["rho"])
    impl.advect_scalar_field(g["rho"], a, g["u"], g["v"], g["is_fluid"], dt, h)
    out["advect_scalar"] = a
    au = np.zeros_like(g["u"])
    impl.advect_velocity_u(g["u"], au, g["v"], g["fmu"], dt, h)
    out["advect_u"] = au
    av = np.zeros_like(g["v"])
    impl.advect_velocity_v(g["v"], av, g["u"], g["fmv"], dt, h)
    out["advect_v"] = av
    vf = g["v"].copy()
    impl.apply_forces(vf, g["rho"], None, dt, -1.0, 5.0, 1.0, 0.0)
    out["apply_forces_v"] = vf
    d = np.zeros_like(g["rho"])
    impl.compute_divergence(g["u"], g["v"], d, g["fmu"], g["fmv"], h)
    out["divergence"] = d
    gu, gv = g["u"].copy(), g["v"].copy()
    impl.apply_pressure_gradient(gu, gv, g["p"], g["fmu"], g["fmv"], 1.0, h, dt)
    impl.enforce_boundary_conditions(gu, gv, g["solid"])
    out["gradient_u"], out["gradient_v"] = gu, gv
    alpha = (-(h * h) * 1.0) / dt
    r = np.zeros_like(g["p"])
    impl.compute_residual(g["p"], g["div"], g["is_fluid"], g["fmu"], g["fmv"], alpha, 0.25, r)
    out["residual"] = r
    args = (g["div"], g["fmu"], g["fmv"], 1.0, h, dt)
    for key, method, kw in (
            ("jacobi_50", "jacobi", dict(iterations=50)),
            ("rbsor_30_auto", "rb_sor_gauss_seidel", dict(iterations=30, omega=0.0)),
            ("multigrid_reference_defaults", "multigrid",
             dict(v_cycles=2, smoother_type="jacobi", num_levels=3, smoother_iterations=2,
                  coarsest_iterations=30)),
            ("multigrid_rbgs_rediscretized_4", "multigrid",
             dict(v_cycles=4, smoother_type="rb_sor_gauss_seidel", num_levels=3,
                  smoother_iterations=2, coarsest_iterations=50, smoother_omega=1.0,
                  coarse_scale="rediscretized"))):
        p = np.zeros_like(g["p"])
        impl.solve(method, p, *args, **kw)
        out[key] = p
    return out


@pytest.mark.parametrize("N", SIZES)
def test_oracle_impl_adapter_matches_fixtures(N):
    """The numpy adapter used by the GPU test, on the oracle."""
    g = load(f"ops_{N}.npz")
    for k, v in _run_ops(OracleImpl(), g).items():
        np.testing.assert_array_equal(v, g[k], err_msg=k)


@pytest.mark.gpu
@pytest.mark.parametrize("N", SIZES)
def test_hip_matches_operator_fixtures(N):
    from impls import HipImpl

    g = load(f"ops_{N}.npz")
    for k, v in _run_ops(HipImpl(), g).items():
        np.testing.assert_array_equal(v, g[k], err_msg=k)


@pytest.mark.gpu
def test_hip_step_matches_step_fixture():
    """Three steps of the 64x64 plume (multigrid to 1e-8): equal to the
    fixture to 1e-9 relative (the tolerance tests may stop on either side of
    a cycle boundary only if a norm lands exactly on the tolerance)."""
    import torch

    from cfd_amd import scenes

    g = load("step_64.npz")
    sc = scenes.plume_2048_mg(N=64, steps=3)
    sim = scenes.make_simulation(sc, dtype="float64", device="cuda")
    np.testing.assert_array_equal(sim.velocity_u.cpu().numpy(), g["u0"])
    for _ in range(3):
        sim.substep(sc.dt)
    torch.cuda.synchronize()
    for k, t in (("density", sim.density), ("u", sim.velocity_u), ("v", sim.velocity_v)):
        got = t.cpu().numpy()
        err = np.abs(got - g[k]).max() / max(np.abs(g[k]).max(), 1e-300)
        assert err <= 1e-9, (k, err)
