"""Pins of the FP64 tape (oracle/tape.py), the checker of the HIP adjoints.

Its forward pieces must equal the C oracle's restatement of the reference
bit for bit (same expressions, float64, no FMA); its dense Poisson solve must
agree with the oracle's multigrid solve run to convergence; the buoyancy has a
closed form; and torch's derivatives of the whole rollout must agree with
central finite differences."""
import numpy as np
import pytest
import torch

import oracle as O
from oracle import tape as T

H, W = 14, 19
DT, HS = 0.01, 1.0 / 16


def _rand(shape, seed, scale=1.0):
    return np.random.default_rng(seed).standard_normal(shape) * scale


def _masks():
    fmu = np.zeros((H, W + 1), dtype=bool)
    fmv = np.zeros((H + 1, W), dtype=bool)
    fmu[:, 1:-1] = True
    fmv[1:-1, :] = True
    return fmu, fmv


def t(a):
    return torch.from_numpy(np.ascontiguousarray(a))


@pytest.mark.parametrize("scale", [0.0, 3.0, 40.0])
def test_advection_equals_oracle(scale):
    fmu, fmv = _masks()
    fmu[3, 5] = False
    fl = np.ones((H, W), dtype=bool)
    fl[2, 7] = False
    src = _rand((H, W), 1)
    u = _rand((H, W + 1), 2, scale)
    v = _rand((H + 1, W), 3, scale)
    ref = np.zeros_like(src)
    O.advect_scalar_field(src, ref, u, v, fl, DT, HS)
    got = T.advect_scalar(t(src), t(u), t(v), t(fl), DT, HS).numpy()
    np.testing.assert_array_equal(got, ref)
    ru = np.zeros_like(u)
    O.advect_velocity_u(u, ru, v, fmu, DT, HS)
    np.testing.assert_array_equal(T.advect_u(t(u), t(v), t(fmu), DT, HS).numpy(), ru)
    rv = np.zeros_like(v)
    O.advect_velocity_v(v, rv, u, fmv, DT, HS)
    np.testing.assert_array_equal(T.advect_v(t(v), t(u), t(fmv), DT, HS).numpy(), rv)


def test_divergence_gradient_boundaries_equal_oracle():
    fmu, fmv = _masks()
    u = _rand((H, W + 1), 4)
    v = _rand((H + 1, W), 5)
    p = _rand((H, W), 6)
    div = np.zeros((H, W))
    O.compute_divergence(u, v, div, fmu, fmv, HS)
    np.testing.assert_array_equal(T.divergence(t(u), t(v), t(fmu), t(fmv), HS).numpy(), div)
    solid = np.zeros((H, W), dtype=bool)
    solid[0, :] = solid[-1, :] = solid[:, 0] = solid[:, -1] = True
    solid[5, 6] = True
    ru, rv = u.copy(), v.copy()
    O.apply_pressure_gradient(ru, rv, p, fmu, fmv, 1.0, HS, DT)
    O.enforce_boundary_conditions(ru, rv, solid)
    gu, gv = T.gradient_and_boundaries(t(u), t(v), t(p), t(fmu), t(fmv), t(solid), 1.0, HS, DT)
    np.testing.assert_array_equal(gu.numpy(), ru)
    np.testing.assert_array_equal(gv.numpy(), rv)


def test_buoyancy_closed_form_and_oracle():
    rho = 0.3 + 0.05 * np.arange(H)[:, None] + 0.01 * np.arange(W)[None, :]
    v = _rand((H + 1, W), 7)
    g, dc = -1.0, 1.0
    got = T.buoyancy(t(v), t(rho), DT, g, dc).numpy()
    d = np.empty((H, W))
    d[0], d[H - 1] = rho[0], rho[H - 1]
    d[1:H - 1] = 0.5 * (rho[:H - 2] + rho[1:H - 1])
    exp = v.copy()
    exp[:H] += DT * (-dc * d * g)
    np.testing.assert_array_equal(got, exp)
    # the reference rounds the force through float32 (forces.py:33): equal to
    # float32 precision, not bit for bit
    ref = v.copy()
    O.apply_forces(ref, rho, None, DT, g, 5.0, dc, 0.0)
    np.testing.assert_allclose(got, ref, rtol=0, atol=DT * 1e-7 * np.abs(d).max())
    assert np.array_equal(got[H], v[H])


def test_dense_poisson_matches_converged_multigrid():
    N = 32
    h = 1.0 / N
    fmu = np.zeros((N, N + 1), dtype=bool)
    fmv = np.zeros((N + 1, N), dtype=bool)
    fmu[:, 1:-1] = True
    fmv[1:-1, :] = True
    u = _rand((N, N + 1), 8) * fmu
    v = _rand((N + 1, N), 9) * fmv
    div = np.zeros((N, N))
    O.compute_divergence(u, v, div, fmu, fmv, h)
    p = np.zeros((N, N))
    O.multigrid_solver(p, div, fmu, fmv, 1.0, h, DT, v_cycles=200,
                       smoother_type="rb_sor_gauss_seidel", num_levels=2,
                       smoother_iterations=2, coarsest_iterations=50, smoother_omega=1.0,
                       coarse_scale="rediscretized", tolerance=1e-14)
    dense = T.DensePoisson(t(fmu), t(fmv), h, 1.0, DT)(t(div)).numpy()
    # equal up to the constant null-space component
    a, b = p - p.mean(), dense - dense.mean()
    assert np.abs(a - b).max() <= 1e-10 * np.abs(a).max()


def _tape_problem(N=12):
    h = 1.0 / N
    return T.TapeSmoke(N, N, h, 0.02, (4, 2, 8, 5)), h


def test_tape_gradient_matches_central_differences():
    """torch's tape through the whole rollout vs central FD (generic point:
    random initial velocity keeps backtraces off grid lines)."""
    model, h = _tape_problem()
    N = model.H
    rng = np.random.default_rng(11)
    u0 = t(rng.standard_normal((N, N + 1)) * 0.5).requires_grad_()
    v0 = t(rng.standard_normal((N + 1, N)) * 0.5).requires_grad_()
    s = t(1.0 + 0.1 * rng.standard_normal(4)).requires_grad_()
    target = t(rng.random((N, N)))
    rho, _, _ = model.rollout(u0, v0, s, 4)
    loss = 0.5 * ((rho - target) ** 2).sum()
    loss.backward()
    du, dv, ds = (rng.standard_normal(x.shape) for x in (u0, v0, s))
    adj = float((u0.grad.numpy() * du).sum() + (v0.grad.numpy() * dv).sum()
                + (s.grad.numpy() * ds).sum())
    eps = 1e-6

    def L(sign):
        with torch.no_grad():
            r, _, _ = model.rollout(u0 + sign * eps * t(du), v0 + sign * eps * t(dv),
                                    s + sign * eps * t(ds), 4)
            return float(0.5 * ((r - target) ** 2).sum())

    fd = (L(1.0) - L(-1.0)) / (2 * eps)
    assert abs(fd - adj) <= 1e-6 * max(abs(fd), 1e-3), (fd, adj)
