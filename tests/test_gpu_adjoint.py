"""HIP adjoint kernels and differentiable rollouts vs the FP64 tape
(oracle/tape.py: torch autograd of the same step on the CPU) and vs central
finite differences.

Per-operator vector-Jacobian products are compared at generic points (random
velocities, so no backtrace lies on a grid line) to 1e-12 relative; whole
rollouts to 1e-6 relative (the bar for gradients), limited by the multigrid
tolerance."""
import numpy as np
import pytest
import torch

from oracle import tape as T

pytestmark = pytest.mark.gpu

H, W = 24, 40
HS, DT = 1.0 / 32, 0.02


def cu(a, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype).cuda()


def cpu(a):
    return torch.from_numpy(np.ascontiguousarray(a))


def rel(a, b):
    a = a.detach().cpu().double().numpy() if isinstance(a, torch.Tensor) else a
    b = b.detach().cpu().double().numpy() if isinstance(b, torch.Tensor) else b
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def masks(h=H, w=W):
    fmu = np.zeros((h, w + 1), dtype=bool)
    fmv = np.zeros((h + 1, w), dtype=bool)
    fmu[:, 1:-1] = True
    fmv[1:-1, :] = True
    return fmu, fmv


@pytest.fixture(scope="module")
def ops():
    from cfd_amd.backends.hip import adjoint as A
    from cfd_amd.backends.hip import operators as P
    return A, P


def test_box_set_and_buoyancy_forward_equal_tape(ops):
    A, _ = ops
    rng = np.random.default_rng(0)
    rho = rng.random((H, W))
    v = rng.standard_normal((H + 1, W))
    s = np.array([0.7, 1.3, 2.1])
    box = (5, 3, 13, 9)
    r = cu(rho)
    A.box_set_indexed(r, box, cu(s), 1)
    exp = T.box_set(cpu(rho), box, cpu(s)[1])
    np.testing.assert_array_equal(r.cpu().numpy(), exp.numpy())
    vv = cu(v)
    A.apply_buoyancy(vv, r, DT, -1.0, 1.0)
    np.testing.assert_array_equal(vv.cpu().numpy(),
                                  T.buoyancy(cpu(v), exp, DT, -1.0, 1.0).numpy())


@pytest.mark.parametrize("mode", ["gather", "scatter"])
@pytest.mark.parametrize("cfl", [0.4, 3.0])
def test_scalar_advection_adjoint_matches_tape(ops, mode, cfl):
    A, _ = ops
    rng = np.random.default_rng(1)
    src = rng.random((H, W))
    u = rng.standard_normal((H, W + 1)) * cfl * HS / DT
    v = rng.standard_normal((H + 1, W)) * cfl * HS / DT
    fl = np.ones((H, W), dtype=bool)
    fl[3, 4] = fl[10, 30] = False
    ybar = rng.standard_normal((H, W))
    ts, tu, tv = (cpu(x).requires_grad_() for x in (src, u, v))
    out = T.advect_scalar(ts, tu, tv, cpu(fl), DT, HS)
    out.backward(cpu(ybar))
    sb = torch.empty((H, W), dtype=torch.float64, device="cuda")
    ub = torch.zeros((H, W + 1), dtype=torch.float64, device="cuda")
    vb = torch.zeros((H + 1, W), dtype=torch.float64, device="cuda")
    A.advect_scalar_adjoint(cu(ybar), cu(src), cu(u), cu(v), cu(fl, torch.bool), DT, HS, sb,
                            ub, vb, mode=mode)
    assert rel(sb, ts.grad) <= 1e-13
    assert rel(ub, tu.grad) <= 1e-12
    assert rel(vb, tv.grad) <= 1e-12


@pytest.mark.parametrize("mode", ["gather", "scatter"])
@pytest.mark.parametrize("cfl", [0.4, 3.0])
def test_velocity_advection_adjoint_matches_tape(ops, mode, cfl):
    A, _ = ops
    rng = np.random.default_rng(2)
    fmu, fmv = masks()
    fmu[6, 9] = False
    fmv[7, 11] = False
    u = rng.standard_normal((H, W + 1)) * cfl * HS / DT
    v = rng.standard_normal((H + 1, W)) * cfl * HS / DT
    ubar = rng.standard_normal((H, W + 1))
    vbar = rng.standard_normal((H + 1, W))
    tu, tv = cpu(u).requires_grad_(), cpu(v).requires_grad_()
    ou = T.advect_u(tu, tv, cpu(fmu), DT, HS)
    ov = T.advect_v(tv, tu, cpu(fmv), DT, HS)
    ((ou * cpu(ubar)).sum() + (ov * cpu(vbar)).sum()).backward()
    us_bar = torch.empty((H, W + 1), dtype=torch.float64, device="cuda")
    vs_bar = torch.empty((H + 1, W), dtype=torch.float64, device="cuda")
    A.advect_velocity_adjoint(cu(ubar), cu(vbar), cu(u), cu(v), cu(fmu, torch.bool),
                              cu(fmv, torch.bool), DT, HS, us_bar, vs_bar, mode=mode)
    assert rel(us_bar, tu.grad) <= 1e-12
    assert rel(vs_bar, tv.grad) <= 1e-12


def test_buoyancy_and_box_adjoints_match_tape(ops):
    A, _ = ops
    rng = np.random.default_rng(3)
    rho = rng.random((H, W))
    v = rng.standard_normal((H + 1, W))
    s = np.array([0.5, 1.5])
    box = (2, 1, 9, 6)
    vbar = rng.standard_normal((H + 1, W))
    rbar = rng.standard_normal((H, W))
    tr, tv, ts = cpu(rho).requires_grad_(), cpu(v).requires_grad_(), cpu(s).requires_grad_()
    r0 = T.box_set(tr, box, ts[1])
    v1 = T.buoyancy(tv, r0, DT, -1.0, 1.0)
    ((v1 * cpu(vbar)).sum() + (r0 * cpu(rbar)).sum()).backward()
    rb = cu(rbar)
    A.apply_buoyancy_adjoint(cu(vbar), rb, DT, -1.0, 1.0)
    sb = torch.zeros(2, dtype=torch.float64, device="cuda")
    A.box_set_adjoint(rb, box, sb, 1)
    assert rel(rb, tr.grad) <= 1e-14
    assert rel(sb, ts.grad) <= 1e-13


def _model(N, tol=1e-12, **kw):
    from cfd_amd.diff import DifferentiableSmoke
    h = 1.0 / N
    box = (N // 2 - N // 8, N // 8, N // 2 + N // 8, N // 8 + N // 8)
    return DifferentiableSmoke(N, N, h, 0.01, box, tolerance=tol, **kw), box, h


def test_projection_adjoint_dot_product():
    m, _, _ = _model(48)
    rng = np.random.default_rng(4)
    a = [cu(rng.standard_normal((48, 49))), cu(rng.standard_normal((49, 48)))]
    b = [cu(rng.standard_normal((48, 49))), cu(rng.standard_normal((49, 48)))]
    pa = [x.clone() for x in a]
    m.project_(*pa)
    ptb = [x.clone() for x in b]
    m.project_adjoint_(*ptb)
    lhs = float((pa[0] * b[0]).sum() + (pa[1] * b[1]).sum())
    rhs = float((a[0] * ptb[0]).sum() + (a[1] * ptb[1]).sum())
    assert abs(lhs - rhs) <= 1e-9 * abs(lhs)


def test_step_vjp_matches_tape():
    N = 32
    m, box, h = _model(N)
    tp = T.TapeSmoke(N, N, h, 0.01, box)
    rng = np.random.default_rng(5)
    rho = rng.random((N, N))
    u = rng.standard_normal((N, N + 1)) * 0.3
    v = rng.standard_normal((N + 1, N)) * 0.3
    s = np.array([1.0, 1.2])
    bars = [rng.standard_normal(x.shape) for x in (rho, u, v)]
    ins = [cpu(x).requires_grad_() for x in (rho, u, v, s)]
    outs = tp.step(ins[0], ins[1], ins[2], ins[3][1])
    sum((o * cpu(b)).sum() for o, b in zip(outs, bars)).backward()
    r2, u3, v3, saved = m.step_forward(cu(rho), cu(u), cu(v), cu(s), 1)
    for got, exp in zip((r2, u3, v3), outs):
        assert rel(got, exp) <= 1e-10
    sb = torch.zeros(2, dtype=torch.float64, device="cuda")
    rb, ub, vb = m.step_backward(saved, *(cu(b) for b in bars), sb, 1)
    assert rel(rb, ins[0].grad) <= 1e-9
    assert rel(ub, ins[1].grad) <= 1e-9
    assert rel(vb, ins[2].grad) <= 1e-9
    assert rel(sb, ins[3].grad) <= 1e-9


@pytest.mark.parametrize("mode", ["gather", "scatter"])
def test_rollout_gradient_matches_tape(mode):
    N, steps = 32, 8
    m, box, h = _model(N, adjoint_mode=mode)
    tp = T.TapeSmoke(N, N, h, 0.01, box)
    rng = np.random.default_rng(6)
    u0 = rng.standard_normal((N, N + 1)) * 0.2
    v0 = rng.standard_normal((N + 1, N)) * 0.2
    s = 1.0 + 0.1 * rng.standard_normal(steps)
    target = rng.random((N, N))
    tin = [cpu(x).requires_grad_() for x in (u0, v0, s)]
    rho, _, _ = tp.rollout(tin[0], tin[1], tin[2], steps)
    (0.5 * ((rho - cpu(target)) ** 2).sum()).backward()
    gin = [cu(x).requires_grad_() for x in (u0, v0, s)]
    rho_g, _, _ = m.rollout(gin[0], gin[1], gin[2], steps, checkpoint_every=3)
    (0.5 * ((rho_g - cu(target)) ** 2).sum()).backward()
    for g, e in zip(gin, tin):
        assert rel(g.grad, e.grad) <= 1e-6


def test_gather_adjoint_is_deterministic_and_matches_scatter():
    N, steps = 48, 6
    grads = {}
    for mode in ("gather", "scatter", "gather"):
        m, box, h = _model(N, adjoint_mode=mode, warm_start_adjoint=False)
        rng = np.random.default_rng(7)
        u0 = cu(rng.standard_normal((N, N + 1)) * 0.5).requires_grad_()
        v0 = cu(rng.standard_normal((N + 1, N)) * 0.5).requires_grad_()
        s = cu(np.ones(steps)).requires_grad_()
        rho, _, _ = m.rollout(u0, v0, s, steps, checkpoint_every=2)
        (0.5 * (rho * rho).sum()).backward()
        g = (u0.grad.clone(), v0.grad.clone(), s.grad.clone())
        if mode in grads:
            for a, b in zip(g, grads[mode]):
                assert torch.equal(a, b), "gather adjoint not bit-reproducible"
        grads[mode] = g
    for a, b in zip(grads["gather"], grads["scatter"]):
        assert rel(a, b) <= 1e-11


def test_kink_slope_is_mean_of_one_sided_slopes(ops):
    """Zero velocity puts every backtrace on a grid line; the adjoint's
    velocity sensitivity must then equal the central difference."""
    A, P = ops
    rng = np.random.default_rng(8)
    src = cu(rng.random((H, W)))
    u = torch.zeros((H, W + 1), dtype=torch.float64, device="cuda")
    v = torch.zeros((H + 1, W), dtype=torch.float64, device="cuda")
    fl = torch.ones((H, W), dtype=torch.bool, device="cuda")
    ybar = cu(rng.standard_normal((H, W)))
    sb = torch.empty_like(src)
    ub, vb = torch.zeros_like(u), torch.zeros_like(v)
    A.advect_scalar_adjoint(ybar, src, u, v, fl, DT, HS, sb, ub, vb)
    du = cu(rng.standard_normal((H, W + 1)))
    eps = 1e-7

    def f(sign):
        out = torch.empty_like(src)
        P.advect_scalar_field(src, out, u + sign * eps * du, v, fl, DT, HS)
        return float((out * ybar).sum())

    fd = (f(1.0) - f(-1.0)) / (2 * eps)
    adj_dir = float((ub * du).sum())
    assert abs(fd - adj_dir) <= 1e-6 * abs(fd)


def test_plume_256_gradient_central_differences():
    """The 256x256, 64-step rollout from zero velocity and unit strengths:
    adjoint directional derivatives vs central differences (32 seeded
    directions, FP64, converged solves) within 1e-6 relative.

    At zero velocity every backtrace lands on a grid line, where the
    bilinear interpolation has a kink: the adjoint takes the mean of the
    one-sided slopes, and a central difference approaches that mean with an
    error that shrinks with eps (measured over the same directions, max /
    median: 3.2e-4 / 5.7e-6 at eps 1e-5, 3.0e-6 / 2.2e-7 at 1e-6, 3.2e-7 /
    2.5e-8 at 1e-7, rounding still far below).  Hence eps = 1e-7, and the
    error must also fall from eps = 1e-6 to 1e-7 (the difference converging
    to the adjoint, not to another value)."""
    from cfd_amd import diff, scenes
    sc = scenes.diff_rollout_256()
    m = diff.from_scene(sc, tolerance=0.0, v_cycles=30)
    N, steps = sc.H, sc.steps
    target = cu(diff.gaussian_target(N, N, sc.cell_size))
    u0 = torch.zeros((N, N + 1), dtype=torch.float64, device="cuda", requires_grad=True)
    v0 = torch.zeros((N + 1, N), dtype=torch.float64, device="cuda", requires_grad=True)
    s = torch.ones(steps, dtype=torch.float64, device="cuda", requires_grad=True)
    rho, _, _ = m.rollout(u0, v0, s, steps, checkpoint_every=8)
    diff.density_loss(rho, target).backward()
    rng = np.random.default_rng(50)
    worst = {1e-6: 0.0, 1e-7: 0.0}
    for _ in range(32):
        du = cu(rng.standard_normal((N, N + 1)))
        dv = cu(rng.standard_normal((N + 1, N)))
        ds = cu(rng.standard_normal(steps))
        adj = float((u0.grad * du).sum() + (v0.grad * dv).sum() + (s.grad * ds).sum())
        for eps in worst:
            with torch.no_grad():
                fd = diff.directional_fd(m, u0.detach(), v0.detach(), s.detach(), steps,
                                         target, du, dv, ds, eps)
            worst[eps] = max(worst[eps], abs(fd - adj) / abs(fd))
    print("gradient vs central differences, max relative error by eps:", worst)
    assert worst[1e-7] <= 1e-6, worst
    assert worst[1e-7] < 0.5 * worst[1e-6], worst


def test_fp32_gradient_tracks_fp64():
    N, steps = 64, 8
    grads = []
    for dt_ in (torch.float64, torch.float32):
        m, box, h = _model(N, tol=1e-8, dtype=dt_)
        rng = np.random.default_rng(9)
        u0 = cu(rng.standard_normal((N, N + 1)) * 0.3, dt_).requires_grad_()
        v0 = cu(rng.standard_normal((N + 1, N)) * 0.3, dt_).requires_grad_()
        s = cu(np.ones(steps), dt_).requires_grad_()
        tgt = cu(rng.random((N, N)), dt_)
        rho, _, _ = m.rollout(u0, v0, s, steps, checkpoint_every=4)
        (0.5 * ((rho - tgt) ** 2).sum()).backward()
        grads.append((u0.grad.double(), v0.grad.double(), s.grad.double()))
    for a, b in zip(grads[1], grads[0]):
        assert rel(a, b) <= 1e-3
