"""Pins of the oracle's step composition (oracle/sim.py) and of the
tolerance stopping rule of its RB-SOR and multigrid solves, independent of
the code they check.

Step composition.  Every operator call of one OracleSmokeSim.substep is
recorded with its arguments, the arrays named by the buffer they are (the
ping-pong halves ``u0``/``u1`` etc. as they were at the start), and compared
with the sequence written out from the reference's substep
(src/cfd/common/simulation.py:186-265 and the helpers :268-321).  A
transposed step (density advected by the old velocity, v advected by the new
u, the solve given ``dt`` instead of ``self.dt``, p not zeroed, dissipation
or the metrics pass dropped or moved) changes the recorded sequence.  The
reference's own substep invariants (tests/cfd/integration/
test_pipeline_invariants.py:80-123) and initial state
(test_config_simulation_init.py:39-68) are restated on the oracle.

Tolerance rule.  The stopping test is ||b - A p|| / ||b|| over fluid cells,
b = alpha * div, checked every ``check_every`` sweeps (RB-SOR) or after
every V-cycle (multigrid), from the initial guess on.  The returned relative
residual is recomputed with a dense A assembled from the masks; the returned
count must be the first checkpoint meeting the tolerance (the previous
checkpoint, rerun as a fixed-count solve, must not meet it), or the cap.
"""
import math

import numpy as np
import pytest

import oracle as O
import oracle.sim as osim

from test_oracle_pins import adjacency, masks

SOLVE_FNS = ("jacobi_pressure_solver", "gauss_seidel_pressure_solver",
             "red_black_sor_gauss_seidel_pressure_solver", "rbsor_solve_tol",
             "multigrid_solver")
RECORDED = ("emit_scalar", "apply_forces", "advect_velocity_u", "advect_velocity_v",
            "advect_scalar_field", "cool_temperature", "compute_divergence",
            "apply_pressure_gradient", "enforce_boundary_conditions", "compute_residual",
            *SOLVE_FNS)


class Recorder:
    def __init__(self, monkeypatch, sim):
        self.calls = []
        self.names = {}
        for attr, name in (("density", "rho0"), ("density_temp", "rho1"),
                           ("velocity_u", "u0"), ("velocity_u_temp", "u1"),
                           ("velocity_v", "v0"), ("velocity_v_temp", "v1"),
                           ("temperature", "T0"), ("temperature_temp", "T1"),
                           ("pressure", "p"), ("divergence", "div"),
                           ("pressure_residual", "res"), ("fmu", "fmu"), ("fmv", "fmv"),
                           ("solid", "solid"), ("is_fluid", "fluid")):
            self.names[id(getattr(sim, attr))] = name
        self.snap = []
        for fn in RECORDED:
            monkeypatch.setattr(O, fn, self._wrap(fn, getattr(O, fn)))

    def _name(self, a):
        if isinstance(a, np.ndarray):
            return self.names.get(id(a), "?array")
        return a

    def _wrap(self, fn, real):
        def call(*args, **kw):
            self.calls.append((fn, tuple(self._name(a) for a in args)))
            # buffers as they are at the call, by name
            self.snap.append({self._name(a): a.copy() for a in args
                              if isinstance(a, np.ndarray)})
            return real(*args, **kw)
        return call


def expected_sequence(sim, dt, h, solve):
    """simulation.py:186-265 for one emitter writing density, temperature,
    u and v (emitters.py:212-220 order), with the oracle's solver call."""
    rho, g, b, amb = sim.fluid_density, sim.gravity, sim.buoyancy, sim.ambient_temperature
    e = sim.emitters[0]
    pe = (e["position"][0], e["position"][1], e["radius"], e["approach"])
    alpha = -(h * h) * rho / dt
    return [
        ("emit_scalar", ("rho0", e["density_value"], dt, *pe)),           # :206-207, :269-271
        ("emit_scalar", ("T0", e["temperature_value"], dt, *pe)),
        ("emit_scalar", ("u0", e["velocity_value"][0], dt, *pe)),
        ("emit_scalar", ("v0", e["velocity_value"][1], dt, *pe)),
        ("apply_forces", ("v0", "rho0", "T0", dt, g, b, 1.0, amb)),      # :210-211, :273-277
        ("advect_velocity_u", ("u0", "u1", "v0", "fmu", dt, h)),         # :214-215, :283
        ("advect_velocity_v", ("v0", "v1", "u0", "fmv", dt, h)),         # :284 (old u)
        ("advect_scalar_field", ("rho0", "rho1", "u1", "v1", "fluid", dt, h)),  # :218-219, :290
        ("advect_scalar_field", ("T0", "T1", "u1", "v1", "fluid", dt, h)),      # :222-223, :296
        ("cool_temperature", ("T1", amb, sim.cooling_rate, dt)),         # :225-226, :304
        ("compute_divergence", ("u1", "v1", "div", "fmu", "fmv", h)),    # :233-234, :300
        solve,                                                           # :239-240, :308-312
        ("apply_pressure_gradient", ("u1", "v1", "p", "fmu", "fmv", rho, h, dt)),  # :242-243
        ("enforce_boundary_conditions", ("u1", "v1", "solid")),          # :246-247, :320-321
        ("compute_divergence", ("u1", "v1", "div", "fmu", "fmv", h)),    # :250-251
        ("compute_residual", ("p", "div", "fluid", "fmu", "fmv", alpha, 0.25, "res")),  # :255-263
    ]


def make_sim(solver, params, H=20, W=24, h=0.5):
    sim = osim.OracleSmokeSim(H, W, h, solver=solver, solver_params=params)
    sim.emitters.append(dict(position=(9.0, 6.0), radius=4.0, density_value=1.0,
                             temperature_value=2.0, velocity_value=(0.7, 1.5),
                             approach="linear"))
    sim.ambient_temperature = 0.25
    rng = np.random.default_rng(7)
    sim.velocity_u[...] = rng.standard_normal(sim.velocity_u.shape)
    sim.velocity_v[...] = rng.standard_normal(sim.velocity_v.shape)
    sim.density[...] = rng.random(sim.density.shape)
    sim.pressure[...] = 5.0  # the step must zero it before the solve
    return sim


SOLVE_CASES = {
    "jacobi": ({"iterations": 7}, "jacobi_pressure_solver"),
    "gauss_seidel": ({"iterations": 5}, "gauss_seidel_pressure_solver"),
    "rb_sor_gauss_seidel": ({"iterations": 6, "omega": 1.3},
                            "red_black_sor_gauss_seidel_pressure_solver"),
    "rb_sor_tolerance": ({"iterations": 400, "tolerance": 1e-5, "check_every": 10},
                         "rbsor_solve_tol"),
    "multigrid": ({"v_cycles": 2, "num_levels": 3}, "multigrid_solver"),
}


@pytest.mark.parametrize("case", sorted(SOLVE_CASES))
def test_step_call_sequence(monkeypatch, case):
    params, fn = SOLVE_CASES[case]
    solver = "rb_sor_gauss_seidel" if case == "rb_sor_tolerance" else case
    sim = make_sim(solver, params)
    dt, h = 0.01, sim.cell_size
    assert dt != sim.dt  # tells the solve's self.dt from the step's dt
    rec = Recorder(monkeypatch, sim)
    sim.substep(dt)
    solve = (fn, ("p", "div", "fmu", "fmv", sim.fluid_density, h, sim.dt))
    assert rec.calls == expected_sequence(sim, dt, h, solve)
    # the swaps of simulation.py:285-286, :292, :298 leave the new fields in place
    names = {v: k for k, v in rec.names.items()}
    for attr, buf in (("density", "rho1"), ("velocity_u", "u1"), ("velocity_v", "v1"),
                      ("temperature", "T1")):
        assert id(getattr(sim, attr)) == names[buf], attr
    seq = [c[0] for c in rec.calls]
    # p = 0 when the solve starts (simulation.py:310)
    assert not rec.snap[seq.index(fn)]["p"].any()


def test_dissipation_applied_once_after_advection(monkeypatch):
    """rho' = advect(rho) * 0.999, applied between the temperature pass and
    the divergence (simulation.py:229-230), exactly once."""
    sim = make_sim("jacobi", {"iterations": 3})
    rec = Recorder(monkeypatch, sim)
    sim.substep(0.01)
    seq = [c[0] for c in rec.calls]
    k_adv = seq.index("advect_scalar_field")
    inputs = rec.snap[k_adv]
    want = np.zeros_like(inputs["rho0"])
    monkeypatch.undo()
    O.advect_scalar_field(inputs["rho0"], want, inputs["u1"], inputs["v1"], inputs["fluid"],
                          0.01, sim.cell_size)
    np.testing.assert_array_equal(sim.density, want * 0.999)


# --------------------------------------------- reference substep invariants
ALL_SOLVERS = ("jacobi", "gauss_seidel", "rb_sor_gauss_seidel", "multigrid")
DEFAULT_PARAMS = {"jacobi": {"iterations": 40}, "gauss_seidel": {"iterations": 40},
                  "rb_sor_gauss_seidel": {"iterations": 40, "omega": 0.0},
                  "multigrid": {"v_cycles": 2, "smoother_type": "jacobi", "num_levels": 4,
                                "smoother_iterations": 2, "coarsest_iterations": 30}}


@pytest.mark.parametrize("method", ALL_SOLVERS)
@pytest.mark.parametrize("iterations", [3, 5, 10])
def test_substep_zero_state_is_noop(method, iterations):
    """test_pipeline_invariants.py:78-96 on the oracle's substep: a 16x16,
    cell size 1, default-parameter simulation stays exactly zero."""
    sim = osim.OracleSmokeSim(16, 16, 1.0, solver=method, solver_params=DEFAULT_PARAMS[method])
    sim.substep(sim.dt)
    for f in ("velocity_u", "velocity_v", "pressure", "divergence"):
        assert not getattr(sim, f).any(), f


@pytest.mark.parametrize("method", ALL_SOLVERS)
def test_substep_single_impulse_reduces_divergence(monkeypatch, method):
    """test_pipeline_invariants.py:99-123 at substep level: one open face
    with u = 1; the divergence the solve sees (after advection) is larger
    than the divergence the metrics pass measures after the projection."""
    sim = osim.OracleSmokeSim(16, 16, 1.0, solver=method, solver_params=DEFAULT_PARAMS[method])
    sim.velocity_u[8, 8] = 1.0
    rec = Recorder(monkeypatch, sim)
    sim.substep(sim.dt)
    seq = [c[0] for c in rec.calls]
    solve = next(k for k, c in enumerate(seq) if c in SOLVE_FNS)
    before = np.abs(rec.snap[solve]["div"]).max()
    after = np.abs(sim.divergence).max()
    assert before > 0.0
    assert after < before, (before, after)


def test_initial_state_and_walls():
    """test_config_simulation_init.py:29-68 on the oracle: shapes, zero
    fields, the one-cell solid ring, and the masks built before the ring
    (simulation.py:102-113), so every cell is fluid to the solver."""
    sim = osim.OracleSmokeSim(10, 16, 0.5)
    assert sim.density.shape == (10, 16) and sim.pressure.shape == (10, 16)
    assert sim.temperature.shape == (10, 16)
    assert sim.velocity_u.shape == (10, 17) and sim.velocity_v.shape == (11, 16)
    for f in ("density", "velocity_u", "velocity_v", "pressure", "temperature"):
        assert not getattr(sim, f).any()
    s = sim.solid
    assert s[0, :].all() and s[-1, :].all() and s[:, 0].all() and s[:, -1].all()
    assert not s[1:-1, 1:-1].any()
    assert sim.is_fluid.all()
    assert not sim.fmu[:, 0].any() and not sim.fmu[:, -1].any() and sim.fmu[:, 1:-1].all()
    assert (sim.dt, sim.density_dissipation, sim.buoyancy, sim.gravity) == (0.016, 0.999, 5.0, -0.5)
    assert (sim.fluid_density, sim.cooling_rate, sim.ambient_temperature) == (1.0, 0.5, 0.0)


# ------------------------------------------------------- tolerance rule
def dense_rel(p, div, fmu, fmv, alpha):
    """||b - A p|| / ||b|| over fluid cells, A = 4 I - N (mirror adjacency)."""
    N = adjacency(fmu, fmv)
    fl = (fmu[:, :-1] | fmu[:, 1:] | fmv[:-1, :] | fmv[1:, :]).ravel()
    b = alpha * div.ravel()
    r = b - (4.0 * p.ravel() - N @ p.ravel())
    return math.sqrt(np.sum(r[fl] ** 2) / np.sum(b[fl] ** 2))


def poisson_case(kind, H, W, seed):
    _, fmu, fmv, _ = masks(H, W, kind, seed=seed)
    rng = np.random.default_rng(seed)
    u = rng.standard_normal((H, W + 1))
    v = rng.standard_normal((H + 1, W))
    div = np.zeros((H, W))
    O.compute_divergence(u, v, div, fmu, fmv, 0.5)
    return fmu, fmv, div


@pytest.mark.parametrize("kind,check_every,omega", [("ring", 7, 1.7), ("obstacles", 5, 0.0),
                                                    ("open", 1, 1.0)])
def test_rbsor_tolerance_rule(kind, check_every, omega):
    H, W, h, rho, dt, tol = 14, 18, 0.5, 1.0, 0.02, 1e-6
    fmu, fmv, div = poisson_case(kind, H, W, 5)
    alpha = O.calculate_alpha(h, rho, dt)
    p = np.zeros((H, W))
    n, rel = O.rbsor_solve_tol(p, div, fmu, fmv, rho, h, dt, omega, tol, check_every, 100000)
    assert n > 0 and n % check_every == 0
    assert rel <= tol
    assert rel == pytest.approx(dense_rel(p, div, fmu, fmv, alpha), rel=1e-9)
    # the same sweeps as a fixed-count solve, and the checkpoint before
    # does not meet the tolerance
    q = np.zeros((H, W))
    O.red_black_sor_gauss_seidel_pressure_solver(q, div, fmu, fmv, rho, h, dt, iterations=n,
                                                 omega=omega)
    np.testing.assert_array_equal(q, p)
    q = np.zeros((H, W))
    O.red_black_sor_gauss_seidel_pressure_solver(q, div, fmu, fmv, rho, h, dt,
                                                 iterations=n - check_every, omega=omega)
    assert dense_rel(q, div, fmu, fmv, alpha) > tol
    # the cap: a run stopped short of the tolerance, at a count that is not
    # a multiple of check_every, reports that count and its residual
    cap = n - check_every + max(1, check_every // 2) if check_every > 1 else n - 1
    q = np.zeros((H, W))
    m, rel_c = O.rbsor_solve_tol(q, div, fmu, fmv, rho, h, dt, omega, tol, check_every, cap)
    assert m == cap and rel_c > tol
    assert rel_c == pytest.approx(dense_rel(q, div, fmu, fmv, alpha), rel=1e-9)


def test_rbsor_tolerance_met_by_initial_guess():
    """Checked before any sweep: a zero right-hand side or a guess already
    within tolerance returns after 0 sweeps."""
    H, W, h, rho, dt = 12, 12, 0.5, 1.0, 0.02
    fmu, fmv, div = poisson_case("ring", H, W, 2)
    p = np.zeros((H, W))
    n, rel = O.rbsor_solve_tol(p, np.zeros((H, W)), fmu, fmv, rho, h, dt, 1.5, 1e-8, 10, 500)
    assert (n, rel) == (0, 0.0)
    n, rel = O.rbsor_solve_tol(p, div, fmu, fmv, rho, h, dt, 1.5, 1e-8, 10, 100000)
    q = p.copy()
    m, rel2 = O.rbsor_solve_tol(q, div, fmu, fmv, rho, h, dt, 1.5, 1e-8, 10, 100000)
    # the oracle's OpenMP norm reduction may combine in another order: 1 ulp
    assert m == 0 and rel2 == pytest.approx(rel, rel=1e-12) and np.array_equal(q, p)


@pytest.mark.parametrize("kind,smoother,scale", [("ring", "rb_sor_gauss_seidel", "reference"),
                                                 ("obstacles", "jacobi", "reference"),
                                                 ("open", "gauss_seidel", "reference"),
                                                 ("open", "rb_sor_gauss_seidel", "rediscretized")])
def test_multigrid_tolerance_rule(kind, smoother, scale):
    H, W, h, rho, dt, tol = 16, 24, 0.5, 1.0, 0.02, 1e-6
    fmu, fmv, div = poisson_case(kind, H, W, 8)
    alpha = O.calculate_alpha(h, rho, dt)
    kw = dict(smoother_type=smoother, num_levels=3, smoother_iterations=2,
              coarsest_iterations=40, smoother_omega=1.0, coarse_scale=scale)
    p = np.zeros((H, W))
    n, rel = O.multigrid_solver(p, div, fmu, fmv, rho, h, dt, v_cycles=2000, tolerance=tol, **kw)
    assert 1 < n < 2000 and rel <= tol
    assert rel == pytest.approx(dense_rel(p, div, fmu, fmv, alpha), rel=1e-9)
    q = np.zeros((H, W))
    O.multigrid_solver(q, div, fmu, fmv, rho, h, dt, v_cycles=n, **kw)
    np.testing.assert_array_equal(q, p)
    q = np.zeros((H, W))
    O.multigrid_solver(q, div, fmu, fmv, rho, h, dt, v_cycles=n - 1, **kw)
    assert dense_rel(q, div, fmu, fmv, alpha) > tol
    q = np.zeros((H, W))
    m, rel_c = O.multigrid_solver(q, div, fmu, fmv, rho, h, dt, v_cycles=n - 1, tolerance=tol,
                                  **kw)
    assert m == n - 1 and rel_c > tol
    assert rel_c == pytest.approx(dense_rel(q, div, fmu, fmv, alpha), rel=1e-9)
