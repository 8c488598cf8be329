"""The reference's own known-answer tests, restated, run against both the
oracle (CPU) and the HIP library (GPU, marked ``gpu``).

Sources (reference paths): tests/cfd/unit/test_advection.py,
test_divergence.py, test_pressure_gradient.py, test_pressure_solver.py,
tests/cfd/integration/test_projection.py, test_pipeline_invariants.py.
Arrays are float64 (the reference tests use float32); tolerances are 1e-12
where the reference uses 1e-6.  The reference's vacuous
``test_zero_velocity_advection`` (it compares the source with its own copy)
is restated with a real assertion.
"""
import numpy as np
import pytest

from impls import HipImpl, OracleImpl

ATOL = 1e-12


@pytest.fixture(params=["oracle", pytest.param("hip", marks=pytest.mark.gpu)])
def impl(request):
    return OracleImpl() if request.param == "oracle" else HipImpl()


def fluid_of(fmu, fmv):
    return fmu[:, :-1] | fmu[:, 1:] | fmv[:-1, :] | fmv[1:, :]


# ------------------------------------------------------------ advection
def test_zero_velocity_advection(impl, small_grid_size, face_masks):
    fmu, fmv = face_masks
    is_fluid = fluid_of(fmu, fmv)
    u = np.zeros(small_grid_size)
    v = np.zeros(small_grid_size)
    src = np.random.default_rng(0).random(small_grid_size)
    dst = np.full(small_grid_size, -7.0)
    impl.advect_scalar_field(src, dst, u, v, is_fluid, 0.1, 1.0)
    np.testing.assert_array_equal(dst[is_fluid], src[is_fluid])
    np.testing.assert_array_equal(dst[~is_fluid], 0.0)


def test_constant_velocity_advection(impl, small_grid_size, face_masks):
    fmu, fmv = face_masks
    is_fluid = fluid_of(fmu, fmv)
    src = np.full(small_grid_size, 7.0)
    u = np.full(small_grid_size, 3.0)
    v = np.full(small_grid_size, -2.0)
    dst = np.zeros(small_grid_size)
    impl.advect_scalar_field(src, dst, u, v, is_fluid, 0.1, 1.0)
    np.testing.assert_allclose(dst[1:-1, 1:-1], src[1:-1, 1:-1], atol=ATOL)


def test_uniform_rightward_velocity_advection(impl, small_grid_size, face_masks):
    H, W = small_grid_size
    fmu, fmv = face_masks
    is_fluid = fluid_of(fmu, fmv)
    u = np.ones(small_grid_size)
    v = np.zeros(small_grid_size)
    src = np.tile(np.arange(W, dtype=np.float64), (H, 1))
    dst = np.zeros(small_grid_size)
    impl.advect_scalar_field(src, dst, u, v, is_fluid, 1.0, 1.0)
    np.testing.assert_allclose(dst[1:-1, 2:-1], src[1:-1, 1:-2], atol=ATOL)


def test_uniform_upward_velocity_advection(impl, small_grid_size, face_masks):
    H, W = small_grid_size
    fmu, fmv = face_masks
    is_fluid = fluid_of(fmu, fmv)
    u = np.zeros(small_grid_size)
    v = np.ones(small_grid_size)
    src = np.tile(np.arange(H, dtype=np.float64)[:, None], (1, W))
    dst = np.zeros(small_grid_size)
    impl.advect_scalar_field(src, dst, u, v, is_fluid, 1.0, 1.0)
    np.testing.assert_allclose(dst[2:-1, 1:-1], src[1:-2, 1:-1], atol=ATOL)


def test_linear_field_bilinear_interpolation(impl, small_grid_size, face_masks):
    H, W = small_grid_size
    fmu, fmv = face_masks
    is_fluid = fluid_of(fmu, fmv)
    jj, ii = np.mgrid[0:H, 0:W].astype(np.float64)
    src = 3.0 * ii + 2.0 * jj
    u = np.full(small_grid_size, 0.25)
    v = np.full(small_grid_size, 0.5)
    dst = np.zeros(small_grid_size)
    impl.advect_scalar_field(src, dst, u, v, is_fluid, 1.0, 1.0)
    expected = 3.0 * (ii - 0.25) + 2.0 * (jj - 0.5)
    np.testing.assert_allclose(dst[1:-1, 1:-1], expected[1:-1, 1:-1], atol=ATOL)


# ----------------------------------------------------------- divergence
@pytest.mark.parametrize("velocity", [(0.0, 0.0), (1.0, 0.0), (0.0, 1.0), (3.5, -2.5)],
                         ids=["zero", "horizontal", "vertical", "arbitrary"])
def test_divergence_constant_velocity(impl, small_grid_size, face_masks, velocity, cell_size):
    H, W = small_grid_size
    u = np.full((H, W + 1), velocity[0])
    v = np.full((H + 1, W), velocity[1])
    div = np.zeros(small_grid_size)
    impl.compute_divergence(u, v, div, face_masks[0], face_masks[1], cell_size)
    np.testing.assert_allclose(div[2:-2, 2:-2], 0.0, atol=ATOL)


def _faces(H, W, fu, fv):
    ju, iu = np.mgrid[0:H, 0:W + 1].astype(np.float64)
    jv, iv = np.mgrid[0:H + 1, 0:W].astype(np.float64)
    return np.ascontiguousarray(fu(iu, ju)), np.ascontiguousarray(fv(iv, jv))


@pytest.mark.parametrize("fu,fv,expected", [
    (lambda i, j: i, lambda i, j: j, 2.0),
    (lambda i, j: 3.0 * i, lambda i, j: 0.0 * i, 3.0),
    (lambda i, j: 0.0 * i, lambda i, j: 2.0 * j, 2.0),
    (lambda i, j: -j, lambda i, j: i, 0.0),
], ids=["linear", "linear_horizontal", "linear_vertical", "rotational"])
def test_divergence_linear_fields(impl, small_grid_size, face_masks, cell_size, fu, fv, expected):
    H, W = small_grid_size
    u, v = _faces(H, W, fu, fv)
    div = np.zeros(small_grid_size)
    impl.compute_divergence(u, v, div, face_masks[0], face_masks[1], cell_size)
    np.testing.assert_allclose(div[2:-2, 2:-2], expected, atol=ATOL)


def test_divergence_quadratic_field(impl, small_grid_size, face_masks, cell_size):
    H, W = small_grid_size
    u, v = _faces(H, W, lambda i, j: i ** 2, lambda i, j: j ** 2)
    div = np.zeros(small_grid_size)
    impl.compute_divergence(u, v, div, face_masks[0], face_masks[1], cell_size)
    jj, ii = np.mgrid[0:H, 0:W]
    expected = 2 * ii + 2 * jj + 2.0
    np.testing.assert_allclose(div[2:-2, 2:-2], expected[2:-2, 2:-2], atol=ATOL)


# ------------------------------------------------------ pressure gradient
def test_constant_pressure_field(impl, small_grid_size, fluid_density, face_masks):
    H, W = small_grid_size
    p = np.full(small_grid_size, 5.0)
    u = np.zeros((H, W + 1))
    v = np.zeros((H + 1, W))
    impl.apply_pressure_gradient(u, v, p, face_masks[0], face_masks[1], fluid_density, 1.0, 1.0)
    np.testing.assert_array_equal(u, 0.0)
    np.testing.assert_array_equal(v, 0.0)


def test_linear_pressure_field(impl, small_grid_size, fluid_density, face_masks):
    H, W = small_grid_size
    jj, ii = np.mgrid[0:H, 0:W].astype(np.float64)
    p = 3.0 * ii - 2.0 * jj
    u = np.zeros((H, W + 1))
    v = np.zeros((H + 1, W))
    impl.apply_pressure_gradient(u, v, p, face_masks[0], face_masks[1], fluid_density, 1.0, 1.0)
    np.testing.assert_allclose(u[1:-1, 2:-2], -3.0 / fluid_density, atol=ATOL)
    np.testing.assert_allclose(v[2:-2, 1:-1], 2.0 / fluid_density, atol=ATOL)


def test_quadratic_pressure_field(impl, small_grid_size, fluid_density, face_masks):
    H, W = small_grid_size
    jj, ii = np.mgrid[0:H, 0:W].astype(np.float64)
    p = ii ** 2 + jj ** 2
    u = np.zeros((H, W + 1))
    v = np.zeros((H + 1, W))
    exp_u = np.zeros((H, W + 1))
    exp_v = np.zeros((H + 1, W))
    for i in range(2, W - 1):
        exp_u[:, i] = -1.0 * (2 * i - 1) / fluid_density
    for j in range(2, H - 1):
        exp_v[j, :] = -1.0 * (2 * j - 1) / fluid_density
    impl.apply_pressure_gradient(u, v, p, face_masks[0], face_masks[1], fluid_density, 1.0, 1.0)
    for col in (1, -2):
        np.testing.assert_array_equal(u[:, col], 0.0)
    for row in (1, -2):
        np.testing.assert_array_equal(v[row, :], 0.0)
    np.testing.assert_array_equal(u[0, :], 0.0)
    np.testing.assert_array_equal(u[-1, :], 0.0)
    np.testing.assert_array_equal(v[:, 0], 0.0)
    np.testing.assert_array_equal(v[:, -1], 0.0)
    np.testing.assert_allclose(u[1:-1, 2:-2], exp_u[1:-1, 2:-2], atol=ATOL)
    np.testing.assert_allclose(v[2:-2, 1:-1], exp_v[2:-2, 1:-1], atol=ATOL)


# ------------------------------------------------------------ projection
def test_divergence_of_pressure_gradient_quadratic(impl, small_grid_size, fluid_density,
                                                   face_masks):
    H, W = small_grid_size
    jj, ii = np.mgrid[0:H, 0:W].astype(np.float64)
    p = ii ** 2 + jj ** 2
    u = np.zeros((H, W + 1))
    v = np.zeros((H + 1, W))
    impl.apply_pressure_gradient(u, v, p, face_masks[0], face_masks[1], fluid_density, 1.0, 1.0)
    div = np.zeros(small_grid_size)
    impl.compute_divergence(u, v, div, face_masks[0], face_masks[1], 1.0)
    np.testing.assert_allclose(div[2:-2, 2:-2], -4.0 / fluid_density, atol=ATOL)


def test_div_grad_matches_pressure_solver_laplacian(impl, small_random_grid, fluid_density,
                                                    cell_size, face_masks):
    H, W = small_random_grid.shape
    p = small_random_grid.copy()
    u = np.zeros((H, W + 1))
    v = np.zeros((H + 1, W))
    impl.apply_pressure_gradient(u, v, p, face_masks[0], face_masks[1], fluid_density, 1.0, 1.0)
    div = np.zeros((H, W))
    impl.compute_divergence(u, v, div, face_masks[0], face_masks[1], cell_size)
    lap = np.zeros((H, W))
    lap[1:-1, 1:-1] = (p[1:-1, :-2] + p[1:-1, 2:] + p[:-2, 1:-1] + p[2:, 1:-1]
                       - 4 * p[1:-1, 1:-1]) / cell_size ** 2
    np.testing.assert_allclose(div[2:-2, 2:-2], -(1.0 / fluid_density) * lap[2:-2, 2:-2],
                               atol=ATOL)


# -------------------------------------------------------- pressure solver
def _restated_residual(p, div, fmu, fmv, alpha):
    """tests/cfd/unit/test_pressure_solver.py:29-43 of the reference."""
    H, W = p.shape
    is_fluid = fluid_of(fmu, fmv)
    r = np.zeros_like(p)
    for j in range(1, H - 1):
        for i in range(1, W - 1):
            if not is_fluid[j, i]:
                continue
            left = p[j, i - 1] if fmu[j, i] else p[j, i]
            right = p[j, i + 1] if fmu[j, i + 1] else p[j, i]
            down = p[j - 1, i] if fmv[j, i] else p[j, i]
            up = p[j + 1, i] if fmv[j + 1, i] else p[j, i]
            r[j, i] = left + right + down + up - 4 * p[j, i] + alpha * div[j, i]
    return np.linalg.norm(r)


def test_pressure_solver_reduces_divergence(impl, small_grid_size, face_masks, fluid_density):
    H, W = small_grid_size
    fmu, fmv = face_masks
    div = np.zeros((H, W))
    div[H // 2, W // 2] = 1.0
    div[H // 2, W // 2 + 1] = -1.0
    alpha = (-(1.0 * 1.0) * fluid_density) / 1.0
    p = np.zeros((H, W))
    r0 = _restated_residual(p, div, fmu, fmv, alpha)
    assert r0 > 1e-6
    impl.solve("jacobi", p, div, fmu, fmv, fluid_density, 1.0, 1.0, iterations=50)
    r1 = _restated_residual(p, div, fmu, fmv, alpha)
    assert r1 < r0 * 0.1


@pytest.mark.parametrize("method", ["jacobi", "gauss_seidel", "rb_sor_gauss_seidel", "multigrid"])
def test_pressure_solver_zero_divergence_stays_zero(impl, small_grid_size, face_masks,
                                                    fluid_density, method):
    fmu, fmv = face_masks
    div = np.zeros(small_grid_size)
    p = np.zeros(small_grid_size)
    kw = {"iterations": 50} if method != "multigrid" else {"v_cycles": 2}
    impl.solve(method, p, div, fmu, fmv, fluid_density, 1.0, 1.0, **kw)
    np.testing.assert_array_equal(p, 0.0)


# ------------------------------------------------- pipeline invariants
@pytest.mark.parametrize("method", ["jacobi", "gauss_seidel", "rb_sor_gauss_seidel", "multigrid"])
def test_zero_velocity_zero_pressure_is_noop(impl, domain, method):
    H, W = domain["height"], domain["width"]
    fmu, fmv = domain["face_mask_u"], domain["face_mask_v"]
    u = np.zeros((H, W + 1))
    v = np.zeros((H + 1, W))
    p = np.zeros((H, W))
    div = np.zeros((H, W))
    impl.compute_divergence(u, v, div, fmu, fmv, 1.0)
    np.testing.assert_array_equal(div, 0.0)
    kw = {"iterations": 40} if method != "multigrid" else {"v_cycles": 2}
    impl.solve(method, p, div, fmu, fmv, 1000.0, 1.0, 1.0, **kw)
    np.testing.assert_array_equal(p, 0.0)
    impl.apply_pressure_gradient(u, v, p, fmu, fmv, 1000.0, 1.0, 1.0)
    np.testing.assert_array_equal(u, 0.0)
    np.testing.assert_array_equal(v, 0.0)
    impl.compute_divergence(u, v, div, fmu, fmv, 1.0)
    np.testing.assert_array_equal(div, 0.0)


@pytest.mark.parametrize("method,kw", [
    ("jacobi", {"iterations": 40}),
    ("gauss_seidel", {"iterations": 40}),
    ("rb_sor_gauss_seidel", {"iterations": 40}),
    ("multigrid", {"v_cycles": 2, "smoother_type": "jacobi", "num_levels": 4,
                   "smoother_iterations": 2, "coarsest_iterations": 30}),
])
def test_single_impulse_reduces_divergence(impl, method, kw):
    """test_pipeline_invariants.py:98-123 on the default simulation's masks
    (outer walls only, simulation.py:102-108), 16x16, cell_size 1, dt 0.016."""
    H = W = 16
    fmu = np.zeros((H, W + 1), dtype=bool)
    fmv = np.zeros((H + 1, W), dtype=bool)
    fmu[:, 1:-1] = True
    fmv[1:-1, :] = True
    u = np.zeros((H, W + 1))
    v = np.zeros((H + 1, W))
    u[8, 8] = 1.0
    div = np.zeros((H, W))
    impl.compute_divergence(u, v, div, fmu, fmv, 1.0)
    before = np.abs(div).max()
    assert before > 0
    p = np.zeros((H, W))
    impl.solve(method, p, div, fmu, fmv, 1.0, 1.0, 0.016, **kw)
    impl.apply_pressure_gradient(u, v, p, fmu, fmv, 1.0, 1.0, 0.016)
    impl.compute_divergence(u, v, div, fmu, fmv, 1.0)
    assert np.abs(div).max() < before
