"""The mixed solver's inner-solve options.  The coarsest level as a
precomputed operator (MultigridParams.coarsest_solve = "operator"): the
red-black sweeps of the coarsest level from a zero guess are linear in its
rhs, so the FP32 V-cycles may apply them as a 256x256 matrix built once per
handle.  The results agree with the sweeps to float32 rounding, tolerance
solves converge alike, and a coarsest level above 16x16 keeps the sweeps.
float32 arithmetic in the FP32 V-cycles (inner_arithmetic = "float32")
reaches the same FP64 tolerance in the same refinements."""
import numpy as np
import pytest
import torch

import oracle as O

pytestmark = pytest.mark.gpu


def problem(H, W, seed=5, block=True):
    """A solid ring and block (scattered solid cells would make the coarsened
    masks -- a coarse cell is solid if any child is -- all solid by 16x16),
    or no solid cell at all (block=False: the domain's edge faces closed)."""
    rng = np.random.default_rng(seed)
    solid = np.zeros((H, W), dtype=bool)
    if block:
        solid[0, :] = solid[-1, :] = True
        solid[:, 0] = solid[:, -1] = True
        solid[H // 4:H // 2, W // 3:W // 2] = True
    fmu = np.zeros((H, W + 1), dtype=bool)
    fmv = np.zeros((H + 1, W), dtype=bool)
    fmu[:, 1:-1] = ~(solid[:, :-1] | solid[:, 1:])
    fmv[1:-1, :] = ~(solid[:-1, :] | solid[1:, :])
    u = rng.standard_normal((H, W + 1))
    v = rng.standard_normal((H + 1, W))
    d = np.zeros((H, W))
    O.compute_divergence(u, v, d, fmu, fmv, 1.0 / H)
    return d, fmu, fmv


MG = dict(smoother_type="rb_sor_gauss_seidel", smoother_omega=1.0, smoother_iterations=2,
          coarsest_iterations=50, coarse_scale="rediscretized", precision="mixed")


def up_to_constant(a, b):
    """max |a - b| after removing their mean difference: with every edge face
    closed the pressure is defined up to a constant, which inner solves of
    different rounding leave differently."""
    d = a - b
    return np.abs(d - d.mean()).max()


def solve(H, W, solver=None, block=True, dt=0.005, **kw):
    from cfd_amd.backends.hip import solvers

    d, fmu, fmv = problem(H, W, block=block)
    out = torch.zeros((H, W), dtype=torch.float64, device="cuda")
    solver = solver or solvers.MultigridSolver()
    solver(out, torch.from_numpy(d).cuda(), torch.from_numpy(fmu).cuda(),
           torch.from_numpy(fmv).cuda(), 1.0, 1.0 / H, dt, **dict(MG, **kw))
    torch.cuda.synchronize()
    return out.cpu().numpy(), dict(solvers.LAST_SOLVE_INFO)


@pytest.mark.parametrize("shape", [(256, 256, 5), (128, 64, 4), (64, 128, 4)],
                         ids=lambda s: "%dx%d" % s[:2])
def test_operator_matches_sweeps(shape):
    h, w, levels = shape
    kw = dict(num_levels=levels, v_cycles=4, inner_cycles=1)
    ref, _ = solve(h, w, coarsest_solve="sweeps", **kw)
    got, _ = solve(h, w, coarsest_solve="operator", **kw)
    scale = np.abs(ref).max()
    assert np.abs(got - ref).max() <= 1e-5 * scale
    assert not np.array_equal(got, ref)  # the operator path did run


def test_operator_tolerance_solve():
    kw = dict(num_levels=5, v_cycles=60, inner_cycles=2, tolerance=1e-8, block=False)
    ref, iref = solve(256, 256, coarsest_solve="sweeps", **kw)
    got, info = solve(256, 256, coarsest_solve="operator", **kw)
    assert iref["converged"] and info["converged"], (iref, info)
    assert info["rel_residual"] <= 1e-8
    assert abs(info["refinements"] - iref["refinements"]) <= 1
    assert up_to_constant(got, ref) <= 1e-6 * np.abs(ref).max()


def test_operator_rebuilt_for_new_sweep_count():
    """One handle, two coarsest sweep counts: the cached operator follows."""
    from cfd_amd.backends.hip import solvers

    mg = solvers.MultigridSolver()
    kw = dict(num_levels=5, v_cycles=3, inner_cycles=1)
    a50, _ = solve(256, 256, mg, coarsest_solve="operator", **kw)
    a20, _ = solve(256, 256, mg, coarsest_solve="operator", coarsest_iterations=20, **kw)
    r20, _ = solve(256, 256, coarsest_solve="sweeps", coarsest_iterations=20, **kw)
    assert np.abs(a20 - r20).max() <= 1e-5 * np.abs(r20).max()
    assert np.abs(a20 - a50).max() > 1e-5 * np.abs(a50).max()


def test_large_coarsest_keeps_the_sweeps():
    """A 25x17 coarsest level (200x136, 4 levels) has more than 256 cells:
    the operator mode falls back to the sweeps, bit for bit."""
    kw = dict(num_levels=5, v_cycles=3, inner_cycles=1)
    ref, _ = solve(200, 136, coarsest_solve="sweeps", **kw)
    got, _ = solve(200, 136, coarsest_solve="operator", **kw)
    np.testing.assert_array_equal(got, ref)


def test_bad_mode_is_rejected():
    with pytest.raises(ValueError, match="coarsest_solve must be one of"):
        solve(64, 64, coarsest_solve="direct", num_levels=3, v_cycles=1)


def test_float32_inner_arithmetic_reaches_the_tolerance():
    kw = dict(num_levels=5, v_cycles=60, inner_cycles=2, tolerance=1e-8, block=False)
    ref, iref = solve(256, 256, **kw)
    for cs in ("sweeps", "operator"):
        got, info = solve(256, 256, inner_arithmetic="float32", coarsest_solve=cs, **kw)
        assert info["converged"] and info["rel_residual"] <= 1e-8, info
        assert abs(info["refinements"] - iref["refinements"]) <= 1
        assert up_to_constant(got, ref) <= 1e-6 * np.abs(ref).max()


def test_float32_inner_arithmetic_differs_from_float64():
    """The option reaches the kernels: a fixed-count solve changes (slightly)."""
    kw = dict(num_levels=5, v_cycles=3, inner_cycles=1)
    a, _ = solve(256, 256, **kw)
    b, _ = solve(256, 256, inner_arithmetic="float32", **kw)
    assert not np.array_equal(a, b)
    assert np.abs(a - b).max() <= 1e-4 * np.abs(a).max()
    with pytest.raises(ValueError, match="inner_arithmetic must be one of"):
        solve(64, 64, inner_arithmetic="float16", num_levels=3, v_cycles=1)


# coarsest_solve = "exact": the FP32 V-cycles apply the coarsest level's exact
# least-squares operator (its host construction is checked against the
# oracle's sweeps in test_coarse_exact.py)

def test_exact_coarsest_converges_in_fewer_refinements():
    """At the same tolerance the exact coarsest level takes fewer refinements
    than the sweeps' operator, and its result is closer to the converged
    solution (an FP64 solve to 1e-13) than the FP64 V-cycles' own result at
    that tolerance: the error left at 1e-8 is the smooth component the 50
    coarsest sweeps leave behind (measured 1.0e-6 against 2.1e-5 relative)."""
    kw = dict(num_levels=5, v_cycles=200, inner_cycles=2, block=False)
    tight, it = solve(256, 256, tolerance=1e-13, **dict(kw, precision="native"))
    assert it["converged"]
    res = {}
    for name, extra in (("native", dict(precision="native")),
                        ("operator", dict(coarsest_solve="operator")),
                        ("exact", dict(coarsest_solve="exact"))):
        got, info = solve(256, 256, tolerance=1e-8, **dict(kw, **extra))
        assert info["converged"] and info["rel_residual"] <= 1e-8, (name, info)
        res[name] = (up_to_constant(got, tight), info)
    assert res["exact"][1]["refinements"] < res["operator"][1]["refinements"], res
    assert res["exact"][0] <= 0.25 * res["native"][0], res


def test_exact_single_level_is_a_direct_solve():
    """A 16x16 grid of one level: the V-cycle is the operator itself, so each
    FP64 refinement gains float32's accuracy (~1e-6 or better)."""
    got, info = solve(16, 16, coarsest_solve="exact", num_levels=1, v_cycles=20,
                      inner_cycles=1, tolerance=1e-11)
    assert info["converged"] and info["rel_residual"] <= 1e-11, info
    assert info["refinements"] <= 4, info
    ref, iref = solve(16, 16, coarsest_solve="sweeps", num_levels=1, v_cycles=20,
                      inner_cycles=1, tolerance=1e-11)
    assert iref["refinements"] > info["refinements"], (iref, info)


def test_exact_operator_cache_follows_the_mode():
    """One handle alternating operator kinds gives each kind's fresh-handle
    result bit for bit."""
    from cfd_amd.backends.hip import solvers

    mg = solvers.MultigridSolver()
    kw = dict(num_levels=5, v_cycles=3, inner_cycles=1)
    fresh = {cs: solve(256, 256, coarsest_solve=cs, **kw)[0] for cs in ("exact", "operator")}
    for cs in ("exact", "operator", "exact"):
        got, _ = solve(256, 256, mg, coarsest_solve=cs, **kw)
        np.testing.assert_array_equal(got, fresh[cs])
    assert not np.array_equal(fresh["exact"], fresh["operator"])


# coarsest_solve = "exact_1024": the FP32 V-cycles stop at the first level of
# at most 1024 cells and apply its exact operator there (k_apply_exact, many
# workgroups); the levels below are not visited

def test_exact_1024_converges_like_exact():
    """256x256, 5 levels: the cycles stop at 32x32.  At the same tolerance the
    solve converges in no more refinements than the 16x16 exact level, to a
    result as close to the tightly converged solution."""
    kw = dict(num_levels=5, v_cycles=200, inner_cycles=2, block=False)
    tight, it = solve(256, 256, tolerance=1e-13, **dict(kw, precision="native"))
    assert it["converged"]
    res = {}
    for cs in ("exact", "exact_1024"):
        got, info = solve(256, 256, tolerance=1e-8, coarsest_solve=cs, **kw)
        assert info["converged"] and info["rel_residual"] <= 1e-8, (cs, info)
        res[cs] = (up_to_constant(got, tight), info)
    assert res["exact_1024"][1]["refinements"] <= res["exact"][1]["refinements"], res
    assert res["exact_1024"][0] <= 2.0 * res["exact"][0] + 1e-9 * np.abs(tight).max(), res


@pytest.mark.parametrize("shape", [(200, 136, 4), (192, 192, 4)], ids=["200x136", "192x192"])
def test_exact_1024_ragged_level(shape):
    """200x136 and 192x192 in 4 levels: the first level of at most 1024 cells
    is the coarsest, 25x17 (425 cells, not a multiple of 4: the product's
    row-by-row path) or 24x24 (576).  Solving it exactly instead of by its 50
    sweeps converges in fewer refinements (measured 4 and 3 against 12), to a
    result at least as close to the tightly converged solution as the FP64
    V-cycles' own at the same tolerance (whose error is the smooth component
    the coarsest sweeps leave behind)."""
    h, w, levels = shape
    kw = dict(num_levels=levels, v_cycles=400, inner_cycles=2, block=False)
    tight, it = solve(h, w, tolerance=1e-13, **dict(kw, precision="native"))
    ref, iref = solve(h, w, tolerance=1e-8, **dict(kw, precision="native"))
    swp, isw = solve(h, w, tolerance=1e-8, coarsest_solve="sweeps", **kw)
    got, info = solve(h, w, tolerance=1e-8, coarsest_solve="exact_1024", **kw)
    assert it["converged"] and iref["converged"], (it, iref)
    assert info["converged"] and info["rel_residual"] <= 1e-8, info
    assert info["refinements"] < isw["refinements"], (info, isw)
    scale = np.abs(tight).max()
    assert up_to_constant(got, tight) <= max(up_to_constant(ref, tight), 1e-7 * scale)


def test_exact_1024_single_level_is_a_direct_solve():
    """A 32x32 grid of one level: the V-cycle is the 1024x1024 operator itself,
    so each FP64 refinement gains float32's accuracy."""
    got, info = solve(32, 32, coarsest_solve="exact_1024", num_levels=1, v_cycles=20,
                      inner_cycles=1, tolerance=1e-11)
    assert info["converged"] and info["rel_residual"] <= 1e-11, info
    assert info["refinements"] <= 4, info


def test_exact_1024_is_red_black_only():
    """With the Jacobi smoother the option keeps the coarsest sweeps, bit for
    bit."""
    kw = dict(num_levels=5, v_cycles=3, inner_cycles=1, smoother_type="jacobi")
    ref, _ = solve(256, 256, coarsest_solve="sweeps", **kw)
    got, _ = solve(256, 256, coarsest_solve="exact_1024", **kw)
    np.testing.assert_array_equal(got, ref)


def test_exact_1024_cache_follows_mode_and_masks():
    """One handle alternating the two exact kinds and two mask sets (every
    level derived again from the masks passed in, rebuild_masks) gives each
    case's fresh-handle result bit for bit."""
    from cfd_amd.backends.hip import solvers

    mg = solvers.MultigridSolver()
    kw = dict(num_levels=5, v_cycles=3, inner_cycles=1, rebuild_masks=True)
    cases = [("exact_1024", True), ("exact", True), ("exact_1024", False), ("exact_1024", True)]
    fresh = {c: solve(256, 256, coarsest_solve=c[0], block=c[1], **kw)[0] for c in set(cases)}
    for cs, block in cases:
        got, _ = solve(256, 256, mg, coarsest_solve=cs, block=block, **kw)
        np.testing.assert_array_equal(got, fresh[(cs, block)])
    assert not np.array_equal(fresh[("exact_1024", True)], fresh[("exact", True)])


def test_exact_1024_operator_serves_every_dt():
    """The 1024-cell operator is built without alpha (the product applies the
    level's alpha), so one handle solving at two time steps gives each time
    step's fresh-handle result bit for bit, and a tolerance solve at the
    second time step still converges."""
    from cfd_amd.backends.hip import solvers

    mg = solvers.MultigridSolver()
    kw = dict(num_levels=5, v_cycles=3, inner_cycles=1, coarsest_solve="exact_1024", block=False)
    fresh = {dt: solve(256, 256, dt=dt, **kw)[0] for dt in (0.005, 0.02)}
    for dt in (0.005, 0.02, 0.005):
        got, _ = solve(256, 256, mg, dt=dt, **kw)
        np.testing.assert_array_equal(got, fresh[dt])
    assert not np.array_equal(fresh[0.005], fresh[0.02])
    got, info = solve(256, 256, mg, dt=0.02, **dict(kw, v_cycles=60, inner_cycles=2,
                                                      tolerance=1e-8))
    assert info["converged"] and info["rel_residual"] <= 1e-8, info
