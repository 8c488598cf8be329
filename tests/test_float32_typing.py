"""Float32 advection under Numba's typing rules, bit for bit.

The reference runs float32 arrays through kernels whose coordinates and
weights are Python floats (float64): in ``lerp(a, b, t) = a + (b - a) * t``
(advection.py:16-19) with a, b float32 array elements and t float64, b - a
is a float32 subtraction and everything after it is float64; the second
lerp of sample_scalar_field (advection.py:48) is float64 throughout; the
backtrace ``u * dt * inv_cell_size`` (advection.py:124-126, :148-150,
:174-176) is float64; the sampled value is rounded to float32 when stored.

``advect`` below restates the three advection operators (advection.py:
112-180) in vectorised numpy with a dtype parameter, applying exactly those
promotions.  In float64 it must equal the C oracle bit for bit (pinned by
the reference's known-answer tests), which checks its structure: offsets,
clamping, masks, operand order.  In float32 it is the typing oracle the HIP
float32 kernels are compared against, bit for bit, on the GPU.
"""
import numpy as np
import pytest

import oracle as O


def sample(f, x, y):
    """sample_scalar_field (advection.py:21-49) at float64 points x, y."""
    h, w = f.shape
    x0 = np.clip(np.floor(x), 0, w - 1).astype(np.int64)
    x1 = np.clip(np.ceil(x), 0, w - 1).astype(np.int64)
    y0 = np.clip(np.floor(y), 0, h - 1).astype(np.int64)
    y1 = np.clip(np.ceil(y), 0, h - 1).astype(np.int64)
    tx = x - x0
    ty = y - y0
    c00, c10 = f[y0, x0], f[y0, x1]
    c01, c11 = f[y1, x0], f[y1, x1]
    # (b - a) in the array's type, then promoted by the float64 weight
    c0 = c00.astype(np.float64) + (c10 - c00).astype(np.float64) * tx
    c1 = c01.astype(np.float64) + (c11 - c01).astype(np.float64) * tx
    return c0 + (c1 - c0) * ty


def advect(kind, src, u, v, mask, dt, cell_size):
    """advect_scalar_field / advect_velocity_u / advect_velocity_v into a new
    array of src's dtype."""
    inv = 1.0 / cell_size
    hh, ww = src.shape
    jj, ii = np.mgrid[0:hh, 0:ww]
    if kind == "scalar":
        px, py = ii + 0.5, jj + 0.5
        uu = sample(u, px, py - 0.5)
        vv = sample(v, px - 0.5, py)
    elif kind == "u":
        px, py = ii.astype(np.float64), jj + 0.5
        uu = src.astype(np.float64)
        vv = sample(v, px - 0.5, py)
    else:
        px, py = ii + 0.5, jj.astype(np.float64)
        uu = sample(u, px, py - 0.5)
        vv = src.astype(np.float64)
    x = px - (uu * dt * inv)
    y = py - (vv * dt * inv)
    if kind == "scalar":
        out = sample(src, x - 0.5, y - 0.5)
    elif kind == "u":
        out = sample(src, x, y - 0.5)
    else:
        out = sample(src, x - 0.5, y)
    return np.where(mask, out, 0.0).astype(src.dtype)


def case(H, W, seed, dtype):
    rng = np.random.default_rng(seed)
    solid = rng.random((H, W)) < 0.15
    solid[0, :] = solid[-1, :] = solid[:, 0] = solid[:, -1] = True
    fmu = np.zeros((H, W + 1), dtype=bool)
    fmv = np.zeros((H + 1, W), dtype=bool)
    fmu[:, 1:-1] = ~(solid[:, :-1] | solid[:, 1:])
    fmv[1:-1, :] = ~(solid[:-1, :] | solid[1:, :])
    fl = fmu[:, :-1] | fmu[:, 1:] | fmv[:-1, :] | fmv[1:, :]
    # velocities large enough that backtraces leave the grid (clamping)
    u = (rng.standard_normal((H, W + 1)) * 6.0).astype(dtype)
    v = (rng.standard_normal((H + 1, W)) * 6.0).astype(dtype)
    q = rng.random((H, W)).astype(dtype)
    return dict(u=u, v=v, q=q, fmu=fmu, fmv=fmv, fl=fl)


SHAPES = ((24, 40, 1), (37, 29, 2), (64, 64, 3))


@pytest.mark.parametrize("H,W,seed", SHAPES)
def test_restatement_float64_equals_oracle(H, W, seed):
    c = case(H, W, seed, np.float64)
    dt, h = 0.7, 0.5
    want = np.zeros_like(c["q"])
    O.advect_scalar_field(c["q"], want, c["u"], c["v"], c["fl"], dt, h)
    np.testing.assert_array_equal(advect("scalar", c["q"], c["u"], c["v"], c["fl"], dt, h), want)
    want = np.zeros_like(c["u"])
    O.advect_velocity_u(c["u"], want, c["v"], c["fmu"], dt, h)
    np.testing.assert_array_equal(advect("u", c["u"], None, c["v"], c["fmu"], dt, h), want)
    want = np.zeros_like(c["v"])
    O.advect_velocity_v(c["v"], want, c["u"], c["fmv"], dt, h)
    np.testing.assert_array_equal(advect("v", c["v"], c["u"], None, c["fmv"], dt, h), want)


def test_float32_typing_differs_from_float64_evaluation():
    """The typing matters: evaluating the float32 inputs' differences in
    float64 instead changes some results, so the GPU comparison below can
    tell the two apart."""
    c = case(64, 64, 3, np.float32)
    f32 = advect("scalar", c["q"], c["u"], c["v"], c["fl"], 0.7, 0.5)
    wide = advect("scalar", c["q"].astype(np.float64), c["u"].astype(np.float64),
                  c["v"].astype(np.float64), c["fl"], 0.7, 0.5).astype(np.float32)
    assert np.count_nonzero(f32 != wide) > 0


@pytest.mark.gpu
@pytest.mark.parametrize("H,W,seed", SHAPES)
def test_hip_float32_advection_bitwise(H, W, seed):
    import torch

    from cfd_amd.backends.hip import operators as hip

    c = case(H, W, seed, np.float32)
    dt, h = 0.7, 0.5
    dev = "cuda"
    t = {k: torch.from_numpy(np.ascontiguousarray(a)).to(dev) for k, a in c.items()}
    out = torch.zeros_like(t["q"])
    hip.advect_scalar_field(t["q"], out, t["u"], t["v"], t["fl"], dt, h)
    np.testing.assert_array_equal(out.cpu().numpy(),
                                  advect("scalar", c["q"], c["u"], c["v"], c["fl"], dt, h))
    out = torch.zeros_like(t["u"])
    hip.advect_velocity_u(t["u"], out, t["v"], t["fmu"], dt, h)
    np.testing.assert_array_equal(out.cpu().numpy(),
                                  advect("u", c["u"], None, c["v"], c["fmu"], dt, h))
    out = torch.zeros_like(t["v"])
    hip.advect_velocity_v(t["v"], out, t["u"], t["fmv"], dt, h)
    np.testing.assert_array_equal(out.cpu().numpy(),
                                  advect("v", c["v"], c["u"], None, c["fmv"], dt, h))
    # the fused u + v launch the simulation uses
    ou, ov = torch.zeros_like(t["u"]), torch.zeros_like(t["v"])
    hip.advect_velocity(t["u"], ou, t["v"], ov, t["fmu"], t["fmv"], dt, h)
    np.testing.assert_array_equal(ou.cpu().numpy(),
                                  advect("u", c["u"], None, c["v"], c["fmu"], dt, h))
    np.testing.assert_array_equal(ov.cpu().numpy(),
                                  advect("v", c["v"], c["u"], None, c["fmv"], dt, h))
