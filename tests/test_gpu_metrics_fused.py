"""divergence_residual: the step's metrics pass (divergence of the projected
velocity, then the residual of p against it) in one launch equals the two
operators called one after the other, bit for bit, in float64 and float32,
on ragged shapes with solid cells (closed faces, non-fluid cells)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def fields(H, W, dtype, seed):
    rng = np.random.default_rng(seed)
    solid = rng.random((H, W)) < 0.15
    solid[0, :] = solid[-1, :] = True
    solid[:, 0] = solid[:, -1] = True
    fmu = np.zeros((H, W + 1), dtype=bool)
    fmv = np.zeros((H + 1, W), dtype=bool)
    fmu[:, 1:-1] = ~(solid[:, :-1] | solid[:, 1:])
    fmv[1:-1, :] = ~(solid[:-1, :] | solid[1:, :])
    fluid = ~solid
    dev = "cuda"
    t = lambda a: torch.from_numpy(a).to(dev)
    return (t(rng.standard_normal((H, W + 1))).to(dtype), t(rng.standard_normal((H + 1, W))).to(dtype),
            t(rng.standard_normal((H, W))).to(dtype), t(fluid), t(fmu), t(fmv))


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32], ids=["f64", "f32"])
@pytest.mark.parametrize("shape", [(64, 64), (37, 53), (1, 7), (130, 66)],
                         ids=lambda s: "%dx%d" % s)
def test_fused_metrics_equal_separate_calls(dtype, shape):
    from cfd_amd.backends.hip import operators as hip

    H, W = shape
    u, v, p, fluid, fmu, fmv = fields(H, W, dtype, seed=H * 1000 + W)
    h, alpha, rb = 1.0 / 64, -0.037, 0.25
    div_a = torch.full((H, W), 9.0, dtype=dtype, device="cuda")
    r_a = torch.full((H, W), 9.0, dtype=dtype, device="cuda")
    hip.compute_divergence(u, v, div_a, fmu, fmv, h)
    hip.compute_residual(p, div_a, fluid, fmu, fmv, alpha, rb, r_a)
    div_b = torch.full((H, W), -3.0, dtype=dtype, device="cuda")
    r_b = torch.full((H, W), -3.0, dtype=dtype, device="cuda")
    hip.divergence_residual(u, v, p, fluid, fmu, fmv, h, alpha, rb, div_b, r_b)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(div_b.cpu().numpy(), div_a.cpu().numpy())
    np.testing.assert_array_equal(r_b.cpu().numpy(), r_a.cpu().numpy())


def test_fused_metrics_rejects_bad_arguments():
    from cfd_amd.backends.hip import operators as hip

    u, v, p, fluid, fmu, fmv = fields(16, 16, torch.float64, 1)
    d = torch.zeros((16, 16), dtype=torch.float64, device="cuda")
    r = torch.zeros((16, 16), dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        hip.divergence_residual(u, v, p, fluid, fmu, fmv, 1.0, -1.0, 0.0, d, r)
    with pytest.raises(ValueError):
        hip.divergence_residual(u, v, p, fluid, fmu, fmv, 1.0, -1.0, 0.25, d,
                                torch.zeros((16, 15), dtype=torch.float64, device="cuda"))
