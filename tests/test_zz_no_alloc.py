"""smk_jacobi_solve_noalloc (static device scratch, no workspace argument)
against smk_jacobi_solve (caller workspace) through the C ABI: bit for bit,
on two calls in a row, the second on a smaller grid, in float64 and
float32, with obstacles in the masks; sizes beyond the static scratch are
rejected before any GPU work."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _case(H, W, seed, dtype):
    rng = np.random.default_rng(seed)
    solid = rng.random((H, W)) < 0.1
    solid[0, :] = solid[-1, :] = solid[:, 0] = solid[:, -1] = True
    fmu = np.zeros((H, W + 1), dtype=bool)
    fmv = np.zeros((H + 1, W), dtype=bool)
    fmu[:, 1:-1] = ~(solid[:, :-1] | solid[:, 1:])
    fmv[1:-1, :] = ~(solid[:-1, :] | solid[1:, :])
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    div = dev(rng.standard_normal((H, W))).to(dtype)
    p0 = dev(rng.standard_normal((H, W))).to(dtype)
    return p0, div, dev(fmu), dev(fmv)


def _solve_ws(p, div, fmu, fmv, iters, code):
    from cfd_amd.backends.hip import _lib

    H, W = p.shape
    nb = _lib.lib().smk_solver_workspace_bytes(H, W, code)
    ws = torch.empty(nb, dtype=torch.uint8, device="cuda")
    _lib.call("smk_jacobi_solve", p.data_ptr(), div.data_ptr(), fmu.data_ptr(), fmv.data_ptr(),
              H, W, 1.0, 1.0 / H, 0.005, iters, ws.data_ptr(), ws.numel(), code,
              _lib.stream_of(p))


def _solve_noalloc(p, div, fmu, fmv, iters, code):
    from cfd_amd.backends.hip import _lib

    H, W = p.shape
    _lib.call("smk_jacobi_solve_noalloc", p.data_ptr(), div.data_ptr(), fmu.data_ptr(),
              fmv.data_ptr(), H, W, 1.0, 1.0 / H, 0.005, iters, code, _lib.stream_of(p))


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_noalloc_jacobi_equals_workspace_jacobi(dtype):
    from cfd_amd.backends.hip.operators import dtype_code

    # the second call is smaller and has an odd iteration count: it reuses
    # the scratch the first left behind and ends with a different parity
    for (H, W, iters, seed) in ((1024, 1024, 40, 1), (200, 136, 13, 2)):
        p0, div, fmu, fmv = _case(H, W, seed, dtype)
        code = dtype_code(p0)
        a, b = p0.clone(), p0.clone()
        _solve_ws(a, div, fmu, fmv, iters, code)
        torch.cuda.synchronize()
        _solve_noalloc(b, div, fmu, fmv, iters, code)
        torch.cuda.synchronize()
        assert torch.equal(a, b), (H, W, dtype)
        assert not torch.equal(a, p0)


def test_noalloc_rejects_grids_beyond_its_scratch():
    from cfd_amd.backends.hip import _lib

    L = _lib.lib()
    p0, div, fmu, fmv = _case(16, 16, 3, torch.float64)
    rc = L.smk_jacobi_solve_noalloc(p0.data_ptr(), div.data_ptr(), fmu.data_ptr(),
                                    fmv.data_ptr(), 1025, 1024, 1.0, 1.0, 1.0, 2, 0, None)
    assert rc == -1
    assert b"exceeds the static scratch" in L.smk_last_error()
