"""Host-side checks that need no GPU: the C ABI exports what
include/smoke_hip.h declares, the registries and dispatch behave like the
reference's (src/cfd/operators/registry.py, src/cfd/solvers/registry.py),
configuration validation, argument validation, and the workload generator."""
import json
import os
import re
import subprocess
import sys

import numpy as np
import pytest
import torch

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "smoke_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(smk_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from cfd_amd.backends.hip import _lib

    L = _lib.load()
    syms = declared_symbols()
    assert len(syms) >= 25
    for s in syms:
        assert hasattr(L, s), f"{s} declared in smoke_hip.h but not exported"
    # the ctypes table types exactly the declared entry points
    assert sorted(_lib.SIGNATURES) == syms
    assert L.smk_abi_version() == _lib.ABI_VERSION == 6


def test_library_rejects_bad_arguments_without_a_gpu():
    from cfd_amd.backends.hip import _lib

    L = _lib.load()
    # argument validation happens before any HIP call: null pointers, bad
    # dtype and bad sizes return SMK_EINVAL with a message
    assert L.smk_divergence(None, None, None, None, None, 4, 4, 1.0, 0, None) == -1
    assert b"null" in L.smk_last_error()
    assert L.smk_scale(None, 10, 2.0, 0, None) == -1
    assert L.smk_solver_workspace_bytes(0, 4, 0) == 0
    assert L.smk_solver_workspace_bytes(64, 64, 0) > 64 * 64 * 9


def test_operator_registry_contract():
    import cfd_amd.operators  # noqa: F401
    from cfd_amd.operators.registrations import OPERATION_KEYS
    from cfd_amd.operators.registry import OPERATOR_REGISTRY, get_operator, register_operator

    for k in OPERATION_KEYS:
        assert (k, "hip") in OPERATOR_REGISTRY
    with pytest.raises(ValueError, match="already registered"):
        register_operator("compute_divergence", "hip", lambda *a: None)
    with pytest.raises(ValueError, match="Available"):
        get_operator("compute_divergence", "numba")


def test_solver_registry_and_dispatch():
    import cfd_amd.solvers  # noqa: F401
    from cfd_amd.solvers.dispatch import solve_pressure
    from cfd_amd.solvers.params import MultigridParams
    from cfd_amd.solvers.registry import SOLVER_REGISTRY, get_solver, register_solver

    assert {m for m, b in SOLVER_REGISTRY if b == "hip"} == {"jacobi", "gauss_seidel",
                                                             "rb_sor_gauss_seidel", "multigrid"}
    with pytest.raises(ValueError, match="already registered"):
        register_solver("jacobi", "hip", lambda *a, **k: None)
    with pytest.raises(ValueError, match="Unknown pressure solver"):
        get_solver("gauss_seidel", "cuda")

    seen = {}

    def fake(p, d, fu, fv, rho, h, dt, **kw):
        seen.update(kw)

    register_solver("multigrid", "recorder", fake)

    class P:
        pressure_solver_method = "multigrid"
        pressure_solver_backend = "recorder"
        solver_params = MultigridParams(v_cycles=7)
        profiler = "prof"

    solve_pressure(None, None, None, None, 1.0, 1.0, 0.1, P())
    assert seen["v_cycles"] == 7 and seen["profiler"] == "prof"
    assert seen["coarse_scale"] == "reference"  # new fields default to the reference


def test_hip_operators_validate_arguments():
    from cfd_amd.backends.hip import operators as hip

    cpu = torch.zeros((4, 4), dtype=torch.float64)
    with pytest.raises(ValueError, match="GPU"):
        hip.compute_divergence(torch.zeros((4, 5), dtype=torch.float64),
                               torch.zeros((5, 4), dtype=torch.float64), cpu,
                               torch.zeros((4, 5), dtype=torch.bool),
                               torch.zeros((5, 4), dtype=torch.bool), 1.0)
    with pytest.raises(ValueError):
        hip.emit_scalar(cpu, 1.0, 0.1, 1, 1, 2, "spiral")


def test_config_validation_and_roundtrip(tmp_path):
    from cfd_amd.common.config import (Backend, GridConfig, PressureSolverConfig,
                                       SimulationConfig)
    from cfd_amd.solvers.params import JacobiParams, MultigridParams

    cfg = SimulationConfig(grid=GridConfig(size=(20, 12), cell_size=0.5),
                           pressure_solver=PressureSolverConfig(
                               method="multigrid", params=MultigridParams(tolerance=1e-8)),
                           dtype="float64")
    cfg.validate_config()
    path = tmp_path / "cfg.json"
    cfg.to_json(str(path))
    back = SimulationConfig.from_json(str(path))
    assert back.grid.size == (20, 12) and back.pressure_solver.params.tolerance == 1e-8
    assert back.operators.advection == Backend.HIP
    bad = SimulationConfig(grid=GridConfig(size=(-1, 16), cell_size=1.0))
    with pytest.raises(ValueError):
        bad.validate_config()
    wrong = SimulationConfig(grid=GridConfig(size=(8, 8), cell_size=1.0),
                             pressure_solver=PressureSolverConfig(method="multigrid",
                                                                  params=JacobiParams()))
    with pytest.raises(ValueError):
        wrong.validate_config()


def test_workload_generator():
    from cfd_amd import scenes

    a = scenes.plume_512_rbsor(N=64)
    b = scenes.plume_512_rbsor(N=64)
    np.testing.assert_array_equal(a.u0, b.u0)
    assert [(x.x0, x.y0) for x in a.boxes] == [(x.x0, x.y0) for x in b.boxes]
    # the curl of a streamfunction is divergence-free away from the outer
    # faces (which are then closed: the projection removes that part)
    fmu, fmv = scenes.outer_wall_face_masks(a.H, a.W)
    d = np.zeros((a.H, a.W))
    O.compute_divergence(a.u0, a.v0, d, fmu, fmv, a.cell_size)
    assert np.abs(d[1:-1, 1:-1]).max() <= 1e-9 * np.abs(a.u0).max() / a.cell_size
    assert np.abs(a.u0).max() > 1.0
    # 128x128 box: 16x8 cells centred at (0.5, 0.1) -> columns [56,72), rows [9,17)
    c1 = scenes.plume_128_jacobi()
    box = c1.boxes[0]
    assert (box.x0, box.x1, box.y0, box.y1) == (56, 72, 9, 17)
    # multigrid depth: 2048 -> 16 is 8 levels
    assert scenes.mg_levels_to(2048) == 8
    assert scenes.plume_2048_mg(N=64).solver_params["num_levels"] == 3


def test_uneven_scene_sizes():
    from cfd_amd import scenes

    sc = scenes.uneven_scenes(count=12)
    for s in sc:
        assert s.H % 16 == 0 and 64 <= s.H <= 1024
        assert s.W in (s.H, 2 * s.H)
        assert len(s.boxes) == 1


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_bench_byte_model():
    sys.path.insert(0, ROOT)
    import bench

    # V(2,2), 8 levels: 132 B x sum 4^-l (l < 7) + 50 coarsest sweeps x 24 B / 4^7
    per_vc = 132.0 * sum(0.25 ** l for l in range(7)) + 1200.0 / 4 ** 7
    assert abs(per_vc - 176.06) < 0.01
    assert bench.algorithmic_bytes_per_cell(10, False, 8) == pytest.approx(
        168 + 48 + 10 * (per_vc + 16) + 16)
    json.dumps({"x": bench.PEAK_HBM_GBS})
