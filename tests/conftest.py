"""Shared fixtures.  The 16x16 ring-domain fixtures restate the reference's
tests/conftest.py:8-83 (float64 here; the reference uses float32)."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "python-game-engine_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long-running")


@pytest.fixture
def small_grid_size():
    return (16, 16)


@pytest.fixture
def small_random_grid():
    rng = np.random.default_rng(seed=42)
    return rng.random((16, 16), dtype=np.float32).astype(np.float64)


def make_face_masks(height, width, solid_mask):
    face_mask_u = np.zeros((height, width + 1), dtype=bool)
    face_mask_v = np.zeros((height + 1, width), dtype=bool)
    face_mask_u[:, 1:-1] = ~(solid_mask[:, :-1] | solid_mask[:, 1:])
    face_mask_v[1:-1, :] = ~(solid_mask[:-1, :] | solid_mask[1:, :])
    return face_mask_u, face_mask_v


def ring_solid(height, width):
    s = np.zeros((height, width), dtype=bool)
    s[0, :] = s[-1, :] = True
    s[:, 0] = s[:, -1] = True
    return s


@pytest.fixture
def face_masks():
    return make_face_masks(16, 16, ring_solid(16, 16))


@pytest.fixture
def domain():
    solid = ring_solid(16, 16)
    fmu, fmv = make_face_masks(16, 16, solid)
    is_fluid = fmu[:, :-1] | fmu[:, 1:] | fmv[:-1, :] | fmv[1:, :]
    return {"height": 16, "width": 16, "solid_mask": solid, "face_mask_u": fmu,
            "face_mask_v": fmv, "is_fluid": is_fluid}


@pytest.fixture
def cell_size():
    return 1.0


@pytest.fixture
def fluid_density():
    return 1.0
