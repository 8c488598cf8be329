"""Many independent smoke scenes of uneven sizes stepped together on the GPU.

Each scene is a plume as in :mod:`cfd_amd.scenes` (one box source, buoyancy
f_y = density, blob initial velocity, walls on the outer ring) and each step
is ``SmokeSimulation2D.substep`` (``src/cfd/common/simulation.py:186-249``)
with the pressure solved by FP32 V-cycles inside FP64 defect correction to a
per-scene relative tolerance.  The library advances all scenes with one
persistent launch per stage over a packed tile list and replays the stages
from HIP graphs (``smk_batch_*`` in ``include/smoke_hip.h``); a scene's fields
come out exactly as a single-scene :class:`SmokeSimulation2D` with the same
settings would leave them.

:meth:`BatchedSmoke.autotune` picks the smoother tile height, the V(s,s)
sweep count and the persistent grid size by timing steps on the batch itself.
"""
from __future__ import annotations

import ctypes
import itertools
import time

import torch

from cfd_amd.backends.hip import _lib
from cfd_amd.backends.hip import operators as hip


def _levels_to(n: int, coarsest: int = 16) -> int:
    levels = 1
    while n > coarsest and n % 2 == 0:
        n //= 2
        levels += 1
    return levels


class BatchedSmoke:
    """Scenes (``cfd_amd.scenes.Scene`` with one box each) in one batch."""

    def __init__(self, scenes, *, tolerance: float = 1e-6, max_cycles: int = 100,
                 smoother_iterations: int = 2, coarsest_iterations: int = 50,
                 coarsest_omega: float = -1.0, use_graphs: bool = True, tile_rows: int = 48,
                 persistent_blocks: int = 0, smoother_omega: float = 1.0,
                 fluid_density: float = 1.0, device="cuda"):
        if not scenes:
            raise ValueError("need at least one scene")
        dts = {float(s.dt) for s in scenes}
        if len(dts) != 1:
            raise ValueError(f"all scenes must share dt, got {sorted(dts)}")
        self.scenes = list(scenes)
        self.device = torch.device(device)
        self.dt = dts.pop()
        f64 = dict(dtype=torch.float64, device=self.device)
        bl = dict(dtype=torch.bool, device=self.device)
        self.fields = []
        arr = (_lib.BatchScene * len(scenes))()
        for k, sc in enumerate(self.scenes):
            H, W = sc.H, sc.W
            if len(sc.boxes) != 1:
                raise ValueError(f"scene {k}: the batch supports one box source per scene")
            f = {
                "density": [torch.zeros((H, W), **f64), torch.zeros((H, W), **f64)],
                "u": torch.as_tensor(sc.u0, **f64).contiguous(),
                "u_tmp": torch.zeros((H, W + 1), **f64),
                "v": torch.as_tensor(sc.v0, **f64).contiguous(),
                "v_tmp": torch.zeros((H + 1, W), **f64),
                "pressure": torch.zeros((H, W), **f64),
                "divergence": torch.zeros((H, W), **f64),
                "face_mask_u": torch.zeros((H, W + 1), **bl),
                "face_mask_v": torch.zeros((H + 1, W), **bl),
                "is_fluid": torch.zeros((H, W), **bl),
                "solid": torch.zeros((H, W), **bl),
            }
            # simulation.py:102-131: masks from the all-fluid solid mask, then the ring
            hip.face_masks_from_solid(f["solid"], f["face_mask_u"], f["face_mask_v"])
            hip.is_fluid_from_masks(f["face_mask_u"], f["face_mask_v"], f["is_fluid"])
            s = f["solid"]
            s[0, :] = True
            s[-1, :] = True
            s[:, 0] = True
            s[:, -1] = True
            self.fields.append(f)
            b = sc.boxes[0]
            e = arr[k]
            e.H, e.W, e.cell_size = H, W, float(sc.cell_size)
            e.num_levels = int(sc.solver_params.get("num_levels", _levels_to(min(H, W))))
            e.density[0] = f["density"][0].data_ptr()
            e.density[1] = f["density"][1].data_ptr()
            for name in ("u", "u_tmp", "v", "v_tmp", "pressure", "divergence", "face_mask_u",
                         "face_mask_v", "is_fluid", "solid"):
                setattr(e, name, f[name].data_ptr())
            e.box_x0, e.box_y0, e.box_x1, e.box_y1 = b.x0, b.y0, b.x1, b.y1
            e.box_value = float(b.value)
        sc0 = self.scenes[0]
        self.params = _lib.BatchParams(
            dt=self.dt, gravity=float(sc0.gravity), buoyancy=float(sc0.buoyancy),
            density_coefficient=float(sc0.density_coefficient), ambient_temperature=0.0,
            fluid_density=float(fluid_density), smoother_iterations=int(smoother_iterations),
            coarsest_iterations=int(coarsest_iterations),
            coarse_scale=_lib.SMK_COARSE_REDISCRETIZED, coarsest_omega=float(coarsest_omega),
            tolerance=float(tolerance),
            max_cycles=int(max_cycles), use_graphs=1 if use_graphs else 0,
            tile_rows=int(tile_rows), persistent_blocks=int(persistent_blocks),
            smoother_omega=float(smoother_omega))
        self._arr = arr
        h = ctypes.c_void_p()
        torch.cuda.synchronize(self.device)
        _lib.call("smk_batch_create", arr, len(self.scenes), ctypes.byref(self.params),
                  self._stream(), ctypes.byref(h))
        self._h = h
        self.last_info: dict = {}
        self.steps_done = 0

    # ------------------------------------------------------------ state
    def _stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    @property
    def cells(self) -> int:
        return sum(s.H * s.W for s in self.scenes)

    def density(self, k: int) -> torch.Tensor:
        return self.fields[k]["density"][_lib.lib().smk_batch_density_index(self._h)]

    def velocity(self, k: int):
        return self.fields[k]["u"], self.fields[k]["v"]

    def pressure(self, k: int) -> torch.Tensor:
        return self.fields[k]["pressure"]

    def step(self) -> dict:
        info = _lib.BatchInfo()
        _lib.call("smk_batch_step", self._h, ctypes.byref(info), self._stream())
        self.steps_done += 1
        self.last_info = {"iterations": info.iterations, "cycles_total": info.cycles_total,
                          "max_rel_residual": info.max_rel_residual}
        return self.last_info

    def scene_stats(self):
        n = len(self.scenes)
        cyc = (ctypes.c_int * n)()
        rel = (ctypes.c_double * n)()
        _lib.call("smk_batch_scene_stats", self._h, cyc, rel, self._stream())
        return list(cyc), list(rel)

    def set_tuning(self, tile_rows: int, smoother_iterations: int, persistent_blocks: int):
        _lib.call("smk_batch_set_tuning", self._h, int(tile_rows), int(smoother_iterations),
                  int(persistent_blocks), self._stream())
        self.params.tile_rows = int(tile_rows)
        self.params.smoother_iterations = int(smoother_iterations)
        self.params.persistent_blocks = int(persistent_blocks)

    @property
    def tuning(self) -> dict:
        return {"tile_rows": self.params.tile_rows,
                "smoother_iterations": self.params.smoother_iterations,
                "persistent_blocks": self.params.persistent_blocks}

    def _snapshot(self):
        idx = _lib.lib().smk_batch_density_index(self._h)
        return idx, [(f["density"][0].clone(), f["density"][1].clone(), f["u"].clone(),
                      f["v"].clone()) for f in self.fields]

    def _restore(self, snap):
        idx, saved = snap
        for f, (d0, d1, u, v) in zip(self.fields, saved):
            f["density"][0].copy_(d0)
            f["density"][1].copy_(d1)
            f["u"].copy_(u)
            f["v"].copy_(v)
        if _lib.lib().smk_batch_density_index(self._h) != idx:
            raise RuntimeError("density parity changed during autotuning")

    def autotune(self, tile_rows=(48, 32), smoother_iterations=(1, 2, 3),
                 persistent_blocks=(1024, 2048, 4096), repeats: int = 2) -> dict:
        """Time ``repeats`` steps (from the same state, restored afterwards)
        for every combination and keep the fastest.  Each candidate runs two
        steps (two parities) so the state and parity come back unchanged."""
        snap = self._snapshot()
        results = []
        for rh, si, pb in itertools.product(tile_rows, smoother_iterations, persistent_blocks):
            self.set_tuning(rh, si, pb)
            best = float("inf")
            for _ in range(repeats):
                self._restore(snap)
                torch.cuda.synchronize(self.device)
                t0 = time.perf_counter()
                self.step()
                self.step()
                torch.cuda.synchronize(self.device)
                best = min(best, (time.perf_counter() - t0) / 2)
            results.append({"tile_rows": rh, "smoother_iterations": si,
                            "persistent_blocks": pb, "ms_per_step": best * 1e3})
        self._restore(snap)
        win = min(results, key=lambda r: r["ms_per_step"])
        self.set_tuning(win["tile_rows"], win["smoother_iterations"], win["persistent_blocks"])
        self.steps_done -= 2 * repeats * len(results)
        return {"chosen": win, "candidates": results}

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.lib().smk_batch_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
