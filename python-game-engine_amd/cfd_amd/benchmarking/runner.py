"""BenchmarkRunner (src/cfd/benchmarking/runner.py:13-84).

Each repetition builds a fresh simulation and profile session, adds the
emitters, runs the warm-up substeps, resets the profiler, runs the measured
substeps (fixed dt 0.016), and records the wall time from before the warm-up
to after the last substep -- the reference's clock, so durations compare.
The simulation's kernels run asynchronously on the GPU, so the device is
synchronized before the clock is read (and before the profiler reset, so no
warm-up kernel is billed to the measured frames); the reference's CPU
backends need no such step.  An exception inside the frames marks the run
"error" with its traceback; the runner goes on to the next repetition.
"""
from __future__ import annotations

import traceback
from time import perf_counter_ns

import torch

from cfd_amd.common.emitters import Emitter2D
from cfd_amd.common.simulation import SmokeSimulation2D
from cfd_amd.profiling import ProfileSession

from .config import BenchmarkConfig
from .result import BenchmarkResult, RunResult, rollup_status
from .writer import ResultWriter

FRAME_DT = 0.016


def _sync(device: torch.device) -> None:
    if device.type == "cuda":
        torch.cuda.synchronize(device)


class BenchmarkRunner:
    def __init__(self, writer: ResultWriter | None):
        self.writer = writer

    def run(self, config: BenchmarkConfig) -> BenchmarkResult:
        runs = []
        for i in range(config.run.repetitions):
            result, raw = self._run_once(config, i)
            runs.append(result)
            if self.writer is not None:
                self.writer.write_raw(benchmark_id=config.name, run_index=i, raw_data=raw)
        out = BenchmarkResult(benchmark_id=config.name, config=config, runs=runs,
                              status=rollup_status([r.status for r in runs]))
        if self.writer is not None:
            self.writer.write_benchmark(result=out)
        return out

    def _run_once(self, config: BenchmarkConfig, run_index: int) -> tuple[RunResult, dict]:
        profiler = ProfileSession(config=config.profiling, grid_shape=config.simulation.grid.size)
        sim = SmokeSimulation2D(config=config.simulation, profiler=profiler)
        for e in config.emitters:
            sim.add_emitter(Emitter2D.from_dict(e))

        status, error = "success", None
        start = perf_counter_ns()
        try:
            for _ in range(config.run.warmup_frames):
                sim.substep(FRAME_DT)
            _sync(sim.device)
            sim.profiler.reset()
            for _ in range(config.run.frames):
                sim.substep(FRAME_DT)
            _sync(sim.device)
        except Exception:
            status, error = "error", traceback.format_exc()
        finally:
            duration_s = (perf_counter_ns() - start) / 1e9

        result = RunResult(run_index=run_index, summary_result=sim.profiler.to_result().to_dict(),
                           duration_s=duration_s, status=status, error=error)
        return result, sim.profiler.to_dict()
