"""Run, benchmark and suite results (src/cfd/benchmarking/result.py).

The dict layouts (``to_dict`` / ``to_summary``) are the reference's key for
key: the status roll-up ("success" / "partial_failure" / "failure"), the
population standard deviation of run durations, and -1.0 for the duration
statistics of a benchmark with no runs.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field
from typing import Any

from .config import BenchmarkConfig


@dataclass
class RunResult:
    run_index: int
    summary_result: dict[str, Any]
    duration_s: float
    status: str = "success"
    error: str | None = None

    def to_dict(self) -> dict:
        return {"run_index": self.run_index,
                "duration_s": self.duration_s,
                "status": self.status,
                "error": self.error,
                "profiler_summary": self.summary_result}

    def __str__(self) -> str:
        return (f"RunResult(run_index={self.run_index}, status={self.status}, "
                f"duration_s={self.duration_s}, error={self.error})")


def rollup_status(statuses: list[str]) -> str:
    """All runs succeeded / some did / none did (result.py:51)."""
    ok = [s == "success" for s in statuses]
    if all(ok):
        return "success"
    return "partial_failure" if any(ok) else "failure"


@dataclass
class BenchmarkResult:
    benchmark_id: str
    config: BenchmarkConfig
    runs: list[RunResult]
    status: str = "success"
    error: str | None = None

    def duration_stats(self) -> tuple[float, float, float]:
        """(mean, total, population std) of the run durations; all -1.0
        without runs.  The mean divides by the run count (result.py:44-47)."""
        if not self.runs:
            return -1.0, -1.0, -1.0
        d = [r.duration_s for r in self.runs if r.duration_s is not None]
        n = len(self.runs)
        mean = sum(d) / n
        return mean, sum(d), (sum((x - mean) ** 2 for x in d) / n) ** 0.5

    def to_summary(self) -> dict:
        mean, total, std = self.duration_stats()
        good = sum(r.status == "success" for r in self.runs)
        return {"benchmark_id": self.benchmark_id,
                "status": rollup_status([r.status for r in self.runs]),
                "error": self.error,
                "config": self.config.to_dict(),
                "summary": {"total_runs": len(self.runs),
                            "successful_runs": good,
                            "failed_runs": len(self.runs) - good,
                            "duration_avg_s": mean,
                            "duration_total_s": total,
                            "duration_std_s": std}}

    def to_dict(self) -> dict:
        return {"benchmark_id": self.benchmark_id,
                "config": self.config.to_dict(),
                "status": self.status,
                "error": self.error,
                "runs": [r.to_dict() for r in self.runs]}

    def __str__(self) -> str:
        return f"BenchmarkResult(name={self.benchmark_id}, config={self.config}, runs={self.runs})"


@dataclass
class SuiteResult:
    execution_id: str
    name: str
    benchmark_results: list[BenchmarkResult] = field(default_factory=list)

    def to_dict(self) -> dict:
        return {"execution_id": self.execution_id,
                "name": self.name,
                "benchmark_results": [b.to_dict() for b in self.benchmark_results]}

    def to_summary(self) -> dict:
        """summary.json (result.py:84-116).  The success count reads each
        benchmark's ``status``, which BenchmarkRunner sets to the roll-up of
        its runs (the reference leaves it at "success", so its count never
        sees a failed benchmark)."""
        rows = []
        for b in self.benchmark_results:
            s = b.to_summary()
            rows.append({"benchmark_id": b.benchmark_id,
                         "status": s["status"],
                         "result": os.path.join("results", f"{b.benchmark_id}.json"),
                         "duration_avg_s": s["summary"]["duration_avg_s"],
                         "duration_total_s": s["summary"]["duration_total_s"],
                         "duration_std_s": s["summary"]["duration_std_s"]})
        ok = sum(b.status == "success" for b in self.benchmark_results)
        return {"execution_id": self.execution_id,
                "name": self.name,
                "total_benchmarks": len(self.benchmark_results),
                "successful_benchmarks": ok,
                "failed_benchmarks": len(self.benchmark_results) - ok,
                "duration_s": sum(r.duration_s for b in self.benchmark_results for r in b.runs
                                  if r.duration_s is not None),
                "benchmarks": rows}
