"""On-disk results layout (src/cfd/benchmarking/writer.py):

    <root>/<execution_id>/config.json            the suite as given
    <root>/<execution_id>/results/<id>.json      BenchmarkResult.to_summary()
    <root>/<execution_id>/raw/<id>/run_<i>.json  ProfileSession.to_dict() per run
    <root>/<execution_id>/summary.json           SuiteResult.to_summary()

JSON is written with indent 4, as the reference writes it.  The execution
folder must not exist yet (a second run in the same second fails rather than
mixing results); ``root`` itself is created when missing.
"""
from __future__ import annotations

import json
import os

from .config import SuiteConfig
from .result import BenchmarkResult, SuiteResult


class ResultWriter:
    def __init__(self, root: str, write_raw: bool = True):
        self.root = root
        self.write_raw_flag = write_raw
        self.execution_id: str | None = None
        self.execution_dir: str | None = None

    def set_execution_id(self, execution_id: str) -> None:
        self.execution_id = execution_id

    def initialize_execution(self, suite_name: str) -> None:
        if self.execution_id is None:
            raise ValueError("Execution ID must be set before initializing execution.")
        os.makedirs(self.root, exist_ok=True)
        self.execution_dir = os.path.join(self.root, self.execution_id)
        os.mkdir(self.execution_dir)
        for sub in ("results", "raw"):
            os.mkdir(os.path.join(self.execution_dir, sub))

    def _path(self, what: str, *parts: str) -> str:
        if self.execution_dir is None:
            raise ValueError(f"Execution directory must be set before writing {what}.")
        return os.path.join(self.execution_dir, *parts)

    @staticmethod
    def _dump(path: str, data: dict) -> None:
        with open(path, "w") as f:
            json.dump(data, f, indent=4)

    def write_config(self, config: SuiteConfig) -> None:
        self._dump(self._path("config", "config.json"), config.to_dict())

    def write_raw(self, benchmark_id: str, run_index: int, raw_data: dict) -> None:
        if not self.write_raw_flag:
            return
        path = self._path("raw data", "raw", benchmark_id, f"run_{run_index}.json")
        os.makedirs(os.path.dirname(path), exist_ok=True)
        self._dump(path, raw_data)

    def write_benchmark(self, result: BenchmarkResult) -> None:
        self._dump(self._path("benchmark results", "results", f"{result.benchmark_id}.json"),
                   result.to_summary())

    def write_suite_result(self, result: SuiteResult) -> None:
        self._dump(self._path("suite results", "summary.json"), result.to_summary())
