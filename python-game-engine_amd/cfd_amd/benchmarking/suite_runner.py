"""SuiteRunner (src/cfd/benchmarking/suite_runner.py:6-46): expand a suite,
run every benchmark in order, write the results layout.

The suite is validated before its execution folder is made, so a bad config
leaves nothing on disk (the reference validates only in ``cfd-bench
validate``)."""
from __future__ import annotations

from datetime import datetime

from .config import SuiteConfig
from .result import SuiteResult
from .runner import BenchmarkRunner
from .writer import ResultWriter


def make_execution_id(suite: SuiteConfig, now: datetime | None = None) -> str:
    """``<name, lower case, spaces and slashes as _>_<YYYYmmddHHMMSS>``."""
    stamp = (now or datetime.now()).strftime("%Y%m%d%H%M%S")
    return f"{suite.name.replace(' ', '_').replace('/', '_').lower()}_{stamp}"


class SuiteRunner:
    def __init__(self, writer: ResultWriter):
        self.writer = writer
        self.runner = BenchmarkRunner(writer)

    def run(self, suite: SuiteConfig) -> SuiteResult:
        suite.validate()
        benchmarks = suite.expand_suite()
        execution_id = make_execution_id(suite)
        self.writer.set_execution_id(execution_id)
        self.writer.initialize_execution(suite.name)
        self.writer.write_config(suite)
        result = SuiteResult(execution_id=execution_id, name=suite.name,
                             benchmark_results=[self.runner.run(b) for b in benchmarks])
        self.writer.write_suite_result(result)
        return result
