"""cfd-bench: benchmark and suite configs, runners and the on-disk results
layout (src/cfd/benchmarking/).

A suite is one JSON definition whose list-valued entries (grid sizes, cell
sizes, field dtypes, solver methods x backends x params, operator backends,
frames, warm-up frames, repetitions) expand into a product of single
benchmarks.  Each benchmark runs ``repetitions`` fresh simulations; results go
to ``<root>/<execution_id>/`` as ``config.json``, ``results/<id>.json``,
``raw/<id>/run_<i>.json`` and ``summary.json``, the same files and keys as the
reference writes, so runs of the numba / cpp / cuda backends there and of the
hip backend here can be compared file for file.
"""
from .config import (SCHEMA_VERSION, BenchmarkConfig, RunConfig, SuiteConfig, make_benchmark_id,
                     validate_emitters)
from .result import BenchmarkResult, RunResult, SuiteResult
from .runner import BenchmarkRunner
from .suite_runner import SuiteRunner
from .writer import ResultWriter

__all__ = ["SCHEMA_VERSION", "BenchmarkConfig", "RunConfig", "SuiteConfig", "make_benchmark_id",
           "validate_emitters", "BenchmarkResult", "RunResult", "SuiteResult", "BenchmarkRunner",
           "SuiteRunner", "ResultWriter"]
