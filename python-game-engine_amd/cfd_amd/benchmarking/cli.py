"""cfd-bench command line (src/cfd/benchmarking/cli.py; the reference's
``cfd-bench`` script, pyproject.toml:30).  Run as
``python -m cfd_amd.benchmarking <command>``:

    run <config>        run a suite or single benchmark config, write results
                        under ./benchmarks/results (or --output)
    validate <config>   check a config, print "Valid suite config: <name>"
    list <resource>     solvers | operators | backends | timings | metrics

A ValueError prints ``<file>:<line> - <function>: <message>`` and exits 1,
as the reference's does.
"""
from __future__ import annotations

import argparse
import os
import sys
import traceback

RESOURCES = ["solvers", "operators", "backends", "timings", "metrics"]


def build_parser() -> argparse.ArgumentParser:
    parser = argparse.ArgumentParser(prog="cfd-bench", description="CFD Benchmarking CLI")
    sub = parser.add_subparsers(dest="command", required=True)
    run = sub.add_parser("run", help="Run a single benchmark config")
    run.add_argument("config", type=str, help="Path to the benchmark config file")
    run.add_argument("--output", type=str, default=None,
                     help="results root (default: ./benchmarks/results)")
    run.add_argument("--no-raw", action="store_true", help="skip raw/<id>/run_<i>.json")
    val = sub.add_parser("validate", help="Validate a benchmark config or suite config")
    val.add_argument("config", type=str,
                     help="Path to the benchmark config or suite config file")
    lst = sub.add_parser("list", help="List registered benchmark components")
    lst.add_argument("resource", type=str, choices=RESOURCES)
    return parser


def _report(e: ValueError, fallback: str) -> int:
    tb = traceback.extract_tb(e.__traceback__)
    if tb:
        last = tb[-1]
        print(f"{last.filename}:{last.lineno} - {last.name}: {e}")
    else:
        print(f"{fallback}: {e}")
    return 1


def run_command(args) -> int:
    from .config import SuiteConfig
    from .suite_runner import SuiteRunner
    from .writer import ResultWriter

    try:
        suite = SuiteConfig.from_json(args.config)
        root = args.output or os.path.join(os.getcwd(), "benchmarks", "results")
        result = SuiteRunner(ResultWriter(root=root, write_raw=not args.no_raw)).run(suite)
        print(result.to_summary())
    except ValueError as e:
        return _report(e, "Invalid suite config")
    return 0


def validate_command(args) -> int:
    from .config import SuiteConfig

    try:
        suite = SuiteConfig.from_json(args.config)
        suite.validate()
        print(f"Valid suite config: {suite.name}")
    except ValueError as e:
        return _report(e, "Invalid suite config")
    return 0


def list_lines(resource: str) -> list[str]:
    """The lines ``cfd-bench list <resource>`` prints."""
    import cfd_amd.operators  # noqa: F401  (fills the registries)
    import cfd_amd.solvers  # noqa: F401
    from cfd_amd.common.config import Backend
    from cfd_amd.operators.registry import OPERATOR_REGISTRY
    from cfd_amd.profiling import REDUCERS, SCALAR_METRICS, SCOPES, SOURCE_STAGES, TIMINGS
    from cfd_amd.solvers.registry import SOLVER_REGISTRY

    if resource == "solvers":
        return ["Registered solvers:"] + [f"  - {m} (backend: {b})" for m, b in SOLVER_REGISTRY]
    if resource == "operators":
        return ["Registered operators:"] + [f"  - {o} (backend: {b})"
                                            for o, b in OPERATOR_REGISTRY]
    if resource == "backends":
        return ["Registered backends:"] + [f"  - {b.value}" for b in Backend]
    if resource == "timings":
        return ["Registered timings:"] + [f"  - {t}" for t in TIMINGS]
    if resource == "metrics":
        return (["Registered scalar metrics:"] + [f"  - {m}" for m in SCALAR_METRICS]
                + ["", "Registered field sources:"]
                + [f"  - {s} (stages: {', '.join(sorted(st))})" for s, st in SOURCE_STAGES.items()]
                + ["Registered reducers:"] + [f"  - {r}" for r in REDUCERS]
                + ["Registered scopes:"] + [f"  - {s}" for s in SCOPES])
    raise ValueError(f"Unknown resource: {resource}")


def list_command(args) -> int:
    try:
        print("\n".join(list_lines(args.resource)))
    except Exception as e:
        print(f"Error listing {args.resource}: {e}")
        return 1
    return 0


def main(argv=None) -> int:
    args = build_parser().parse_args(argv)
    commands = {"run": run_command, "validate": validate_command, "list": list_command}
    try:
        return commands[args.command](args)
    except ValueError as e:
        print(f"Error: {e}")
        return 1
    except KeyboardInterrupt:
        print("Execution interrupted by user.")
        return 1


if __name__ == "__main__":
    sys.exit(main())
