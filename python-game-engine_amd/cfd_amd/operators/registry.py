"""(operation, backend) -> callable registry.

Same contract as src/cfd/operators/registry.py:6-34: a duplicate key and an
unknown key both raise ValueError, the latter listing what is registered.
"""
from collections.abc import Callable
from typing import Any

OperatorFn = Callable[..., Any]

OPERATOR_REGISTRY: dict[tuple[str, str], OperatorFn] = {}


def register_operator(operation: str, backend: str, fn: OperatorFn) -> None:
    key = (operation, str(backend))
    if key in OPERATOR_REGISTRY:
        raise ValueError(f"Operator already registered: operation={operation!r}, backend={backend!r}")
    OPERATOR_REGISTRY[key] = fn


def get_operator(operation: str, backend: str) -> OperatorFn:
    fn = OPERATOR_REGISTRY.get((operation, str(backend)))
    if fn is None:
        available = [f"{op}/{b}" for op, b in OPERATOR_REGISTRY]
        raise ValueError(f"Unknown operator operation={operation!r}, backend={backend!r}. "
                         f"Available: {available}")
    return fn
