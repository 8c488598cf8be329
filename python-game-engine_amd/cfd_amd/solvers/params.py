"""Solver parameter dataclasses (src/cfd/solvers/params.py:5-76).

Field names and defaults are the reference's; fields marked *new* extend it
and default to the reference behaviour:

* SORParams.tolerance / check_every -- stop RB-SOR at a relative residual
  (the reference always runs ``iterations`` sweeps).
* MultigridParams.smoother_omega -- force the RB-SOR smoother's omega (the
  reference always uses the automatic omega per level, solver.py:107).
* MultigridParams.coarse_scale -- "reference" keeps restrict(r)/alpha(2h)
  (solver.py:124-125); "rediscretized" uses restrict(r)/alpha(h).
* MultigridParams.tolerance -- V-cycles until the relative residual is met
  (``v_cycles`` is then the cap).
* MultigridParams.precision / inner_cycles -- "mixed" runs FP32 V-cycles
  inside FP64 defect correction.
"""
from dataclasses import dataclass
from typing import Union


@dataclass
class JacobiParams:
    iterations: int = 40

    def string(self):
        return f"JacobiParams_iters_{self.iterations}"

    def to_fn(self):
        return f"i_{self.iterations}"


@dataclass
class GaussSeidelParams:
    iterations: int = 40

    def string(self):
        return f"GaussSeidelParams_iters_{self.iterations}"

    def to_fn(self):
        return f"i_{self.iterations}"


@dataclass
class SORParams:
    iterations: int = 40
    omega: float = 0.0  # 0.0 -> automatic optimal omega
    tolerance: float = 0.0  # new
    check_every: int = 20  # new
    precision: str = "native"  # new: "mixed" = FP32 sweeps in FP64 defect correction

    def string(self):
        return f"SORParams_iters_{self.iterations}_omega_{self.omega}"

    def to_fn(self):
        return f"i_{self.iterations}_o_{self.omega}"


@dataclass
class MultigridParams:
    smoother_type: str = "jacobi"
    smoother_backend: str = "hip"
    num_levels: int = 4
    v_cycles: int = 2
    smoother_iterations: int = 2
    coarsest_iterations: int = 30
    rebuild_masks: bool = False
    smoother_omega: float = 0.0  # new
    coarse_scale: str = "reference"  # new
    tolerance: float = 0.0  # new
    precision: str = "native"  # new
    inner_cycles: int = 1  # new
    coarsest_omega: float = 0.0  # new: 0 = smoother_omega; < 0 optimal for the coarsest size
    # new: "sweeps" (the reference's), or "operator": the mixed solver's FP32
    # V-cycles apply the coarsest sweeps as their precomputed matrix, or
    # "exact": they apply the coarsest level's exact least-squares solution
    # operator (the limit of the sweeps) instead, or "exact_1024": they stop
    # at the first level of at most 1024 cells and apply its exact operator
    # (one multi-workgroup product replaces that level and those below)
    coarsest_solve: str = "sweeps"
    # new: "float64" (the reference's expression, as a float32 field's solve
    # evaluates it) or "float32": the mixed solver's FP32 V-cycles smooth in
    # float32 arithmetic
    inner_arithmetic: str = "float64"

    def string(self):
        return (f"MultigridParams_smoother_{self.smoother_type}_{self.smoother_backend}"
                f"_vcycles_{self.v_cycles}_levels_{self.num_levels}"
                f"_smoother_iters_{self.smoother_iterations}_coarsest_{self.coarsest_iterations}")

    def to_fn(self):
        return (f"st_{self.smoother_type}_{self.smoother_backend}_v{self.v_cycles}"
                f"_l{self.num_levels}_si{self.smoother_iterations}_ci{self.coarsest_iterations}")


SolverParams = Union[JacobiParams, GaussSeidelParams, SORParams, MultigridParams]

DEFAULT_SOLVER_PARAMS: dict[str, SolverParams] = {
    "jacobi": JacobiParams(),
    "gauss_seidel": GaussSeidelParams(),
    "rb_sor_gauss_seidel": SORParams(),
    "multigrid": MultigridParams(),
}

SOLVER_PARAM_TYPES: dict[str, type] = {
    "jacobi": JacobiParams,
    "gauss_seidel": GaussSeidelParams,
    "rb_sor_gauss_seidel": SORParams,
    "multigrid": MultigridParams,
}
