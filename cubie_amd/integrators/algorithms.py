"""Algorithm resolution: alias / tableau -> step family, tableau, defaults.

Same alias space as the reference's registry (integrators/algorithms/
__init__.py: ``_ALGORITHM_REGISTRY`` plus every tableau alias) and the same
controller defaults per family (generic_erk.py:77-88 ERK_ADAPTIVE_DEFAULTS,
generic_dirk.py:69-104, generic_firk.py:71-104, generic_rosenbrock_w.py:
71-106, backwards_euler.py:55, crank_nicolson.py:41, explicit_euler.py:41).
Explicit Euler is the one-stage tableau of the generic ERK stepper (the
arithmetic is identical). Backward Euler, its predictor-corrector variant
and Crank-Nicolson keep their own steppers, as in the reference: their
Newton starting guesses and Crank-Nicolson's two-solve error estimate are
not expressible as a DIRK tableau. Their tableaus below carry only the
metadata (order, nodes, whether an error estimate exists).
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Any, Dict, Optional, Tuple, Union

from . import tableaus as tb

ERK_ADAPTIVE_DEFAULTS = {
    "step_controller": "pi", "kp": 0.7, "ki": -0.4, "deadband_min": 1.0,
    "deadband_max": 1.0, "min_gain": 0.2, "max_gain": 10.0, "safety": 0.9,
}
IMPLICIT_ADAPTIVE_DEFAULTS = {
    "step_controller": "gustafsson", "deadband_min": 1.0, "deadband_max": 1.2,
    "min_gain": 0.2, "max_gain": 8.0, "safety": 0.9,
}
FIXED_DEFAULTS = {"step_controller": "fixed"}

EXPLICIT_EULER = tb.ERKTableau(name="euler", a=((0.0,),), b=(1.0,), c=(0.0,), order=1)
BACKWARD_EULER = tb.DIRKTableau(name="backwards_euler", a=((1.0,),), b=(1.0,), c=(1.0,), order=1)
CRANK_NICOLSON = tb.DIRKTableau(
    name="crank_nicolson", a=((0.0, 0.0), (0.5, 0.5)), b=(0.5, 0.5),
    b_hat=(1.0, 0.0), c=(0.0, 1.0), order=2)

_FAMILY_BY_NAME = {
    "euler": ("erk", EXPLICIT_EULER),
    "erk": ("erk", tb.DEFAULT_ERK_TABLEAU),
    "dirk": ("dirk", tb.DEFAULT_DIRK_TABLEAU),
    "firk": ("firk", tb.DEFAULT_FIRK_TABLEAU),
    "rosenbrock": ("rosenbrock", tb.DEFAULT_ROSENBROCK_TABLEAU),
    "backwards_euler": ("backwards_euler", BACKWARD_EULER),
    "backwards_euler_pc": ("backwards_euler_pc", BACKWARD_EULER),
    "crank_nicolson": ("crank_nicolson", CRANK_NICOLSON),
}


@dataclass(frozen=True)
class ResolvedAlgorithm:
    name: str
    family: str          # erk | dirk | firk | rosenbrock | backwards_euler[_pc] | crank_nicolson
    tableau: tb.ButcherTableau

    @property
    def is_adaptive(self) -> bool:
        return self.tableau.has_error_estimate

    @property
    def is_implicit(self) -> bool:
        return self.family != "erk"

    @property
    def is_linear(self) -> bool:
        """Linearly implicit (Rosenbrock): one linear solve per stage."""
        return self.family == "rosenbrock"

    @property
    def order(self) -> int:
        return self.tableau.order

    def controller_defaults(self) -> Dict[str, Any]:
        if not self.is_adaptive:
            return dict(FIXED_DEFAULTS)
        if self.family == "erk":
            return dict(ERK_ADAPTIVE_DEFAULTS)
        return dict(IMPLICIT_ADAPTIVE_DEFAULTS)


def _registry() -> Dict[str, Tuple[str, tb.ButcherTableau]]:
    reg = dict(_FAMILY_BY_NAME)
    for k, v in tb.ERK_TABLEAU_REGISTRY.items():
        reg[k.lower()] = ("erk", v)
    for k, v in tb.DIRK_TABLEAU_REGISTRY.items():
        reg[k.lower()] = ("dirk", v)
    for k, v in tb.FIRK_TABLEAU_REGISTRY.items():
        reg[k.lower()] = ("firk", v)
    for k, v in tb.ROSENBROCK_TABLEAUS.items():
        reg[k.lower()] = ("rosenbrock", v)
    return reg


def resolve_algorithm(algorithm: Union[str, tb.ButcherTableau]) -> ResolvedAlgorithm:
    """Return the family and tableau for an alias or a tableau instance."""
    if isinstance(algorithm, tb.ButcherTableau):
        fam = algorithm.family
        if fam == "generic":
            raise TypeError("Tableau instance must be an ERK/DIRK/FIRK/Rosenbrock tableau")
        return ResolvedAlgorithm(algorithm.name, fam, algorithm)
    if not isinstance(algorithm, str):
        raise TypeError(f"Expected algorithm name or tableau, got {type(algorithm).__name__}")
    key = algorithm.lower()
    reg = _registry()
    if key not in reg:
        raise ValueError(f"Unknown algorithm '{algorithm}'.")
    fam, tab = reg[key]
    return ResolvedAlgorithm(key, fam, tab)


def algorithm_is_adaptive(alias: str) -> bool:
    return resolve_algorithm(alias).is_adaptive


ALL_ALGORITHM_STEP_PARAMETERS = {
    "algorithm", "tableau", "newton_atol", "newton_rtol", "newton_max_iters",
    "krylov_atol", "krylov_rtol", "krylov_max_iters", "krylov_residual_reduction",
    "krylov_residual_floor", "linear_correction_type", "preconditioner_order",
    "preconditioner_type",
    "newton_damping", "newton_max_backtracks", "attempt_dense_prediction",
}
