"""The reference's Julia-matching protocol, restated for this engine.

Golden data: DifferentialEquations.jl runs vendored by the reference
(tests/integrated_numerical_tests/julia_reference/data/, README there), of
which tests/golden/julia_lorenz_subset.npz keeps the subset used here (see
tests/golden/make_julia_subset.py): the Lorenz ensemble (sigma 10, beta 8/3,
1024 rho values, y0 = (1, 0, 0), t in [0, 1]) solved by Vern9 at 1e-13 in
Float64, and the per-algorithm Float32 Julia finals at the pinned fixed dt /
adaptive tolerance.

Protocol (julia_reference/ne_gate.py:43-67, 125-247 and
test_julia_reference.py:69-173): an engine's Float32 final states "match"
when the RMS distance to the Julia finals is at most MATCH_FRACTION times the
Julia solution's own RMS error against the Vern9 golden. Fixed-step solves
use step_controller fixed with atol 1e-6 / rtol 1e-3 (the inner Newton and
Krylov tolerances derive from them); adaptive solves use dt0 0.01,
dt_min 1e-6, dt_max 0.5, atol = rtol = the pin, and Julia's resolved
controller constants mapped onto the PI controller (kp = beta1 (p+1),
ki = -beta2 (p+1), safety gamma, gain bounds qmin / qmax, deadband
1 / qsteady) or, for Radau's predictive controller, Gustafsson with the same
safety.
"""

import csv
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")

MATCH_FRACTION = 2.0
DT0, DT_MIN, DT_MAX = 0.01, 1e-6, 0.5
ATOL_FIXED, RTOL_FIXED = 1e-6, 1e-3
DURATION = 1.0

# Every algorithm of the reference's mutual table that this engine runs.
ALIASES = ["dormand-prince-54", "bogacki-shampine-32", "classical-rk4", "heun-21",
           "cash-karp-54", "fehlberg-45", "tsit5", "dormand-prince-853", "vern7",
           "ros3p", "radau_iia_5", "kvaerno3", "kvaerno5", "l_stable_sdirk_4",
           "rodas3p", "rosenbrock23_sciml", "sdirk_2_2", "trapezoidal_dirk",
           "implicit_midpoint", "backwards_euler", "crank_nicolson", "euler"]


def load_golden():
    return dict(np.load(os.path.join(GOLDEN, "julia_lorenz_subset.npz")))


def controller_constants():
    out = {}
    with open(os.path.join(GOLDEN, "julia_controller_constants.csv"), newline="") as f:
        for row in csv.DictReader(f):
            entry = {"controller": row["controller"]}
            for key in ("beta1", "beta2", "qmin", "qmax", "gamma", "qsteady_min",
                        "qsteady_max", "order"):
                raw = row.get(key, "")
                entry[key] = float(raw) if raw not in ("", None) else None
            out[row["cubie_alias"]] = entry
    return out


def matched_controller(constants, alias):
    c = constants.get(alias)
    if c is None:
        return None
    order = c["order"]
    if c["controller"] == "PIController":
        return {"step_controller": "pi", "kp": c["beta1"] * (order + 1),
                "ki": -c["beta2"] * (order + 1), "safety": c["gamma"],
                "min_gain": c["qmin"], "max_gain": c["qmax"],
                "deadband_min": 1.0 / c["qsteady_max"], "deadband_max": 1.0 / c["qsteady_min"]}
    if c["controller"] == "PredictiveController":
        return {"step_controller": "gustafsson", "safety": c["gamma"]}
    return None


def fixed_settings(golden, alias):
    pin = float(golden[f"fixed_{alias}_pin"])
    return dict(algorithm=alias, step_controller="fixed", dt=pin, atol=ATOL_FIXED,
                rtol=RTOL_FIXED)


def adaptive_settings(golden, constants, alias):
    key = f"adaptive_{alias}_pin"
    matched = matched_controller(constants, alias)
    if key not in golden or matched is None:
        return None
    pin = float(golden[key])
    out = dict(algorithm=alias, dt=DT0, dt_min=DT_MIN, dt_max=DT_MAX, atol=pin, rtol=pin)
    out.update(matched)
    return out


def inputs(golden):
    """(3, 1024) initial states and (1, 1024) rho, in the engine's layout."""
    rho = golden["golden_rho"].astype(np.float64)
    n = rho.size
    inits = np.zeros((3, n))
    inits[0] = 1.0
    return inits, rho[None, :]


def ensemble_error(final, golden_states):
    diff = np.asarray(final, dtype=np.float64) - golden_states
    return float(np.sqrt(np.mean(diff ** 2)))


def matches(final, julia_final, golden_states):
    """(ok, ratio, report) of the protocol's comparison for (N, 3) finals."""
    err_c = ensemble_error(final, golden_states)
    err_j = ensemble_error(julia_final, golden_states)
    rms = float(np.sqrt(np.mean((np.asarray(final, np.float64) - julia_final) ** 2)))
    if not (np.isfinite(err_c) and np.isfinite(err_j) and np.isfinite(rms)) or err_j <= 0:
        return False, np.inf, f"invalid: err_engine={err_c} err_julia={err_j} rms={rms}"
    ratio = rms / err_j
    return ratio <= MATCH_FRACTION, ratio, (
        f"err_engine={err_c:.3e} err_julia={err_j:.3e} rms_diff={rms:.3e} ratio={ratio:.3g}")
