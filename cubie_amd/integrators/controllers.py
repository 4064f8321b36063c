"""Host-side step-controller settings.

Resolves the controller kind, dt / dt_min / dt_max and the gain constants
exactly as the reference's controller factories do before compiling:
``_resolve_step_params`` / ``_ensure_sane_bounds``
(step_control/adaptive_step_controller.py:157-230), config defaults
(AdaptiveStepControlConfig, adaptive_step_controller.py:44-80; fixed dt
default 1e-3, fixed_step_controller.py), PI exponents ``k/((order+1)*2)``
(adaptive_PI_controller.py:152-153), PID ``k/(2*(order+1))``
(adaptive_PID_controller.py:156-158), I / Gustafsson ``1/(2*(order+1))``
and Gustafsson's ``(1 + 2*target)*safety`` numerator
(gustafsson_controller.py:173-190). Tolerances broadcast to one entry per
state (tol_converter).
"""

from __future__ import annotations

import math
import warnings
from dataclasses import dataclass, field
from typing import Any, Dict, Optional

import numpy as np

CONTROLLER_IDS = {"fixed": 0, "i": 1, "pi": 2, "pid": 3, "gustafsson": 4}

ALL_STEP_CONTROLLER_PARAMETERS = {
    "precision", "n", "step_controller", "dt", "dt_min", "dt_max", "atol", "rtol",
    "algorithm_order", "min_gain", "max_gain", "safety", "kp", "ki", "kd",
    "deadband_min", "deadband_max", "newton_target_iters", "timestep_memory_location",
}


def _tol_array(value, n: int, name: str) -> np.ndarray:
    arr = np.atleast_1d(np.asarray(value, dtype=np.float64))
    if arr.size == 1:
        arr = np.full(n, float(arr[0]))
    if arr.size != n:
        raise ValueError(f"{name} must have length 1 or {n}, got {arr.size}")
    if np.any(arr < 0):
        raise ValueError(f"{name} must be non-negative")
    return arr


@dataclass
class ControllerSettings:
    kind: str
    precision: type
    n: int
    dt: float
    dt_min: float
    dt_max: float
    atol: np.ndarray
    rtol: np.ndarray
    algorithm_order: int = 1
    min_gain: float = 0.3
    max_gain: float = 2.0
    safety: float = 0.9
    kp: float = 0.7
    ki: float = -0.4
    kd: float = 0.0
    deadband_min: float = 1.0
    deadband_max: float = 1.2
    newton_target_iters: int = 20
    user_params: Dict[str, Any] = field(default_factory=dict)

    @property
    def is_adaptive(self) -> bool:
        return self.kind != "fixed"

    @property
    def id(self) -> int:
        return CONTROLLER_IDS[self.kind]

    # Compile-time constants, cast the way the reference casts them.
    def constants(self) -> Dict[str, float]:
        p = self.precision
        k = self.algorithm_order
        out = {
            "safety": float(p(self.safety)),
            "min_gain": float(p(self.min_gain)),
            "max_gain": float(p(self.max_gain)),
            "dt_min": float(p(self.dt_min)),
            "dt_max": float(p(self.dt_max)),
            "deadband_min": float(p(self.deadband_min)),
            "deadband_max": float(p(self.deadband_max)),
            "inv_n": float(p(1.0 / self.n)),
            "order_exponent": float(p(1.0 / (2 * (k + 1)))),
            "kp_exp": 0.0, "ki_exp": 0.0, "kd_exp": 0.0,
            "gain_numerator": float(p(p(1 + 2 * int(self.newton_target_iters)) * p(self.safety))),
        }
        if self.kind == "pi":
            out["kp_exp"] = float(p(self.kp / ((k + 1) * 2)))
            out["ki_exp"] = float(p(self.ki / ((k + 1) * 2)))
        elif self.kind == "pid":
            out["kp_exp"] = float(p(self.kp / (2 * (k + 1))))
            out["ki_exp"] = float(p(self.ki / (2 * (k + 1))))
            out["kd_exp"] = float(p(self.kd / (2 * (k + 1))))
        out["deadband_disabled"] = (out["deadband_min"] == 1.0 and out["deadband_max"] == 1.0)
        return out


def make_controller(settings: Dict[str, Any], precision, n: int, algorithm_order: int) -> ControllerSettings:
    """Build controller settings from merged defaults + user settings."""
    s = dict(settings)
    kind = str(s.pop("step_controller", "fixed")).lower()
    if kind not in CONTROLLER_IDS:
        raise ValueError(f"Unknown controller type: {kind}")
    prec = precision
    atol = _tol_array(s.pop("atol", 1e-6), n, "atol")
    rtol = _tol_array(s.pop("rtol", 1e-6), n, "rtol")
    dt = s.pop("dt", None)
    user = {}
    if kind == "fixed":
        dt_val = 1e-3 if dt is None else float(dt)
        for key in ("dt_min", "dt_max"):
            s.pop(key, None)
        cs = ControllerSettings(kind, prec, n, dt_val, dt_val, dt_val, atol, rtol,
                                algorithm_order=algorithm_order)
        _apply_gains(cs, s)
        return cs
    # adaptive: _resolve_step_params
    if dt is not None:
        user["dt"] = dt
    if "dt_min" in s:
        user["dt_min"] = s["dt_min"]
    if "dt_max" in s:
        user["dt_max"] = s["dt_max"]
    if dt is not None:
        s.setdefault("dt_min", dt / 100)
        s.setdefault("dt_max", dt * 100)
    dt_min = float(s.pop("dt_min", 1e-6))
    dt_max = float(s.pop("dt_max", 1.0))
    if dt is None:
        if "dt_min" in user and "dt_max" in user:
            dt = math.sqrt(float(user["dt_min"]) * float(user["dt_max"]))
    cs = ControllerSettings(kind, prec, n, 0.0, dt_min, dt_max, atol, rtol,
                            algorithm_order=algorithm_order, user_params=user)
    cs.dt = float(dt) if dt is not None else math.sqrt(dt_min * dt_max)
    _apply_gains(cs, s)
    _ensure_sane_bounds(cs)
    return cs


def _apply_gains(cs: ControllerSettings, s: Dict[str, Any]) -> None:
    for key in ("min_gain", "max_gain", "safety", "kp", "ki", "kd", "deadband_min",
                "deadband_max"):
        if key in s and s[key] is not None:
            setattr(cs, key, float(s.pop(key)))
    if "newton_target_iters" in s:
        cs.newton_target_iters = int(s.pop("newton_target_iters"))
    for key in ("algorithm_order", "n", "precision", "timestep_memory_location"):
        s.pop(key, None)
    if not (0 < cs.min_gain <= 1):
        raise ValueError("min_gain must lie in (0, 1]")
    if cs.max_gain < 1:
        raise ValueError("max_gain must be >= 1")
    if not (0 < cs.safety <= 1):
        raise ValueError("safety must lie in (0, 1]")


def _ensure_sane_bounds(cs: ControllerSettings) -> None:
    p = cs.precision
    dt, dt_min, dt_max = float(p(cs.dt)), float(p(cs.dt_min)), float(p(cs.dt_max))
    u_min = cs.user_params.get("dt_min") is not None
    u_max = cs.user_params.get("dt_max") is not None
    if dt_max < dt_min and u_min and u_max:
        raise ValueError(f"dt_max ({dt_max}) < dt_min ({dt_min}). Provide compatible bounds.")
    if dt < dt_min and u_min:
        raise ValueError(f"dt ({dt}) < dt_min ({dt_min}). Provide a compatible dt or adjust dt_min.")
    if dt > dt_max and u_max:
        raise ValueError(f"dt ({dt}) > dt_max ({dt_max}). Provide a compatible dt or adjust dt_max.")
    if dt_max < dt_min and not u_max:
        cs.dt_max = dt_min * 100
    if dt_max < dt_min and not u_min:
        cs.dt_min = dt_max / 100
    if dt < dt_min and not u_min:
        cs.dt_min = dt / 100
    if dt > dt_max and not u_max:
        cs.dt_max = dt * 100
