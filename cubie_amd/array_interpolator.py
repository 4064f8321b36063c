"""Driver arrays -> piecewise-polynomial coefficients (host side).

Same input dictionary, defaults and coefficient layout as the reference's
ArrayInterpolator (array_interpolator.py:150-560, 899-1010): samples on a
uniform grid given by ``"time"`` or ``"driver_sample_period"`` + ``"t0"``;
``"order"`` (default 3); ``"wrap"`` (default True, boundary "periodic");
without wrapping the default boundary is "clamped" and the table gains a
zero-valued ghost segment on each side, so the driver ramps to zero outside
the sampled window. Coefficients are stored [segment][driver][power] in
ascending powers of the local coordinate tau in [0, 1), which the device
evaluates with Horner's rule (csrc/device/drivers.hpp).

Order 1 uses the closed form c0 = y_k, c1 = y_{k+1} - y_k; higher orders
solve the global spline system (value at both segment ends, C^1..C^{order-1}
continuity, boundary rows) with one dense solve.
"""

from __future__ import annotations

import math
from typing import Any, Dict, Optional

import numpy as np

CONFIG_KEYS = ("wrap", "order", "boundary_condition")
TIME_KEYS = ("time", "driver_sample_period", "t0", "dt")


class ArrayInterpolator:
    def __init__(self, precision, input_dict: Dict[str, Any], driver_names=None):
        self.precision = np.dtype(precision).type
        self.order = 3
        self.wrap = True
        self.boundary_condition: Optional[str] = None
        self.t0 = 0.0
        self.driver_sample_period = 1e-16
        self.driver_names = list(driver_names) if driver_names is not None else None
        self.input_array = np.zeros((0, 0))
        self.coefficients = None
        self.update_from_dict(input_dict)

    def update_from_dict(self, input_dict: Dict[str, Any]) -> None:
        cfg = {k: v for k, v in input_dict.items() if k in CONFIG_KEYS}
        inputs = {k: v for k, v in input_dict.items() if k not in CONFIG_KEYS and k not in TIME_KEYS}
        times = {k: v for k, v in input_dict.items() if k in TIME_KEYS}
        if "order" in cfg:
            self.order = int(cfg["order"])
            if self.order < 1:
                raise ValueError("order must be >= 1")
        if "wrap" in cfg:
            self.wrap = bool(cfg["wrap"])
        if self.driver_names is not None:
            if set(inputs) != set(self.driver_names):
                raise ValueError(
                    f"input symbols in inputs_dict ({set(inputs)}) do not match drivers "
                    f"symbols in system ({set(self.driver_names)}).")
            names = self.driver_names
        else:
            names = list(inputs)
        cols = []
        length = None
        for nm in names:
            a = np.asarray(inputs[nm], dtype=np.float64).reshape(-1)
            if length is None:
                length = a.size
            elif a.size != length:
                raise ValueError("All driver arrays must have the same length")
            cols.append(a)
        if not cols:
            raise ValueError("No driver arrays supplied")
        if length < 2:
            raise ValueError("Driver arrays need at least two samples")
        self.input_array = np.stack(cols, axis=1)
        self._resolve_time(times)
        bc = cfg.get("boundary_condition", None)
        if bc is None:
            bc = "periodic" if self.wrap else "clamped"
        if bc not in ("natural", "periodic", "clamped", "not-a-knot"):
            raise ValueError(f"Unknown boundary condition {bc}")
        self.boundary_condition = bc
        self.coefficients = self._compute_coefficients()

    def _resolve_time(self, times: Dict[str, Any]) -> None:
        if "time" in times:
            t = np.asarray(times["time"], dtype=np.float64).reshape(-1)
            if t.size != self.input_array.shape[0]:
                raise ValueError("time array length must match the driver samples")
            d = np.diff(t)
            if not np.allclose(d, d[0], rtol=1e-6, atol=0.0):
                raise ValueError("Driver samples must be uniformly spaced in time")
            self.t0 = float(t[0])
            self.driver_sample_period = float((t[-1] - t[0]) / (t.size - 1))
        else:
            period = times.get("driver_sample_period", times.get("dt"))
            if period is None:
                raise ValueError("Either a time array or driver_sample_period must be provided.")
            self.driver_sample_period = float(period)
            self.t0 = float(times.get("t0", 0.0))

    @property
    def num_inputs(self) -> int:
        return self.input_array.shape[1]

    @property
    def pad_clamped(self) -> bool:
        return (not self.wrap) and self.boundary_condition == "clamped"

    @property
    def num_segments(self) -> int:
        base = self.input_array.shape[0] - 1
        return base + 2 if self.pad_clamped else base

    # device constants, cast the way the reference casts them
    def device_constants(self) -> Dict[str, float]:
        p = self.precision
        res = p(self.driver_sample_period)
        inv = p(p(1.0) / res)
        start = p(self.t0)
        ev = p(start - (res if self.pad_clamped else p(0.0)))
        return {"inv_resolution": float(inv), "evaluation_start": float(ev)}

    def _compute_coefficients(self) -> np.ndarray:
        p = self.precision
        base = self.input_array.astype(p)
        k = self.order
        if self.pad_clamped:
            zero = np.zeros((1, base.shape[1]), dtype=p)
            y = np.vstack((zero, base, zero))
        else:
            y = base
        if self.boundary_condition == "periodic":
            if not self.wrap:
                raise ValueError("Periodic boundary conditions require wrap=True")
            if not np.allclose(y[0], y[-1]):
                raise ValueError("Periodic boundary conditions require the first and last samples to match.")
        nseg = y.shape[0] - 1
        if k == 1:
            c = np.empty((nseg, y.shape[1], 2), dtype=p)
            c[:, :, 0] = y[:-1]
            c[:, :, 1] = y[1:] - y[:-1]
            return np.ascontiguousarray(c)
        return self._solve_spline(y.astype(np.float64), k, nseg).astype(p)

    def _solve_spline(self, y: np.ndarray, k: int, nseg: int) -> np.ndarray:
        """Global spline fit in the local coordinate tau in [0, 1] per segment."""
        ncoef = nseg * (k + 1)
        M = np.zeros((ncoef, ncoef))
        R = np.zeros((ncoef, y.shape[1]))
        # falling factorial: d^m/dtau^m tau^j at tau=1 = j!/(j-m)!
        ff = lambda j, m: math.factorial(j) // math.factorial(j - m) if j >= m else 0
        row = 0
        for s in range(nseg):
            M[row, s * (k + 1)] = 1.0
            R[row] = y[s]
            row += 1
        for s in range(nseg):
            M[row, s * (k + 1): (s + 1) * (k + 1)] = 1.0
            R[row] = y[s + 1]
            row += 1
        for s in range(nseg - 1):
            for m in range(1, k):
                for j in range(m, k + 1):
                    M[row, s * (k + 1) + j] = ff(j, m)
                M[row, (s + 1) * (k + 1) + m] -= ff(m, m)
                row += 1
        bc = self.boundary_condition
        need = k - 1
        if bc == "periodic":
            for m in range(1, k):
                for j in range(m, k + 1):
                    M[row, (nseg - 1) * (k + 1) + j] = ff(j, m)
                M[row, m] -= ff(m, m)
                row += 1
        elif bc in ("natural", "clamped"):
            m = 2 if bc == "natural" else 1
            added = 0
            while added < need and m <= k:
                M[row, m] = ff(m, m)
                row += 1
                added += 1
                if added == need:
                    break
                for j in range(m, k + 1):
                    M[row, (nseg - 1) * (k + 1) + j] = ff(j, m)
                row += 1
                added += 1
                m += 1
        else:  # not-a-knot: vanishing differences of the top coefficient
            added = 0
            for order_d in range(1, k):
                if added >= need:
                    break
                for off in range(order_d + 1):
                    coef = (-1) ** (order_d - off) * math.comb(order_d, off)
                    M[row, off * (k + 1) + k] = coef
                row += 1
                added += 1
                if added >= need:
                    break
                for off in range(order_d + 1):
                    coef = (-1) ** (order_d - off) * math.comb(order_d, off)
                    seg = nseg - 1 - (order_d - off)
                    M[row, seg * (k + 1) + k] = coef
                row += 1
                added += 1
        if row != ncoef:
            raise ValueError("Failed to assemble a square spline system")
        sol = np.linalg.solve(M, R)
        return np.ascontiguousarray(sol.reshape(nseg, k + 1, -1).transpose(0, 2, 1))

    def evaluate(self, t) -> np.ndarray:
        """Host evaluation (same arithmetic as the device), for checks/plots."""
        p = self.precision
        c = self.device_constants()
        t = np.asarray(t, dtype=p)
        scaled = (t - p(c["evaluation_start"])) * p(c["inv_resolution"])
        fl = np.floor(scaled)
        idx = fl.astype(np.int64)
        nseg = self.num_segments
        if self.wrap:
            seg = np.mod(idx, nseg)
            tau = (scaled - fl).astype(p)
            inr = np.ones_like(seg, dtype=bool)
        else:
            inr = (scaled >= 0) & (scaled <= nseg)
            seg = np.clip(idx, 0, nseg - 1)
            tau = (scaled - seg.astype(p)).astype(p)
        out = np.zeros(t.shape + (self.num_inputs,), dtype=p)
        for i in range(self.num_inputs):
            acc = np.zeros(t.shape, dtype=p)
            for kk in range(self.order, -1, -1):
                acc = acc * tau + self.coefficients[seg, i, kk]
            out[..., i] = np.where(inr, acc, p(0))
        return out
