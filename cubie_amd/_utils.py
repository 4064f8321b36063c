"""Small helpers shared across the package."""

from __future__ import annotations

import math
from typing import Iterable

import numpy as np

_ALLOWED = (np.float32, np.float64)


def precision_from(precision) -> type:
    """Normalise a dtype-like to ``np.float32`` or ``np.float64``."""
    dt = np.dtype(precision)
    if dt == np.float32:
        return np.float32
    if dt == np.float64:
        return np.float64
    raise ValueError(f"Unsupported precision {precision!r}; use float32 or float64")


def c_literal(value: float, precision) -> str:
    """Exact C++ literal of ``value`` after casting to ``precision``.

    Hexadecimal floating literals round-trip exactly, so the device sees the
    same constant the host computed (the reference casts every compile-time
    constant with ``precision(value)``).
    """
    prec = precision_from(precision)
    v = float(prec(value))
    if math.isnan(v):
        return "__builtin_nan(\"\")" if prec is np.float64 else "__builtin_nanf(\"\")"
    if math.isinf(v):
        inf = "__builtin_huge_val()" if prec is np.float64 else "__builtin_huge_valf()"
        return inf if v > 0 else "(-" + inf + ")"
    text = float.hex(v)
    if prec is np.float32:
        return f"{text}f"
    return text


def c_double(value: float) -> str:
    """Exact C++ double literal."""
    v = float(value)
    if math.isinf(v):
        return "__builtin_huge_val()" if v > 0 else "(-__builtin_huge_val())"
    if math.isnan(v):
        return "__builtin_nan(\"\")"
    return float.hex(v)


def c_array(values: Iterable[float], precision) -> str:
    vals = list(values)
    if not vals:
        vals = [0.0]
    return "{" + ", ".join(c_literal(v, precision) for v in vals) + "}"


def c_int_array(values: Iterable[int]) -> str:
    vals = [int(v) for v in values]
    if not vals:
        vals = [0]
    return "{" + ", ".join(str(v) for v in vals) + "}"
