"""Run-grid construction: label dictionaries or arrays -> (variable, run).

Follows the reference's BatchInputHandler semantics
(batchsolving/BatchInputHandler.py:174-470, 526-945): a dict sweeps the named
variables (``"combinatorial"`` = Cartesian product of de-duplicated values,
``"verbatim"`` = columns paired as given) and every other variable keeps its
default; a 2-D array of shape (n_vars, n_runs) is used as-is; the state and
parameter grids are combined by Cartesian product of runs (combinatorial) or
paired with single-run broadcasting (verbatim).
"""

from __future__ import annotations

from itertools import product
from typing import Dict, List, Optional, Sequence, Tuple, Union

import numpy as np


def unique_cartesian_product(arrays: List[np.ndarray]) -> np.ndarray:
    dedup = [list(dict.fromkeys(np.asarray(a).reshape(-1).tolist())) for a in arrays]
    return np.array([list(t) for t in product(*dedup)], dtype=np.float64).T


def combine_grids(grid1: np.ndarray, grid2: np.ndarray, kind: str = "combinatorial"):
    if kind == "combinatorial":
        return np.repeat(grid1, grid2.shape[1], axis=1), np.tile(grid2, (1, grid1.shape[1]))
    if kind == "verbatim":
        r1, r2 = grid1.shape[1], grid2.shape[1]
        if r1 == 1 and r2 > 1:
            grid1 = np.repeat(grid1, r2, axis=1)
        elif r2 == 1 and r1 > 1:
            grid2 = np.repeat(grid2, r1, axis=1)
        if grid1.shape[1] != grid2.shape[1]:
            raise ValueError(
                "For 'verbatim', both grids must have the same number of runs, or one grid "
                "must have exactly 1 run so it can be broadcast to match the other.")
        return grid1, grid2
    raise ValueError(f"Unknown grid type '{kind}'. Use 'combinatorial' or 'verbatim'.")


class SystemValues:
    """Ordered names + default values of one variable category."""

    def __init__(self, names: Sequence[str], defaults: Dict[str, float], precision):
        self.names = list(names)
        self.precision = precision
        self.values_array = np.array([defaults.get(n, 0.0) for n in self.names], dtype=np.float64)

    @property
    def empty(self) -> bool:
        return len(self.names) == 0

    def get_indices(self, keys) -> np.ndarray:
        out = []
        for k in keys:
            if isinstance(k, (int, np.integer)):
                if not 0 <= int(k) < len(self.names):
                    raise KeyError(f"index {k} out of range")
                out.append(int(k))
            else:
                if k not in self.names:
                    raise KeyError(f"'{k}' is not a variable of this category ({self.names})")
                out.append(self.names.index(k))
        return np.asarray(out, dtype=np.int64)


class BatchInputHandler:
    def __init__(self, system):
        self.system = system

    def _values(self):
        s = self.system
        return (SystemValues(s.state_names, s.initial_values, s.precision),
                SystemValues(s.parameter_names, s.parameters, s.precision))

    def _single(self, data, values: SystemValues, kind: str) -> np.ndarray:
        if values.empty:
            if data is not None and np.asarray(data if not isinstance(data, dict) else list(data)).size:
                raise ValueError("Grid values were provided but the system has no settable "
                                 "variables of this type.")
            return np.empty((0, 1))
        if data is None:
            return values.values_array[:, None].copy()
        if isinstance(data, dict):
            req = {k: np.atleast_1d(np.asarray(v, dtype=np.float64)) for k, v in data.items()}
            req = {k: v for k, v in req.items() if v.size > 0}
            if not req:
                return values.values_array[:, None].copy()
            idx = values.get_indices(list(req.keys()))
            if kind == "combinatorial":
                grid = unique_cartesian_product(list(req.values()))
            elif kind == "verbatim":
                lens = {v.size for v in req.values()}
                lmax = max(lens)
                if not lens <= {1, lmax}:
                    raise ValueError("verbatim sweeps need equal-length value arrays")
                grid = np.array([np.broadcast_to(v, (lmax,)) for v in req.values()])
            else:
                raise ValueError(f"Unknown grid type '{kind}'")
            full = np.repeat(values.values_array[:, None], grid.shape[1], axis=1)
            full[idx, :] = grid
            return full
        arr = np.asarray(data, dtype=np.float64)
        if arr.size == 0:
            return values.values_array[:, None].copy()
        if arr.ndim == 1:
            if arr.size != len(values.names):
                raise ValueError(f"Expected {len(values.names)} values, got {arr.size}")
            return arr[:, None].copy()
        if arr.ndim != 2 or arr.shape[0] != len(values.names):
            raise ValueError(f"Array input must have shape ({len(values.names)}, n_runs)")
        return arr

    def __call__(self, states=None, params=None, kind: str = "combinatorial") -> Tuple[np.ndarray, np.ndarray]:
        sv, pv = self._values()
        prec = self.system.precision
        # Prebuilt (variable, run) arrays of equal run count pass straight through.
        if (isinstance(states, np.ndarray) and isinstance(params, np.ndarray) and states.ndim == 2
                and params.ndim == 2 and states.shape[0] == len(sv.names)
                and params.shape[0] == len(pv.names) and states.shape[1] == params.shape[1]):
            return (np.ascontiguousarray(states, dtype=prec), np.ascontiguousarray(params, dtype=prec))
        s = self._single(states, sv, kind)
        p = self._single(params, pv, kind)
        s, p = combine_grids(s, p, kind)
        return np.ascontiguousarray(s, dtype=prec), np.ascontiguousarray(p, dtype=prec)
