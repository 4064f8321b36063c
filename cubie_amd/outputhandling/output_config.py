"""Output selection: which states / observables are saved and summarised.

Follows the reference's OutputConfig semantics (outputhandling/
output_config.py:120-690): ``output_types`` holds the direct outputs
``"state"``, ``"observables"``, ``"time"``, ``"iteration_counters"`` and any
summary metric names; index arrays select variables; summarised indices
default to the saved ones; unknown output types warn and are ignored.
"""

from __future__ import annotations

import warnings
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

from .summary_metrics import summary_metrics

DIRECT_TYPES = ("state", "observables", "time", "iteration_counters")

ALL_OUTPUT_FUNCTION_PARAMETERS = {
    "output_types", "saved_state_indices", "saved_observable_indices",
    "summarised_state_indices", "summarised_observable_indices",
    "sample_summaries_every", "precision",
}


def _idx(v) -> np.ndarray:
    if v is None:
        return np.zeros(0, dtype=np.int64)
    return np.asarray(v, dtype=np.int64).reshape(-1)


@dataclass
class OutputConfig:
    max_states: int
    max_observables: int
    output_types: Tuple[str, ...] = ("state",)
    saved_state_indices: Optional[np.ndarray] = None
    saved_observable_indices: Optional[np.ndarray] = None
    summarised_state_indices: Optional[np.ndarray] = None
    summarised_observable_indices: Optional[np.ndarray] = None

    def __post_init__(self):
        ot = self.output_types
        if ot is None:
            ot = ()
        elif isinstance(ot, str):
            ot = (ot,)
        self.output_types = tuple(ot)
        self.flags = {k: k in self.output_types for k in DIRECT_TYPES}
        self.summary_types: List[str] = []
        for item in self.output_types:
            if item in DIRECT_TYPES:
                continue
            if any(item.startswith(n) for n in summary_metrics.implemented_metrics):
                self.summary_types.append(item)
            else:
                warnings.warn(f"Summary type '{item}' is not implemented. Ignoring.")
        ss = self.saved_state_indices
        so = self.saved_observable_indices
        self.saved_state_indices = (np.arange(self.max_states) if ss is None else _idx(ss))
        self.saved_observable_indices = (np.arange(self.max_observables) if so is None else _idx(so))
        if self.summarised_state_indices is None:
            self.summarised_state_indices = self.saved_state_indices.copy()
        else:
            self.summarised_state_indices = _idx(self.summarised_state_indices)
        if self.summarised_observable_indices is None:
            self.summarised_observable_indices = self.saved_observable_indices.copy()
        else:
            self.summarised_observable_indices = _idx(self.summarised_observable_indices)
        for name, arr, hi in (("saved_state_indices", self.saved_state_indices, self.max_states),
                              ("saved_observable_indices", self.saved_observable_indices, self.max_observables),
                              ("summarised_state_indices", self.summarised_state_indices, self.max_states),
                              ("summarised_observable_indices", self.summarised_observable_indices, self.max_observables)):
            if arr.size and (arr.min() < 0 or arr.max() >= hi):
                raise ValueError(f"{name} out of range [0, {hi})")
            if len(set(arr.tolist())) != arr.size:
                raise ValueError(f"Duplicate entries in {name}")
        if not any(self.flags.values()) and not self.summary_types:
            raise ValueError(
                "At least one output type must be enabled (state, observables, time, "
                "iteration_counters, summaries)")
        self.metrics = summary_metrics.parse(self.summary_types)

    # -- flags ---------------------------------------------------------------
    @property
    def save_state(self) -> bool:
        return self.flags["state"] and self.saved_state_indices.size > 0

    @property
    def save_observables(self) -> bool:
        return self.flags["observables"] and self.saved_observable_indices.size > 0

    @property
    def save_time(self) -> bool:
        return self.flags["time"]

    @property
    def save_counters(self) -> bool:
        return self.flags["iteration_counters"]

    @property
    def save_summaries(self) -> bool:
        return len(self.metrics) > 0

    @property
    def summarise_state(self) -> bool:
        return self.save_summaries and self.summarised_state_indices.size > 0

    @property
    def summarise_observables(self) -> bool:
        return self.save_summaries and self.summarised_observable_indices.size > 0

    @property
    def time_domain_outputs(self) -> bool:
        return self.save_state or self.save_observables or self.save_time or self.save_counters

    @property
    def n_saved_states(self) -> int:
        return int(self.saved_state_indices.size) if self.save_state else 0

    @property
    def n_saved_observables(self) -> int:
        return int(self.saved_observable_indices.size) if self.save_observables else 0

    @property
    def n_summarised_states(self) -> int:
        return int(self.summarised_state_indices.size) if self.summarise_state else 0

    @property
    def n_summarised_observables(self) -> int:
        return int(self.summarised_observable_indices.size) if self.summarise_observables else 0

    @property
    def summaries_buffer_height_per_var(self) -> int:
        return sum(summary_metrics[n].buffer_len(p) for n, p in self.metrics)

    @property
    def summaries_output_height_per_var(self) -> int:
        return sum(summary_metrics[n].output_len(p) for n, p in self.metrics)

    @property
    def summary_legend_per_variable(self) -> List[str]:
        return summary_metrics.legend(self.summary_types)

    @property
    def state_output_width(self) -> int:
        return self.n_saved_states + (1 if self.save_time else 0)
