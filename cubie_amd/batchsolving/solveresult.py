"""Result container in the reference's layout.

Arrays keep the ``("time", "variable", "run")`` stride order of the
reference's SolveResult (batchsolving/solveresult.py:320-760): the state
buffer carries the time in its last column when ``"time"`` is requested,
``time_domain_array`` stacks states then observables, ``summaries_array``
stacks state then observable summaries, runs with a nonzero status are
overwritten with NaN (except the time column) unless
``nan_error_trajectories=False``, and ``status_messages`` decodes the status
words.
"""

from __future__ import annotations

from typing import Dict, List, Optional

import numpy as np

from ..result_codes import decode_status_codes


class SolveResult:
    stride_order = ("time", "variable", "run")

    def __init__(self, state, observables, state_summaries, observable_summaries,
                 status_codes, iteration_counters, *, save_time: bool, active: Dict[str, bool],
                 time_domain_legend: Dict[int, str], summaries_legend: Dict[int, str],
                 solve_settings: Optional[dict] = None, nan_error_trajectories: bool = True,
                 kernel_ms: Optional[float] = None):
        self._state = state
        self._observables = observables
        self._state_summaries = state_summaries
        self._observable_summaries = observable_summaries
        self.status_codes = status_codes
        self._iteration_counters = iteration_counters
        self._save_time = save_time
        self._active = active
        self.time_domain_legend = time_domain_legend
        self.summaries_legend = summaries_legend
        self.solve_settings = solve_settings or {}
        self.kernel_ms = kernel_ms
        if nan_error_trajectories:
            self._mask_error_runs()

    def _mask_error_runs(self) -> None:
        codes = self.status_codes
        if codes is None or codes.size == 0:
            return
        bad = np.where(codes != 0)[0]
        if bad.size == 0:
            return
        for arr in (self._state_less_time if self._active["state"] else None,
                    self._observables if self._active["observables"] else None,
                    self._state_summaries if self._active["state_summaries"] else None,
                    self._observable_summaries if self._active["observable_summaries"] else None):
            if arr is not None and arr.size:
                arr[:, :, bad] = np.nan

    @property
    def _state_less_time(self):
        if self._state is None:
            return None
        return self._state[:, :-1, :] if self._save_time else self._state

    @property
    def state(self):
        return self._state if (self._active["state"] or self._save_time) else None

    @property
    def observables(self):
        return self._observables if self._active["observables"] else None

    @property
    def state_summaries(self):
        return self._state_summaries if self._active["state_summaries"] else None

    @property
    def observable_summaries(self):
        return self._observable_summaries if self._active["observable_summaries"] else None

    @property
    def iteration_counters(self):
        return self._iteration_counters if self._active["iteration_counters"] else None

    @property
    def time(self):
        if not self._save_time or self._state is None:
            return None
        return self._state[:, -1, :]

    @staticmethod
    def _combine(first, second):
        act = [a for a in (first, second) if a is not None and a.size > 0]
        if not act:
            return np.array([])
        if len(act) == 1:
            return act[0]
        return np.concatenate(act, axis=1)

    @property
    def time_domain_array(self):
        return self._combine(self._state_less_time if self._active["state"] else None,
                             self._observables if self._active["observables"] else None)

    @property
    def summaries_array(self):
        return self._combine(self._state_summaries if self._active["state_summaries"] else None,
                             self._observable_summaries if self._active["observable_summaries"] else None)

    @property
    def as_numpy(self) -> dict:
        ic = self.iteration_counters
        return {
            "time": None if self.time is None else np.array(self.time, copy=True),
            "time_domain_array": np.array(self.time_domain_array, copy=True),
            "summaries_array": np.array(self.summaries_array, copy=True),
            "time_domain_legend": dict(self.time_domain_legend),
            "summaries_legend": dict(self.summaries_legend),
            "iteration_counters": None if ic is None else np.array(ic, copy=True),
        }

    @property
    def status_messages(self) -> Dict[int, List[str]]:
        return decode_status_codes(self.status_codes)

    @property
    def accepted_steps(self) -> Optional[np.ndarray]:
        """Per-run accepted steps (steps - rejections) summed over saves."""
        ic = self.iteration_counters
        if ic is None:
            return None
        return (ic[:, 2, :].astype(np.int64) - ic[:, 3, :]).sum(axis=0)

    def close(self) -> None:
        pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class DeviceSolveResult:
    """Device-resident outputs (torch tensors) plus the stream they ran on."""

    stride_order = ("time", "variable", "run")

    def __init__(self, outputs, stream, save_time: bool, active: Dict[str, bool]):
        self.state = outputs.state if (active["state"] or save_time) else None
        self.observables = outputs.observables if active["observables"] else None
        self.state_summaries = outputs.state_summaries if active["state_summaries"] else None
        self.observable_summaries = outputs.observable_summaries if active["observable_summaries"] else None
        self.iteration_counters = outputs.iteration_counters if active["iteration_counters"] else None
        self.status_codes = outputs.status_codes
        self.stream = stream
        self.save_time = save_time
        self.active_outputs = active
