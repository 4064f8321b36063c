"""Loop timing: save / summary schedules, row counts and stop times.

Host arithmetic shared with the device loop, restating
SingleIntegratorRunCore._process_loop_timing (:245-313),
set_summary_timing_from_duration (:330-351), ODELoopConfig.samples_per_summary
(ode_loop_config.py:246-290) and SingleIntegratorRun._regular_event_count /
output_length / save_stop_time / summary_stop_time / summaries_length
(SingleIntegratorRun.py:153-325). The allocation and the device stop times
come from the same event count, so host and device always agree on the
number of rows.
"""

from __future__ import annotations

import math
import warnings
from dataclasses import dataclass
from typing import Optional

import numpy as np


@dataclass
class LoopTiming:
    precision: type
    save_every: Optional[float] = None
    summarise_every: Optional[float] = None
    sample_summaries_every: Optional[float] = None
    save_last: bool = False
    save_regularly: bool = False
    summarise_regularly: bool = False
    duration_dependent: bool = False
    # Explicit save times (offsets from t0 + settling, increasing), in place
    # of save_every: row 0 at the start, row k at save_times[k-1]. An engine
    # extension (the reference saves on a uniform grid only).
    save_times: Optional[tuple] = None

    @property
    def save_table(self) -> bool:
        return self.save_times is not None

    @classmethod
    def resolve(cls, precision, save_every, summarise_every, sample_summaries_every,
                has_time_domain: bool, has_summaries: bool, save_times=None) -> "LoopTiming":
        save_last = False
        duration_dependent = False
        if save_times is not None:
            if save_every is not None:
                raise ValueError("save_every and save_times are mutually exclusive")
            st = np.asarray(save_times, dtype=np.float64).reshape(-1)
            if st.size == 0 or not np.all(np.isfinite(st)) or st[0] <= 0 or np.any(np.diff(st) <= 0):
                raise ValueError("save_times must be positive, finite and strictly increasing")
            save_times = tuple(float(v) for v in st)
        if has_time_domain and save_every is None and save_times is None:
            save_last = True
        if has_summaries:
            if summarise_every is None:
                duration_dependent = True
            elif sample_summaries_every is None:
                sample_summaries_every = summarise_every / 10.0
        else:
            summarise_every = None
            sample_summaries_every = None
        t = cls(precision=precision, save_every=save_every, summarise_every=summarise_every,
                sample_summaries_every=sample_summaries_every, save_last=save_last,
                save_regularly=(save_every is not None or save_times is not None) and has_time_domain,
                summarise_regularly=summarise_every is not None and has_summaries,
                duration_dependent=duration_dependent,
                save_times=save_times if has_time_domain else None)
        if duration_dependent:
            warnings.warn(
                "Summary metrics were requested with no summarise_every or "
                "sample_summaries_every timing. Sample_summaries_every was set to "
                "duration / 100 by default. If duration changes, the kernel will need "
                "to recompile. Set timing parameters explicitly to avoid this.",
                UserWarning, stacklevel=3)
        return t

    def with_duration(self, duration: float) -> "LoopTiming":
        """Summary timing derived from the duration when none was given."""
        if not self.duration_dependent:
            return self
        return LoopTiming(self.precision, self.save_every, float(duration),
                          float(duration) / 100, self.save_last, self.save_regularly,
                          True, True, self.save_times)

    # precision-cast views (ODELoopConfig properties)
    def _p(self, v):
        return None if v is None else self.precision(v)

    @property
    def save_every_p(self):
        return self._p(self.save_every)

    @property
    def summarise_every_p(self):
        return self._p(self.summarise_every)

    @property
    def sample_every_p(self):
        return self._p(self.sample_summaries_every)

    @property
    def samples_per_summary(self) -> int:
        se, ss = self.summarise_every_p, self.sample_every_p
        if se is None or ss is None:
            return 0
        raw = se / ss
        spp = int(round(float(raw)))
        deviation = abs(float(raw) - spp)
        if deviation <= 0.01:
            adjusted = spp * ss
            if adjusted != self.summarise_every:
                warnings.warn(
                    f"summarise_every adjusted from {self.summarise_every} to {adjusted}, the "
                    f"nearest integer multiple of sample_summaries_every ({ss})")
            return spp
        raise ValueError(
            f"summarise_every ({self.summarise_every}) must be an integer multiple of "
            f"sample_summaries_every ({ss}). Under these settings, summaries are "
            f"calculated every {float(raw):.2f} updates, and the calculation can't run "
            "between samples.")

    # -- event counting ------------------------------------------------------
    def regular_event_count(self, duration: float, interval: float) -> int:
        p = self.precision
        total = np.float64(p(duration)) / np.float64(p(interval))
        allowance = min(4.0 * float(np.finfo(p).eps) * float(total), 0.49)
        return int(np.floor(total + allowance))

    def save_table_starts(self, settling_time: float, t0: float) -> float:
        """Start of the save schedule in the working precision."""
        p = self.precision
        return float(p(p(settling_time) + p(t0)))

    def save_table_time(self, k: int, settling_time: float, t0: float) -> float:
        """Time of save row k >= 1 as the device forms it:
        real(double(start) + save_times[k - 1])."""
        start = np.float64(self.save_table_starts(settling_time, t0))
        return float(self.precision(start + np.float64(self.save_times[k - 1])))

    def output_length(self, duration: float) -> int:
        if self.save_table:
            return 1 + len(self.save_times)
        regular = 0
        if self.save_every is not None:
            regular = self.regular_event_count(duration, self.save_every)
        return regular + 1 + (1 if self.save_last else 0)

    def save_stop_time(self, duration: float, settling_time: float, t0: float) -> float:
        start = np.float64(settling_time) + np.float64(t0)
        if self.save_table:
            return self.save_table_time(len(self.save_times), settling_time, t0)
        if self.save_every is None:
            return float(start + np.float64(self.precision(duration)))
        events = self.regular_event_count(duration, self.save_every)
        return float(start + (events + 0.5) * np.float64(self.save_every_p))

    def summary_stop_time(self, duration: float, settling_time: float, t0: float) -> float:
        start = np.float64(settling_time) + np.float64(t0)
        if self.sample_summaries_every is None:
            return float(start + np.float64(self.precision(duration)))
        events = self.regular_event_count(duration, self.sample_summaries_every)
        return float(start + (events + 0.5) * np.float64(self.sample_every_p))

    def summaries_length(self, duration: float) -> int:
        if self.summarise_every is None:
            return 0
        updates = self.regular_event_count(duration, self.sample_summaries_every)
        return updates // max(self.samples_per_summary, 1)
