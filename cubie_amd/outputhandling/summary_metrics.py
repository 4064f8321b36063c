"""Summary-metric registry.

Same metric names, buffer / output sizes, parameter syntax (``"peaks[3]"``),
combined-metric substitution and legends as the reference's registry
(outputhandling/summarymetrics/metrics.py:98-560, built-ins registered in
summarymetrics/__init__.py order). The device bodies of the built-ins are
hand-written in csrc/device/outputs.hpp and selected by id; a plugin metric
registered with :func:`register_metric` supplies its own HIP C++ update and
save bodies, which are compiled into the kernel.
"""

from __future__ import annotations

import re
import warnings
from dataclasses import dataclass
from typing import Callable, Dict, List, Optional, Sequence, Tuple, Union

SizeSpec = Union[int, Callable[[int], int]]


@dataclass
class SummaryMetric:
    """A summary metric: sizes, naming and (for plugins) device code.

    ``update_body`` / ``save_body`` are HIP C++ statement blocks for plugin
    metrics; they see ``value``, ``buf`` (real*), ``idx`` (int sample index),
    ``param`` (int) in update and ``buf``, ``out``, ``spp`` (samples per
    window), ``param``, ``sample_every`` in save.
    """

    name: str
    buffer_size: SizeSpec
    output_size: SizeSpec
    unit_modification: str = "[unit]"
    output_names: Optional[List[str]] = None
    metric_id: int = -1
    update_body: Optional[str] = None
    save_body: Optional[str] = None

    def buffer_len(self, param: int) -> int:
        return int(self.buffer_size(param)) if callable(self.buffer_size) else int(self.buffer_size)

    def output_len(self, param: int) -> int:
        return int(self.output_size(param)) if callable(self.output_size) else int(self.output_size)


class SummaryMetrics:
    """Registry and request parser."""

    _COMBINED = {
        frozenset(["mean", "std", "rms"]): "mean_std_rms",
        frozenset(["mean", "std"]): "mean_std",
        frozenset(["std", "rms"]): "std_rms",
        frozenset(["max", "min"]): "extrema",
        frozenset(["dxdt_max", "dxdt_min"]): "dxdt_extrema",
        frozenset(["d2xdt2_max", "d2xdt2_min"]): "d2xdt2_extrema",
    }

    def __init__(self):
        self._metrics: Dict[str, SummaryMetric] = {}
        self._names: List[str] = []

    def register_metric(self, metric: SummaryMetric) -> None:
        if metric.name in self._metrics:
            raise ValueError(f"Metric '{metric.name}' is already registered.")
        if metric.metric_id < 0:
            metric.metric_id = 1000 + len(self._names)
        self._metrics[metric.name] = metric
        self._names.append(metric.name)

    @property
    def implemented_metrics(self) -> List[str]:
        return list(self._names)

    def __getitem__(self, name: str) -> SummaryMetric:
        return self._metrics[name]

    # -- request parsing ---------------------------------------------------
    def parse(self, request: Sequence[str]) -> List[Tuple[str, int]]:
        """Return ``[(metric_name, param), ...]`` after combination."""
        parsed: List[Tuple[str, int]] = []
        for item in request:
            m = re.fullmatch(r"\s*([A-Za-z0-9_]+)\s*(?:\[\s*(-?\d+)\s*\])?\s*", item)
            if not m:
                warnings.warn(f"Metric '{item}' is not registered. Skipping.", stacklevel=3)
                continue
            name, param = m.group(1), int(m.group(2)) if m.group(2) else 0
            parsed.append((name, param))
        names = [n for n, _ in parsed]
        params = {n: p for n, p in parsed}
        result: List[Tuple[str, int]] = []
        used = set()
        combos = sorted(self._COMBINED.items(), key=lambda kv: len(kv[0]), reverse=True)
        for name in names:
            if name in used:
                continue
            found = False
            for members, combined in combos:
                if name in members and members.issubset(names):
                    if combined in self._metrics and combined not in [r[0] for r in result]:
                        result.append((combined, 0))
                        used.update(members)
                        found = True
                        break
            if not found:
                result.append((name, params[name]))
                used.add(name)
        validated = []
        for name, param in result:
            if name not in self._metrics:
                warnings.warn(f"Metric '{name}' is not registered. Skipping.", stacklevel=3)
                continue
            metric = self._metrics[name]
            if callable(metric.buffer_size) and param == 0:
                warnings.warn(
                    f"Metric '{name}' has a callable size but parameter is set to 0. "
                    "This results in a size of 0, which is likely not what you want",
                    UserWarning, stacklevel=3)
            validated.append((name, param))
        return validated

    def buffer_height(self, request) -> int:
        return sum(self._metrics[n].buffer_len(p) for n, p in self.parse(request))

    def output_height(self, request) -> int:
        return sum(self._metrics[n].output_len(p) for n, p in self.parse(request))

    def legend(self, request) -> List[str]:
        out = []
        for n, p in self.parse(request):
            m = self._metrics[n]
            k = m.output_len(p)
            if m.output_names is not None:
                out.extend(m.output_names)
            elif k == 1:
                out.append(n)
            else:
                out.extend(f"{n}_{i + 1}" for i in range(k))
        return out

    def unit_modifications(self, request) -> List[str]:
        out = []
        for n, p in self.parse(request):
            m = self._metrics[n]
            out.extend([m.unit_modification] * m.output_len(p))
        return out


summary_metrics = SummaryMetrics()

_BUILTINS = [
    # name, id, buffer, output, unit, output_names  (registration order of the
    # reference's summarymetrics/__init__.py, then the spike counter)
    ("mean", 0, 1, 1, "[unit]", None),
    ("max", 1, 1, 1, "[unit]", None),
    ("rms", 2, 1, 1, "[unit]", None),
    ("peaks", 3, lambda n: 3 + n, lambda n: n, "s", None),
    ("std", 4, 3, 1, "[unit]", None),
    ("min", 5, 1, 1, "[unit]", None),
    ("max_magnitude", 6, 1, 1, "[unit]", None),
    ("extrema", 7, 2, 2, "[unit]", ["max", "min"]),
    ("negative_peaks", 8, lambda n: 3 + n, lambda n: n, "s", None),
    ("mean_std_rms", 9, 3, 3, "[unit]", ["mean", "std", "rms"]),
    ("mean_std", 10, 3, 2, "[unit]", ["mean", "std"]),
    ("std_rms", 11, 3, 2, "[unit]", ["std", "rms"]),
    ("dxdt_max", 12, 2, 1, "[unit]*s^-1", None),
    ("dxdt_min", 13, 2, 1, "[unit]*s^-1", None),
    ("dxdt_extrema", 14, 3, 2, "[unit]*s^-1", ["dxdt_max", "dxdt_min"]),
    ("d2xdt2_max", 15, 3, 1, "[unit]*s^-2", None),
    ("d2xdt2_min", 16, 3, 1, "[unit]*s^-2", None),
    ("d2xdt2_extrema", 17, 4, 2, "[unit]*s^-2", ["d2xdt2_max", "d2xdt2_min"]),
    # Upward crossings of a threshold (param = threshold in the variable's
    # units, integer); counts spikes of a membrane potential.
    ("spike_count", 18, 2, 1, "count", None),
]
for _n, _i, _b, _o, _u, _names in _BUILTINS:
    summary_metrics.register_metric(SummaryMetric(_n, _b, _o, _u, _names, metric_id=_i))


def register_metric(registry: SummaryMetrics = summary_metrics):
    """Class decorator registering a plugin metric (reference metrics.py:98-122).

    The decorated class is instantiated with no arguments and must expose
    ``name``, ``buffer_size``, ``output_size`` and the HIP C++ bodies
    ``update_body`` / ``save_body``.
    """

    def decorator(cls):
        inst = cls()
        metric = SummaryMetric(
            name=inst.name, buffer_size=inst.buffer_size, output_size=inst.output_size,
            unit_modification=getattr(inst, "unit_modification", "[unit]"),
            output_names=getattr(inst, "output_names", None),
            update_body=inst.update_body, save_body=inst.save_body)
        registry.register_metric(metric)
        return cls

    return decorator
