"""Launch-shape autotuner for the integration kernels.

The per-system loop kernels have two free launch parameters that the
problem does not fix:

* ``block_size``: threads per workgroup. It is compiled into the kernel's
  ``__launch_bounds__``, so it bounds the registers the compiler may use.
* ``blocks_per_cu``: workgroups resident per CU for the persistent
  work-queue kernel (its grid is CUs x blocks_per_cu). Fewer resident waves
  leave the queue more work to balance; more waves hide more latency.

On a given GPU the best pair depends on the system's register footprint and
on how uneven the batch is. The tuner times each candidate on a sample of the
actual batch (the first ``sample_runs`` systems, over a shortened duration),
with the kernel's own hipEvent timing, and returns the fastest. Results are
memoised per kernel configuration for the life of the process. Candidates
whose compiled kernel would not fit the register file at one wave per SIMD
are skipped by the hardware limit check in :meth:`Module.info`.
"""

from __future__ import annotations

import math
from typing import Dict, Iterable, Optional, Tuple

import numpy as np

_MEMO: Dict[tuple, dict] = {}


def autotune(solver, inits, params, duration: float, *, sample_runs: int = 131072,
             sample_duration: Optional[float] = None,
             block_sizes: Iterable[int] = (64, 128, 256),
             blocks_per_cu: Iterable[int] = (1, 2, 4, 8),
             kernel_mode: str = "persistent", rtol_per_run=None, duration_per_run=None,
             apply: bool = True) -> dict:
    """Time every (block_size, blocks_per_cu) candidate on a sample of the
    batch and return ``{"block_size", "blocks_per_cu", "ms", "table"}``.

    With ``apply`` the winner is written into the solver's kernel settings,
    so later ``solve`` calls launch with it.
    """
    kernel = solver.kernel
    n = int(np.asarray(inits).shape[1])
    m = min(n, int(sample_runs))
    sd = float(sample_duration) if sample_duration is not None else float(duration)
    key = (kernel.config_key(sd), m, tuple(block_sizes), tuple(blocks_per_cu), kernel_mode)
    if key in _MEMO:
        best = _MEMO[key]
    else:
        ins = np.asarray(inits)[:, :m]
        prs = np.asarray(params)[:, :m] if np.asarray(params).size else params
        rt = None if rtol_per_run is None else np.asarray(rtol_per_run)[:m]
        dpr = None if duration_per_run is None else np.minimum(np.asarray(duration_per_run)[:m], sd)
        old = (kernel.block_size, kernel.blocks_per_cu)
        table = []
        try:
            for bs in block_sizes:
                kernel.block_size = int(bs)
                bpcs = blocks_per_cu if kernel_mode == "persistent" else (old[1],)
                for bpc in bpcs:
                    kernel.blocks_per_cu = int(bpc)
                    kernel.run(ins, prs, duration=sd, kernel_mode=kernel_mode, timed=True,
                               rtol_per_run=rt, duration_per_run=dpr)          # warm (compile / load)
                    _, ms, _ = kernel.run(ins, prs, duration=sd, kernel_mode=kernel_mode,
                                          timed=True, rtol_per_run=rt, duration_per_run=dpr)
                    table.append({"block_size": int(bs), "blocks_per_cu": int(bpc), "ms": float(ms)})
        finally:
            kernel.block_size, kernel.blocks_per_cu = old
        win = min(table, key=lambda r: r["ms"])
        best = dict(win, table=table, sample_runs=m, sample_duration=sd)
        _MEMO[key] = best
    if apply:
        kernel.block_size = best["block_size"]
        kernel.blocks_per_cu = best["blocks_per_cu"]
    return best
