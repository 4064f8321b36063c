"""Per-run status bit vocabulary (reference result_codes.py:22-88).

Same names and values as the reference's ``CUBIE_RESULT_CODES``; the device
kernels set these bits in each run's int32 status word.
``MAX_LOOP_ITERS_EXCEEDED`` (reserved in the reference) is set by this
engine's optional attempted-step cap. ``ADJOINT_TAPE_FULL`` is this
engine's own: the recording pass of the discrete adjoint ran out of tape
or checkpoint space for that run.
"""

from __future__ import annotations

from enum import IntFlag
from typing import Dict, Iterable, List, Optional


class CUBIE_RESULT_CODES(IntFlag):
    SUCCESS = 0
    MAX_NEWTON_ITERATIONS_EXCEEDED = 2
    MAX_LINEAR_ITERATIONS_EXCEEDED = 4
    STEP_TOO_SMALL = 8
    DT_EFF_EFFECTIVELY_ZERO = 16
    MAX_LOOP_ITERS_EXCEEDED = 32
    STAGNATION = 64
    BICGSTAB_BREAKDOWN = 128
    NEWTON_DIVERGENCE = 256
    ADJOINT_TAPE_FULL = 4096


def decode_status_codes(status_codes: Optional[Iterable[int]]) -> Dict[int, List[str]]:
    """Map run index -> names of the flags set in its status word.

    Successful runs are omitted, so an empty mapping means every run
    succeeded.
    """
    out: Dict[int, List[str]] = {}
    if status_codes is None:
        return out
    flags = [f for f in CUBIE_RESULT_CODES if f.value != 0]
    for i, code in enumerate(status_codes):
        code = int(code)
        if code == 0:
            continue
        out[i] = [f.name for f in flags if code & f.value == f.value]
    return out
