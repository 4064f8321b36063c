"""Pins the oracle to the reference's Julia golden vectors.

The oracle integrates the Julia-gate Lorenz ensemble in Float32 for every
algorithm of the reference's mutual table, at the pinned fixed dt and (for
adaptive tableaus with a matched controller) the pinned tolerance, and must
pass the reference's own matching rule: RMS distance to the Julia finals at
most twice the Julia solution's error against the Vern9 golden. In Float64
the oracle's Dormand-Prince at tolerance 1e-12 must land on the Vern9
golden itself.
"""

import numpy as np
import pytest

import cubie_amd as ca
import julia_gate as jg
from systems import lorenz_julia

GOLDEN = jg.load_golden()
CONSTANTS = jg.controller_constants()
ADAPTIVE = [a for a in jg.ALIASES
            if ca.integrators.algorithms.algorithm_is_adaptive(a)
            and jg.adaptive_settings(GOLDEN, CONSTANTS, a) is not None]


def oracle_final(oracle_module, settings, precision=np.float32):
    inits, params = jg.inputs(GOLDEN)
    solver = ca.Solver(lorenz_julia(precision), save_every=jg.DURATION, output_types=["state"],
                       **settings)
    run = oracle_module.OracleRun(solver.kernel, "lorenz_julia", jg.DURATION)
    out = run.solve(inits, params, threads=8)
    assert np.all(out["status_codes"] == 0)
    return out["state"][-1].T      # (N, 3) state at t = 1


@pytest.mark.parametrize("alias", jg.ALIASES)
def test_oracle_fixed_step_matches_julia(oracle_module, alias):
    final = oracle_final(oracle_module, jg.fixed_settings(GOLDEN, alias))
    ok, ratio, report = jg.matches(final, GOLDEN[f"fixed_{alias}_final"], GOLDEN["golden_states"])
    assert ok, f"{alias}: {report}"


@pytest.mark.parametrize("alias", ADAPTIVE)
def test_oracle_adaptive_matches_julia(oracle_module, alias):
    final = oracle_final(oracle_module, jg.adaptive_settings(GOLDEN, CONSTANTS, alias))
    ok, ratio, report = jg.matches(final, GOLDEN[f"adaptive_{alias}_final"], GOLDEN["golden_states"])
    assert ok, f"{alias}: {report}"


def test_oracle_fp64_lands_on_vern9_golden(oracle_module):
    settings = dict(algorithm="dormand-prince-54", atol=1e-12, rtol=1e-12, dt=1e-3)
    final = oracle_final(oracle_module, settings, precision=np.float64)
    gold = GOLDEN["golden_states"]
    rel = np.abs(final - gold).max() / np.abs(gold).max()
    assert rel <= 1e-10, rel
