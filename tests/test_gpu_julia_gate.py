"""The reference's Julia-matching gate run on the GPU path.

Same protocol as tests/test_oracle_julia.py (Float32 Lorenz ensemble of
1024 rho values, pinned dt / tolerance per algorithm, RMS distance to the
Julia finals at most twice Julia's own error against the Vern9 golden),
with the HIP kernels producing the finals; additionally the GPU finals must
sit within the same bound of the oracle's.
"""

import numpy as np
import pytest

import cubie_amd as ca
import julia_gate as jg
from systems import lorenz_julia

pytestmark = pytest.mark.gpu

GOLDEN = jg.load_golden()
CONSTANTS = jg.controller_constants()
ADAPTIVE = [a for a in jg.ALIASES
            if ca.integrators.algorithms.algorithm_is_adaptive(a)
            and jg.adaptive_settings(GOLDEN, CONSTANTS, a) is not None]


def gpu_and_oracle(oracle_module, settings):
    inits, params = jg.inputs(GOLDEN)
    solver = ca.Solver(lorenz_julia(np.float32), save_every=jg.DURATION, output_types=["state"],
                       **settings)
    res = solver.solve(inits, params, duration=jg.DURATION, nan_error_trajectories=False)
    assert np.all(res.status_codes == 0), res.status_messages
    ref = oracle_module.OracleRun(solver.kernel, "lorenz_julia", jg.DURATION).solve(
        inits, params, threads=8)
    return res.state[-1].T, ref["state"][-1].T


def check(final, oracle_final, julia_final):
    gold = GOLDEN["golden_states"]
    ok, ratio, report = jg.matches(final, julia_final, gold)
    print(report)
    assert ok, report
    # engine vs oracle: well inside the Julia bound (in FP32 an adaptive run
    # can flip an accept / reject decision against the oracle, so this is not
    # a rounding-level bar)
    err_j = jg.ensemble_error(julia_final, gold)
    rms = float(np.sqrt(np.mean((final.astype(np.float64) - oracle_final) ** 2)))
    assert rms <= 0.25 * err_j, (rms, err_j)


@pytest.mark.parametrize("alias", jg.ALIASES)
def test_gpu_fixed_step_matches_julia(oracle_module, alias):
    final, orc = gpu_and_oracle(oracle_module, jg.fixed_settings(GOLDEN, alias))
    check(final, orc, GOLDEN[f"fixed_{alias}_final"])


@pytest.mark.parametrize("alias", ADAPTIVE)
def test_gpu_adaptive_matches_julia(oracle_module, alias):
    final, orc = gpu_and_oracle(oracle_module, jg.adaptive_settings(GOLDEN, CONSTANTS, alias))
    check(final, orc, GOLDEN[f"adaptive_{alias}_final"])
