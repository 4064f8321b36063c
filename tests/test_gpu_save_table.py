"""Explicit save-time tables on the GPU: 200 log-spaced saves of Robertson
kinetics over [0, 1e5] (the stiff kinetics configuration), against the
oracle given the same table."""

import numpy as np
import pytest

import cubie_amd as ca
from systems import robertson, robertson_batch, normwise_rel

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("alg", ["ros3p", "radau"])
def test_robertson_log_spaced_saves_match_oracle(oracle_module, alg):
    n = 128
    inits, params = robertson_batch(n, seed=71)
    table = np.logspace(-5, 5, 200)
    kw = dict(algorithm=alg, atol=1e-10, rtol=1e-6, linear_correction_type="bicgstab",
              save_times=table, output_types=["state", "time", "iteration_counters"])
    solver = ca.Solver(robertson(np.float64), **kw)
    res = solver.solve(inits, params, duration=1e5, nan_error_trajectories=False)
    orc = oracle_module.OracleRun(solver.kernel, "robertson", 1e5).solve(inits, params, threads=8)
    assert res.state.shape == (201, 4, n)
    assert np.array_equal(res.status_codes, orc["status_codes"])
    assert np.all(res.status_codes == 0)
    # every row at its tabulated time, bit for bit
    assert np.array_equal(res.state[1:, 3, :], np.repeat(table[:, None], n, axis=1))
    err = normwise_rel(res.state[:, :3, :], orc["state"][:, :3, :])
    print(alg, "log-spaced saves vs oracle normwise rel err", err)
    assert err <= 1e-9
    assert np.abs(res.state[:, :3, :].sum(axis=1) - 1.0).max() < 1e-8
