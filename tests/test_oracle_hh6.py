"""The oracle's six-parameter Hodgkin-Huxley (reversal potentials as
parameters, the adjoint workload's system) reduces to the three-parameter
one when those parameters equal the constants."""

import numpy as np

import bench
import cubie_amd as ca
from systems import hodgkin_huxley, hh_driver, hh_batch


def test_hh6_equals_hh_with_constant_reversals(oracle_module):
    inits, params = hh_batch(8, seed=70)
    p6 = np.concatenate([params, np.tile([[50.0], [-77.0], [-54.387]], (1, 8))])
    kw = dict(algorithm="dormand-prince-54", atol=1e-8, rtol=1e-8, save_every=1.0,
              output_types=["state"])
    out = {}
    for name, sysm, p in [("hodgkin_huxley", hodgkin_huxley(np.float64), params),
                          ("hodgkin_huxley6", bench.hh6_system(np.float64), p6)]:
        s = ca.Solver(sysm, **kw)
        s._configure_drivers(hh_driver())
        out[name] = oracle_module.OracleRun(s.kernel, name, 20.0).solve(inits, p, threads=4)
    assert np.array_equal(out["hodgkin_huxley"]["state"], out["hodgkin_huxley6"]["state"])
    assert np.all(out["hodgkin_huxley6"]["status_codes"] == 0)
