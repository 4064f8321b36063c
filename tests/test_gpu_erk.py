"""GPU parity of the explicit path (ERK steppers + controllers + outputs).

Each case integrates the same seeded batch with the HIP kernel and with the
CPU oracle (the restated reference loop) and compares saved trajectories,
counters and status words.
"""

import numpy as np
import pytest

import cubie_amd as ca
from cubie_amd.batchsolving.kernel import BatchSolverKernel

pytestmark = pytest.mark.gpu

LORENZ = ["dx = sigma*(y - x)", "dy = x*(rho - z) - y", "dz = x*y - beta*z"]


def lorenz_system(precision):
    return ca.create_ODE_system(LORENZ, states={"x": 1.0, "y": 0.0, "z": 25.0},
                                parameters={"sigma": 10.0, "rho": 28.0, "beta": 8.0 / 3.0},
                                precision=precision, name="lorenz")


def lorenz_batch(n, seed=1):
    rng = np.random.default_rng(seed)
    base = np.array([10.0, 28.0, 8.0 / 3.0])
    params = base[:, None] * rng.uniform(0.9, 1.1, size=(3, n))
    inits = rng.uniform(-1.0, 1.0, size=(3, n)) + np.array([0.0, 0.0, 25.0])[:, None]
    return inits, params


def run_both(solver, oracle_module, system_id, inits, params, duration, **kw):
    res = solver.solve(inits, params, duration=duration, nan_error_trajectories=False, **kw)
    run = oracle_module.OracleRun(solver.kernel, system_id, duration)
    ref = run.solve(inits, params, threads=8)
    return res, ref


def normwise_rel(a, b):
    """max |a - b| / max |b| over each (variable, run) trajectory."""
    num = np.abs(a - b).max(axis=0)
    den = np.maximum(np.abs(b).max(axis=0), 1e-300)
    return float((num / den).max())


def _acc(ic):
    return (ic[:, 2, :].astype(np.int64) - ic[:, 3, :]).sum(axis=0)


def test_lorenz_dopri5_fp64_bitwise_parity(oracle_module):
    """Lorenz batch (1024 systems), DOPRI5 + PI, rtol 1e-6 / atol 1e-8,
    t in [0, 10], saves every 0.01, with the deterministic controller pow.

    Bar: every saved value, time, counter and status word equal to the
    oracle's bit for bit (reported as max |diff| == 0).
    """
    n = 1024
    inits, params = lorenz_batch(n)
    solver = ca.Solver(fp_contract=False, system=lorenz_system(np.float64), algorithm="dormand-prince-54",
                       atol=1e-8, rtol=1e-6, save_every=0.01, deterministic_pow=True,
                       output_types=["state", "time", "iteration_counters"])
    res, ref = run_both(solver, oracle_module, "lorenz", inits, params, 10.0)
    assert res.state.shape == (1001, 4, n)
    assert np.all(res.status_codes == 0)
    assert np.array_equal(res.status_codes, ref["status_codes"])
    assert np.array_equal(res.iteration_counters, ref["iteration_counters"])
    assert float(np.abs(res.state - ref["state"]).max()) == 0.0


def test_lorenz_dopri5_fp64_native_pow(oracle_module):
    """Same batch with the device library's pow (production build).

    pow differs from the host libm in the last bit, so step sizes differ by
    an ulp and chaos (largest Lyapunov exponent ~1-2 for these parameters)
    amplifies it over t in [0, 10]. Bar: identical step counts and statuses,
    accepted steps within 1 % per run, trajectories within 1e-9 normwise on
    t <= 2 and within 1e-3 on the whole horizon.
    """
    n = 1024
    inits, params = lorenz_batch(n)
    solver = ca.Solver(fp_contract=False, system=lorenz_system(np.float64), algorithm="dormand-prince-54",
                       atol=1e-8, rtol=1e-6, save_every=0.01,
                       output_types=["state", "time", "iteration_counters"])
    res, ref = run_both(solver, oracle_module, "lorenz", inits, params, 10.0)
    assert np.array_equal(res.status_codes, ref["status_codes"])
    assert np.array_equal(res.state[:, 3, :], ref["state"][:, 3, :])
    a, b = _acc(res.iteration_counters), _acc(ref["iteration_counters"])
    assert np.all(np.abs(a - b) <= 0.01 * b)
    assert normwise_rel(res.state[:201, :3], ref["state"][:201, :3]) <= 1e-9
    assert normwise_rel(res.state[:, :3], ref["state"][:, :3]) <= 1e-3


@pytest.mark.parametrize("alg", ["classical-rk4", "heun-21", "ralston-33", "euler"])
def test_fixed_step_erk_fp64_bitwise(oracle_module, alg):
    """Errorless tableaus run fixed-step; FP64 results equal the oracle to 1e-12."""
    n = 256
    inits, params = lorenz_batch(n, seed=3)
    solver = ca.Solver(fp_contract=False, system=lorenz_system(np.float64), algorithm=alg, dt=0.001,
                       save_every=0.05, output_types=["state", "iteration_counters"])
    res, ref = run_both(solver, oracle_module, "lorenz", inits, params, 2.0)
    err = normwise_rel(res.state, ref["state"])
    assert err <= 1e-12, err
    assert np.array_equal(res.iteration_counters, ref["iteration_counters"])


@pytest.mark.parametrize("alg", ["bogacki-shampine-32", "tsit5", "cash-karp-54", "fehlberg-45",
                                 "dormand-prince-853", "vern7"])
def test_adaptive_erk_fp64(oracle_module, alg):
    n = 256
    inits, params = lorenz_batch(n, seed=4)
    solver = ca.Solver(fp_contract=False, system=lorenz_system(np.float64), algorithm=alg, atol=1e-8, rtol=1e-6,
                       save_every=0.1, deterministic_pow=True,
                       output_types=["state", "iteration_counters"])
    res, ref = run_both(solver, oracle_module, "lorenz", inits, params, 2.0)
    assert normwise_rel(res.state, ref["state"]) <= 1e-12
    assert np.array_equal(res.iteration_counters, ref["iteration_counters"])


@pytest.mark.parametrize("controller", ["i", "pi", "pid", "gustafsson"])
def test_controllers_fp64(oracle_module, controller):
    n = 128
    inits, params = lorenz_batch(n, seed=5)
    extra = {"kd": 0.1} if controller == "pid" else {}
    solver = ca.Solver(fp_contract=False, system=lorenz_system(np.float64), algorithm="dormand-prince-54",
                       step_controller=controller, atol=1e-7, rtol=1e-5, dt=0.01,
                       save_every=0.1, deterministic_pow=True,
                       output_types=["state", "iteration_counters"], **extra)
    res, ref = run_both(solver, oracle_module, "lorenz", inits, params, 3.0)
    assert normwise_rel(res.state, ref["state"]) <= 1e-12
    assert np.array_equal(res.iteration_counters, ref["iteration_counters"])
    assert np.array_equal(res.status_codes, ref["status_codes"])


def test_lorenz_fp32_matches_oracle(oracle_module):
    """Float32 stage arithmetic: oracle and kernel both run in float32."""
    n = 512
    inits, params = lorenz_batch(n, seed=6)
    solver = ca.Solver(fp_contract=False, system=lorenz_system(np.float32), algorithm="dormand-prince-54",
                       atol=1e-6, rtol=1e-4, save_every=0.01, deterministic_pow=True,
                       output_types=["state", "time", "iteration_counters"])
    res, ref = run_both(solver, oracle_module, "lorenz", inits, params, 1.0)
    assert float(np.abs(res.state - ref["state"]).max()) == 0.0
    assert np.all(res.status_codes == 0)
