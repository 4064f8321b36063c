"""Benchmark: accepted steps/s of the batch integration kernel on one MI355X.

One "step" of this benchmark is one launch of the integration kernel over
the whole synthetic batch (every system from t0 to t_end), inputs already
resident in HBM. The timed region is K launches bracketed by
torch.cuda.synchronize(); each launch is also timed with a pair of hipEvents
recorded on the launch stream (the roofline uses their mean).

value = total accepted steps (sum over systems of steps - rejections, read
from the kernel's own iteration counters) / mean kernel time.

Prints one JSON line (see the repository's benchmark contract).
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"f64": 78.6, "f32": 157.3}   # MI355X published dense peaks, TFLOP/s
HBM_PEAK = 8000.0                    # GB/s


# ---------------------------------------------------------------------------
# Workloads
# ---------------------------------------------------------------------------

def lorenz_workload(n, precision=np.float64):
    import cubie_amd as ca
    sysm = ca.create_ODE_system(
        ["dx = sigma*(y - x)", "dy = x*(rho - z) - y", "dz = x*y - beta*z"],
        states={"x": 1.0, "y": 0.0, "z": 25.0},
        parameters={"sigma": 10.0, "rho": 28.0, "beta": 8.0 / 3.0},
        precision=precision, name="lorenz")
    rng = np.random.default_rng(1234)
    params = np.array([10.0, 28.0, 8.0 / 3.0])[:, None] * rng.uniform(0.9, 1.1, (3, n))
    inits = rng.uniform(-1, 1, (3, n)) + np.array([0.0, 0.0, 25.0])[:, None]
    solver_kw = dict(algorithm="dormand-prince-54", atol=1e-8, rtol=1e-6, save_every=0.01,
                     output_types=["state", "iteration_counters"])
    return {"name": "lorenz_dopri5", "system": sysm, "inits": inits, "params": params,
            "duration": 10.0, "solver_kw": solver_kw, "oracle_system": "lorenz",
            "config": {"workload": "Lorenz DOPRI5 rtol 1e-6 atol 1e-8, t in [0,10], save every 0.01",
                       "n_systems": n}}


HH_EQUATIONS = """
alpha_m = 0.1*(V + 40)/(1 - exp(-(V + 40)/10))
beta_m = 4*exp(-(V + 65)/18)
alpha_h = 0.07*exp(-(V + 65)/20)
beta_h = 1/(1 + exp(-(V + 35)/10))
alpha_n = 0.01*(V + 55)/(1 - exp(-(V + 55)/10))
beta_n = 0.125*exp(-(V + 65)/80)
dV = (I_ext - gNa*m**3*h*(V - ENa) - gK*n**4*(V - EK) - gL*(V - EL))/Cm
dm = alpha_m*(1 - m) - beta_m*m
dh = alpha_h*(1 - h) - beta_h*h
dn = alpha_n*(1 - n) - beta_n*n
"""


def hh_system(precision):
    import cubie_amd as ca
    return ca.create_ODE_system(
        HH_EQUATIONS, states={"V": -65.0, "m": 0.0529, "h": 0.5961, "n": 0.3177},
        parameters={"gNa": 120.0, "gK": 36.0, "gL": 0.3},
        constants={"ENa": 50.0, "EK": -77.0, "EL": -54.387, "Cm": 1.0},
        drivers=["I_ext"], precision=precision, name="hodgkin_huxley")


def hh_workload(n, precision=np.float64, duration=500.0, dt=0.01):
    """Hodgkin-Huxley batch: conductances uniform within +-30 % of 120/36/0.3,
    driver = 1000 samples (0.5 ms apart) of 10 sin(2 pi t / 50) + N(0, 2),
    linearly interpolated; L-stable SDIRK3 at fixed dt with BiCGSTAB; summary
    outputs only (mean V, max V, spike count above 0 mV) over the whole run,
    sampled every 0.05 ms; iteration counters at the end for the step tally."""
    rng = np.random.default_rng(2024)
    params = np.array([120.0, 36.0, 0.3])[:, None] * rng.uniform(0.7, 1.3, (3, n))
    inits = np.tile(np.array([[-65.0], [0.0529], [0.5961], [0.3177]]), (1, n))
    drng = np.random.default_rng(7)
    t = np.arange(1000) * 0.5
    current = 10.0 * np.sin(2.0 * np.pi * t / 50.0) + drng.normal(0.0, 2.0, 1000)
    solver_kw = dict(algorithm="l_stable_dirk_3", dt=dt, atol=1e-6, rtol=1e-6,
                     linear_correction_type="bicgstab", save_every=duration,
                     summarise_every=duration, sample_summaries_every=0.05,
                     summarise_variables=["V"], save_variables=["V"],
                     output_types=["iteration_counters", "mean", "max", "spike_count[0]"])
    return {"name": "hh_sdirk3", "system": hh_system(precision), "inits": inits, "params": params,
            "duration": duration, "solver_kw": solver_kw, "oracle_system": "hodgkin_huxley",
            "drivers": {"I_ext": current, "dt": 0.5, "order": 1, "wrap": False},
            "config": {"workload": f"Hodgkin-Huxley L-stable SDIRK3 + BiCGSTAB, dt {dt} ms, "
                                   f"t in [0,{duration:g}] ms, summaries mean/max/spike count of V",
                       "n_systems": n}}


WORKLOADS = {"lorenz": lorenz_workload, "hh": hh_workload}


# ---------------------------------------------------------------------------
# Flop model (algorithmic work per attempted step)
# ---------------------------------------------------------------------------

def flops_per_attempted_step(kernel, spec):
    """Algorithmic flops of one attempted step (add/mul/div = 1, FMA = 2,
    transcendental = 1), from the generated system's IR counts and the
    tableau's nonzero structure."""
    tab = kernel.algorithm.tableau
    n = kernel.system.num_states
    ops = spec.op_counts
    f_rhs = ops["dxdt"]["flops"] + ops["dxdt"]["transcendental"]
    s = tab.stage_count
    rhs_evals = s - (1 if tab.first_same_as_last else 0)
    nnz_a = sum(1 for i in range(s) for j in range(i) if tab.a[i][j] != 0.0)
    stage_forms = 2 * n * nnz_a + 2 * n * (s - 1)            # sum a_ij k_j, *dt + y
    sol = 0 if tab.b_matches_a_row is not None else 2 * n * sum(1 for b in tab.b if b) + 2 * n
    err = 0
    if kernel.algorithm.is_adaptive:
        d = tab.d
        err = 2 * n * sum(1 for x in d if x) + n
    norm = 7 * n + 2                                       # |e|, max, tol, ratio, square, sum
    ctrl = 12                                             # 2 pow + clamps + products
    return rhs_evals * f_rhs + stage_forms + sol + err + norm + ctrl


# ---------------------------------------------------------------------------

def cpu_baseline(work, solver, seconds_target=15.0):
    """Oracle (CPU restatement) on a bounded sample, all host threads."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    orc.build()
    threads = os.cpu_count() or 1
    run = orc.OracleRun(solver.kernel, work["oracle_system"], work["duration"])
    n_total = work["inits"].shape[1]
    sample = min(n_total, 256 * threads)
    t = time.perf_counter()
    res = run.solve(work["inits"][:, :sample], work["params"][:, :sample], threads=threads)
    el = time.perf_counter() - t
    # grow the sample once to reach a 10-30 s measurement when the first is short
    if el < seconds_target / 4 and sample < n_total:
        sample = int(min(n_total, sample * max(2, seconds_target / max(el, 1e-3))))
        t = time.perf_counter()
        res = run.solve(work["inits"][:, :sample], work["params"][:, :sample], threads=threads)
        el = time.perf_counter() - t
    ic = res["iteration_counters"]
    acc = int((ic[:, 2, :].astype(np.int64) - ic[:, 3, :]).sum())
    return {"value": acc / el, "unit": "accepted steps/s", "cores": threads, "kind": "port",
            "sample": f"{sample} of {n_total} systems, {el:.1f} s on {threads} host threads"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="lorenz", choices=sorted(WORKLOADS))
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--compile-only", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              "(systems are independent: a multi-GPU run would shard the batch with no exchange).",
              file=sys.stderr)
        sys.exit(2)

    import cubie_amd as ca
    work = WORKLOADS[args.workload](args.n)
    solver = ca.Solver(work["system"], **work["solver_kw"])
    if args.compile_only:
        solver.kernel.compile_only(work["duration"])
        return

    import torch
    kernel = solver.kernel
    inits = torch.from_numpy(work["inits"].astype(kernel.precision)).cuda()
    params = torch.from_numpy(work["params"].astype(kernel.precision)).cuda()
    dtype = "f64" if kernel.precision is np.float64 else "f32"

    def launch():
        return kernel.run(inits, params, duration=work["duration"], timed=True)

    for _ in range(args.warmup):
        out, ms, info = launch()
    torch.cuda.synchronize()
    times = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out, ms, info = launch()
        times.append(ms)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ic = out.iteration_counters.cpu().numpy()
    steps_total = int(ic[:, 2, :].astype(np.int64).sum())
    accepted = int((ic[:, 2, :].astype(np.int64) - ic[:, 3, :]).sum())
    status = out.status_codes.cpu().numpy()
    mean_ms = float(np.mean(times))
    value = accepted / (mean_ms * 1e-3)
    spec = info["spec"]
    fps = flops_per_attempted_step(kernel, spec)
    achieved = steps_total * fps / (mean_ms * 1e-3) / 1e12
    line = {
        "metric": "accepted steps/s",
        "value": value,
        "unit": "accepted steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mean_ms,
        "wall_ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic",
        "config": dict(work["config"], kernel_mode=kernel.kernel_mode,
                       precision_mode=kernel.precision_mode,
                       failed_runs=int((status != 0).sum()),
                       attempted_steps=steps_total, accepted_steps=accepted,
                       kernel_info=info["module"].info()),
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK[dtype],
                     "unit": "TFLOP/s", "frac": achieved / PEAK[dtype], "traffic": None,
                     "flops_per_attempted_step": fps},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(work, solver)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
