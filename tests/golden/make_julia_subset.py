"""Write tests/golden/julia_lorenz_subset.npz from the reference's vendored
Julia data (tests/integrated_numerical_tests/julia_reference/data/
julia_reference_ne.npz, DifferentialEquations.jl, see its README).

Run once in a container where the reference checkout is present; the
subset (golden FP64 Vern9 finals, the rho grid, and the Float32 finals at
the gate's pinned dt / tolerance for the algorithms this engine tests) is
committed so the checks never need the reference at run time.
"""

import os
import sys

import numpy as np

SRC = "/root/reference/tests/integrated_numerical_tests/julia_reference/data/julia_reference_ne.npz"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "julia_lorenz_subset.npz")

ALIASES = ["dormand-prince-54", "bogacki-shampine-32", "classical-rk4", "heun-21",
           "cash-karp-54", "fehlberg-45", "tsit5", "dormand-prince-853", "vern7",
           "ros3p", "radau_iia_5", "kvaerno3", "kvaerno5", "l_stable_sdirk_4",
           "rodas3p", "rosenbrock23_sciml", "sdirk_2_2", "trapezoidal_dirk",
           "implicit_midpoint", "backwards_euler", "crank_nicolson", "euler"]

# Gate protocol constants (ne_gate.py): floor / cap relative to the golden scale.
FLOOR_REL, PIN_CAP_REL, PIN_MARGIN_MULT = 4e-6, 3e-2, 4.0
DTS = [2.0 ** -k for k in range(1, 14)]
TOLS = [10.0 ** -k for k in range(2, 7)]


def err(final, golden):
    return float(np.sqrt(np.mean((final.astype(np.float64) - golden) ** 2)))


def main():
    if not os.path.exists(SRC):
        sys.exit("reference data not available here")
    d = np.load(SRC)
    golden = d["golden_states"].astype(np.float64)
    scale = float(np.sqrt(np.mean(golden ** 2)))
    floor, cap = FLOOR_REL * scale, PIN_CAP_REL * scale
    out = {"golden_rho": d["golden_rho"], "golden_states": golden}
    for alias in ALIASES:
        key = f"fixed_{alias}_dts"
        if key in d.files:
            dts = d[key]
            fin = d[f"fixed_{alias}_finals"]
            errs = {float(dt): err(fin[i], golden) for i, dt in enumerate(dts)}
            pin = None
            for dt in DTS:
                e = errs.get(dt)
                if e is not None and np.isfinite(e) and PIN_MARGIN_MULT * floor <= e < cap:
                    pin = dt
                elif pin is not None:
                    break
            if pin is not None:
                i = [float(x) for x in dts].index(pin)
                out[f"fixed_{alias}_pin"] = np.array(pin)
                out[f"fixed_{alias}_final"] = fin[i]
        key = f"adaptive_{alias}_tols"
        if key in d.files:
            tols = d[key]
            fin = d[f"adaptive_{alias}_finals"]
            errs = {float(t): err(fin[i], golden) for i, t in enumerate(tols)}
            ok = [t for t in TOLS if t in errs and np.isfinite(errs[t])
                  and PIN_MARGIN_MULT * floor <= errs[t] < cap]
            if ok:
                pin = min(ok)
                i = [float(x) for x in tols].index(pin)
                out[f"adaptive_{alias}_pin"] = np.array(pin)
                out[f"adaptive_{alias}_final"] = fin[i]
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, len(out), "arrays")


if __name__ == "__main__":
    main()
