"""Dense stage prediction for implicit Runge-Kutta steps (host side).

The converged stage increments of a step are h * f at the stage times, i.e.
samples of the derivative curve. Reading the Lagrange interpolant through
those samples ahead at the next step's stage times (and rescaling by the
step-size ratio r = h_next / h_prev) gives the next Newton starting guess.
Every entry of that read-ahead matrix is a polynomial in r with no constant
term, so the matrix is stored as per-power coefficient matrices C_p with
P(r) = sum_p C_p r^(p+1); the device evaluates the polynomials in Horner
form and applies P(r) per state (csrc/device/dirk.hpp).

Same construction, preconditions and defaults as the reference's
DenseStagePredictor (integrators/stage_predictors.py:83-300): repeated stage
nodes contribute one sample (the last stage at that node), a tableau whose
equal-step read-ahead amplifies sample error more than 10x the ideal
Gauss-Legendre spread does not predict, an explicit first stage is not
predicted, and prediction applies only after an accepted step whose ratio is
within the tableau's calibrated ceiling for the precision.
"""

from __future__ import annotations

from typing import Tuple

import numpy as np
from numpy.polynomial import polynomial as P
from numpy.polynomial.legendre import leggauss

MAX_NODE_AMPLIFICATION_RATIO = 10.0


def sample_stages(c) -> Tuple[int, ...]:
    """Stage sampled at each distinct node (last stage sharing the node)."""
    last = {}
    for stage, node in enumerate(c):
        last[node] = stage
    return tuple(last.values())


def source_stages(c) -> Tuple[int, ...]:
    """History row each stage's starting guess reads (first stage at its node)."""
    latest, out = {}, []
    for stage, node in enumerate(c):
        out.append(latest.get(node, stage))
        latest[node] = stage
    return tuple(out)


def readahead_matrix(nodes, ratio: float, samples) -> np.ndarray:
    nodes = np.asarray(nodes, dtype=np.float64)
    s = nodes.size
    dn = np.asarray([nodes[i] for i in samples])
    out = np.zeros((s, s))
    for ti, target in enumerate(1.0 + ratio * nodes):
        for bi, bnode in enumerate(dn):
            w = ratio
            for oi, onode in enumerate(dn):
                if oi != bi:
                    w *= (target - onode) / (bnode - onode)
            out[ti, samples[bi]] = w
    return out


def supports_prediction(tableau) -> bool:
    nodes = np.asarray(tableau.c, dtype=np.float64)
    samples = sample_stages(tableau.c)
    k = len(samples)
    amp = float(np.abs(readahead_matrix(nodes, 1.0, samples)).sum(axis=1).max())
    g = 0.5 * (leggauss(k)[0] + 1.0)
    ref = float(np.abs(readahead_matrix(g, 1.0, tuple(range(k)))).sum(axis=1).max())
    return amp <= MAX_NODE_AMPLIFICATION_RATIO * ref


def ratio_coefficients(tableau) -> np.ndarray:
    """(powers, stages, stages) with P(r) = sum_p C[p] r^(p+1)."""
    nodes = np.asarray(tableau.c, dtype=np.float64)
    s = nodes.size
    samples = sample_stages(tableau.c)
    dn = [nodes[i] for i in samples]
    k = len(dn)
    shift = P.Polynomial([1.0, 1.0])
    weights = []
    for bi in range(k):
        others = [dn[j] for j in range(k) if j != bi]
        coef = P.polyfromroots(others)
        coef = coef / P.polyval(dn[bi], coef)
        shifted = P.Polynomial(coef)(shift)
        c = np.zeros(k)
        c[: shifted.coef.size] = shifted.coef
        weights.append(c)
    out = np.zeros((k, s, s))
    for p in range(k):
        for ti in range(s):
            npow = nodes[ti] ** p
            for bi in range(k):
                out[p, ti, samples[bi]] = weights[bi][p] * npow
    return out


def dense_prediction_enabled(tableau, precision, attempt: bool = True) -> bool:
    return bool(attempt and tableau.dense_prediction_ratio_limit(precision) > 0.0
                and supports_prediction(tableau))
