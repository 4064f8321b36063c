"""Discrete adjoint on the GPU (explicit RK and Rosenbrock-W).

Pinned three ways: a closed form (a linear system with fixed steps, where
dL/dy0 = (R^N)^T c with R the method's one-step matrix, built here in NumPy
from the tableau), central finite differences of the GPU forward solve at
fixed steps (where the discrete adjoint is the exact derivative of the map),
and central differences of the adaptive solve (where the frozen-step adjoint
differs from the derivative of the adaptive map by about the tolerance)."""

import numpy as np
import pytest

import cubie_amd as ca
from cubie_amd.batchsolving import adjoint as adj
from cubie_amd.integrators.algorithms import resolve_algorithm
from systems import hodgkin_huxley, hh_driver, hh_batch

pytestmark = pytest.mark.gpu

LINEAR = ["dx0 = p0*x0 + x1", "dx1 = -x0 + p1*x1", "dx2 = x1 - p2*x2"]


def linear_system():
    return ca.create_ODE_system(LINEAR, states={"x0": 1.0, "x1": 0.0, "x2": 0.5},
                                parameters={"p0": -0.5, "p1": -0.2, "p2": 1.0},
                                precision=np.float64, name="linear_adj")


def linear_matrix(p):
    return np.array([[p[0], 1.0, 0.0], [-1.0, p[1], 0.0], [0.0, 1.0, -p[2]]])


def one_step_matrix(alg, M, h):
    """R with y_{k+1} = R y_k for y' = M y, from the tableau (NumPy)."""
    tab = resolve_algorithm(alg).tableau
    A, b = np.array(tab.a, dtype=float), np.array(tab.b, dtype=float)
    s, n = len(b), M.shape[0]
    R = np.zeros((n, n))
    for col in range(n):
        y = np.eye(n)[:, col]
        K = np.zeros((s, n))
        for i in range(s):
            K[i] = M @ (y + h * (A[i, :i] @ K[:i]))
        R[:, col] = y + h * (b @ K)
    return R


@pytest.mark.parametrize("alg", ["dormand-prince-54", "rk4"])
def test_erk_adjoint_linear_closed_form(alg):
    n_runs = 64
    rng = np.random.default_rng(61)
    params = np.array([-0.5, -0.2, 1.0])[:, None] * rng.uniform(0.8, 1.2, (3, n_runs))
    inits = rng.uniform(-1, 1, (3, n_runs))
    h, steps = 2.0 ** -7, 128     # binary step: t lands exactly on t_end
    solver = ca.Solver(linear_system(), algorithm=alg, step_controller="fixed", dt=h,
                       output_types=["state"])
    rec = adj.record(solver.kernel, inits, params, duration=h * steps)
    assert int(rec.outputs.status_codes.abs().sum()) == 0
    c = rng.normal(size=(3, n_runs))
    g = np.zeros(tuple(rec.outputs.state.shape))
    g[-1] = c                                  # L = sum_runs c . y(t_end)
    gi, gp, _ = adj.backward(rec, grad_state=__import__("torch").from_numpy(g))
    gi = gi.cpu().numpy()
    worst = 0.0
    for r in range(n_runs):
        R = one_step_matrix(alg, linear_matrix(params[:, r]), h)
        expect = np.linalg.matrix_power(R, steps).T @ c[:, r]
        worst = max(worst, np.abs(gi[:, r] - expect).max() / np.abs(expect).max())
    print(alg, "closed-form dL/dy0 rel err", worst)
    assert worst <= 1e-12


def _fd_batch(inits, params, rel=1e-6):
    """Central-difference batch: for every run and every input (states then
    parameters) two perturbed copies; returns the batch and the step sizes."""
    n, p, N = inits.shape[0], params.shape[0], inits.shape[1]
    X = np.concatenate([inits, params])
    m = n + p
    Xb = np.repeat(X[:, :, None, None], m, axis=2).repeat(2, axis=3)   # (m, N, m, 2)
    steps = rel * np.maximum(np.abs(X), 1e-3)                          # (m, N)
    for k in range(m):
        Xb[k, :, k, 0] += steps[k]
        Xb[k, :, k, 1] -= steps[k]
    Xb = Xb.reshape(m, -1)
    return Xb[:n], Xb[n:], steps


def _loss_terms(state, summ, c_state, c_mean):
    """Per-run loss: c_state . final saved state + c_mean * mean summaries."""
    val = (c_state * state[-1]).sum(axis=0)
    if summ is not None:
        val = val + (c_mean * summ).sum(axis=(0, 1))
    return val


def _hh_solver(alg, fixed, tol=1e-8):
    kw = dict(algorithm=alg, save_every=None, summarise_every=20.0, sample_summaries_every=0.5,
              output_types=["state", "mean"], summarise_variables=["V", "m"])
    if fixed:
        kw.update(step_controller="fixed", dt=2.0 ** -7)
    else:
        kw.update(atol=tol, rtol=tol)
    if alg == "ros3p":
        kw.update(linear_correction_type="bicgstab")
    s = ca.Solver(hodgkin_huxley(np.float64), **kw)
    s._configure_drivers(hh_driver())
    return s


@pytest.mark.parametrize("alg,fixed,bar", [
    ("dormand-prince-54", True, 1e-5),
    ("ros3p", True, 1e-5),
    ("dormand-prince-54", False, 1e-4),
    # ROS3P's second-order error estimate makes its step sizes depend more
    # strongly on the inputs; the frozen-step gradient differs from the
    # derivative of the adaptive map by more than for DOPRI5
    ("ros3p", False, 1e-3),
])
def test_hh_adjoint_vs_central_differences(alg, fixed, bar):
    """dL/d(y0, g) against central differences of the GPU's own FP64 forward
    solve (relative step 1e-6), HH with a sampled driver over 40 ms."""
    import torch
    N = 8
    inits, params = hh_batch(N, seed=62)
    solver = _hh_solver(alg, fixed)
    rng = np.random.default_rng(63)
    rec = adj.record(solver.kernel, inits, params, duration=40.0)
    out = rec.outputs
    assert int(out.status_codes.abs().sum()) == 0
    state = out.state.cpu().numpy()
    summ = out.state_summaries.cpu().numpy()
    c_state = rng.normal(size=state.shape[1:2] + (1,))
    c_mean = rng.normal(size=(1,) + summ.shape[1:2] + (1,))
    gs = np.zeros_like(state)
    gs[-1] = c_state
    gss = np.broadcast_to(c_mean, summ.shape).copy()
    gi, gp, _ = adj.backward(rec, torch.from_numpy(gs), torch.from_numpy(gss))
    grad = np.concatenate([gi.cpu().numpy(), gp.cpu().numpy()])          # (m, N)

    xi, xp, steps = _fd_batch(inits, params)
    res = solver.solve(xi, xp, duration=40.0, nan_error_trajectories=False)
    assert np.all(res.status_codes == 0)
    L = _loss_terms(res.state, res.state_summaries, c_state, c_mean)
    L = L.reshape(N, -1, 2)                                             # (N, m, 2)
    fd = ((L[:, :, 0] - L[:, :, 1]) / (2 * steps.T)).T                  # (m, N)
    err = np.abs(grad - fd).max(axis=0) / np.abs(fd).max(axis=0)
    print(alg, "fixed" if fixed else "adaptive", "adjoint vs FD rel err per run", err)
    assert err.max() <= bar


def test_solve_adjoint_autograd_matches_backward():
    """torch.autograd through solve_adjoint gives what backward() gives."""
    import torch
    N = 16
    inits, params = hh_batch(N, seed=64)
    solver = _hh_solver("dormand-prince-54", True)
    rec = adj.record(solver.kernel, inits, params, duration=20.0)
    x0 = torch.tensor(inits, dtype=torch.float64, device="cuda", requires_grad=True)
    p = torch.tensor(params, dtype=torch.float64, device="cuda", requires_grad=True)
    state, summ, status = ca.solve_adjoint(solver, x0, p, duration=20.0)
    assert int(status.abs().sum()) == 0
    target = torch.linspace(-70.0, -60.0, N, dtype=torch.float64, device="cuda")
    loss = ((summ[:, 0, :] - target) ** 2).sum() + (state[-1, 1, :] ** 2).sum()
    loss.backward()
    gs = torch.zeros_like(rec.outputs.state)
    gs[-1, 1, :] = 2 * rec.outputs.state[-1, 1, :]
    gss = torch.zeros_like(rec.outputs.state_summaries)
    gss[:, 0, :] = 2 * (rec.outputs.state_summaries[:, 0, :] - target)
    gi, gp, _ = adj.backward(rec, gs, gss)
    assert torch.allclose(x0.grad, gi, rtol=1e-12, atol=0)
    assert torch.allclose(p.grad, gp, rtol=1e-12, atol=0)


def test_adjoint_rejects_nondifferentiable_summary():
    import torch
    s = ca.Solver(hodgkin_huxley(np.float64), algorithm="dormand-prince-54",
                  step_controller="fixed", dt=2.0 ** -7, summarise_every=5.0,
                  sample_summaries_every=0.5, output_types=["mean", "peaks[2]"],
                  summarise_variables=["V"])
    s._configure_drivers(hh_driver())
    inits, params = hh_batch(4)
    rec = adj.record(s.kernel, inits, params, duration=10.0)
    g = torch.ones_like(rec.outputs.state_summaries)
    with pytest.raises(NotImplementedError):
        adj.backward(rec, None, g)


def _rows_and_extrema_case(precision, save_every=2.0, summarise_every=5.0):
    kw = dict(algorithm="dormand-prince-54", step_controller="fixed", dt=2.0 ** -7,
              save_every=save_every, summarise_every=summarise_every, sample_summaries_every=0.25,
              output_types=["state", "mean", "max", "rms"], summarise_variables=["V", "h"])
    s = ca.Solver(hodgkin_huxley(precision), **kw)
    s._configure_drivers(hh_driver())
    return s


def _rows_and_extrema_grad(s, inits, params, c_state, c_summ, duration=10.0):
    import torch
    prec = s.kernel.precision
    rec = adj.record(s.kernel, inits.astype(prec), params.astype(prec), duration=duration)
    out = rec.outputs
    assert int(out.status_codes.abs().sum()) == 0
    gs = np.broadcast_to(c_state, tuple(out.state.shape)).astype(prec).copy()
    gss = np.broadcast_to(c_summ, tuple(out.state_summaries.shape)).astype(prec).copy()
    gi, gp, _ = adj.backward(rec, torch.from_numpy(gs), torch.from_numpy(gss))
    return np.concatenate([gi.cpu().numpy(), gp.cpu().numpy()]).astype(np.float64), out


def test_adjoint_saved_rows_and_max_rms_summaries():
    """A loss on several saved rows and on the max / rms / mean summaries of
    two states, DOPRI5 at fixed steps, FP64: the adjoint gradient against
    central differences of the forward solve."""
    N = 8
    inits, params = hh_batch(N, seed=65)
    s = _rows_and_extrema_case(np.float64)
    rng = np.random.default_rng(66)
    c_state = rng.normal(size=(6, 4, 1))
    c_summ = rng.normal(size=(2, 6, 1))
    grad, out = _rows_and_extrema_grad(s, inits, params, c_state, c_summ)
    assert tuple(out.state.shape)[:2] == (6, 4) and tuple(out.state_summaries.shape)[:2] == (2, 6)
    xi, xp, steps = _fd_batch(inits, params)
    res = s.solve(xi, xp, duration=10.0, nan_error_trajectories=False)
    L = (c_state * res.state).sum(axis=(0, 1)) + (c_summ * res.state_summaries).sum(axis=(0, 1))
    L = L.reshape(N, -1, 2)
    fd = ((L[:, :, 0] - L[:, :, 1]) / (2 * steps.T)).T
    err = np.abs(grad - fd).max(axis=0) / np.abs(fd).max(axis=0)
    print("rows + max/rms/mean summaries: adjoint vs FD", err)
    assert err.max() <= 1e-5


def test_adjoint_fp32_matches_fp64_adjoint():
    """The FP32 adjoint (FP32 tape, replay and step adjoints) against the FP64
    adjoint of the same problem on the same FP32-representable inputs, over
    the first 3 ms, while every run is still below -55 mV: saved rows every
    1 ms and the mean / max / rms of V and h over two 1.5 ms windows. Max
    gradients land on the sample the FP32 forward pass kept, which the tape
    records. (Near and across a spike the derivative of V with respect to
    the conductances is set by the spike time, which the FP32 forward
    solution itself shifts: a run of this batch that starts its upstroke at
    4 ms shows 9e-3 between the two adjoints at 4 ms, and over 10 ms a run
    that spikes at ~7 ms differs from central differences of the FP64
    solve by 2e-2 on the saved rows and 0.17 on the rms of V in FP32, where
    the FP64 adjoint matches them to 1e-7.)"""
    N = 8
    inits, params = hh_batch(N, seed=63)
    inits = inits.astype(np.float32).astype(np.float64)
    params = params.astype(np.float32).astype(np.float64)
    rng = np.random.default_rng(68)
    c_state = rng.normal(size=(4, 4, 1))
    c_summ = rng.normal(size=(2, 6, 1))
    g = {}
    for prec in (np.float32, np.float64):
        s = _rows_and_extrema_case(prec, save_every=1.0, summarise_every=1.5)
        g[prec], out = _rows_and_extrema_grad(s, inits, params, c_state, c_summ, duration=3.0)
        assert tuple(out.state.shape)[:2] == (4, 4) and tuple(out.state_summaries.shape)[:2] == (2, 6)
        assert float(out.state_summaries[:, 1].max()) < -55.0     # max V of every window
    err = np.abs(g[np.float32] - g[np.float64]).max(axis=0) / np.abs(g[np.float64]).max(axis=0)
    print("fp32 adjoint vs fp64 adjoint", err)
    assert err.max() <= 1e-3
