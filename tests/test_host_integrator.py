"""The device integrator's per-ray loop, run on the host (prr_debug_trace_rays,
prr_debug_trace_log; CPU only), against the oracle's loop on the same
initial states.

For the 8-component metrics both integrate the same equations in affine
parameter with the same controller (accept when errn <= 1 or h == hmin,
h <- clamp(0.9 h errn^-1/5, hmin, hmax), kernels/solver.cuh:175-188), the same
located escape (D21) and the same wraps, so they take the same steps: equal
accepted and attempted step counts, except where the rounding of an error
estimate that sits at the accept threshold flips one decision.
"""
import math

import numpy as np
import pytest

import oracle as O
from parrelray_amd import _abi as A
from parrelray_amd import api
from parrelray_amd import configs as CF

RNG = np.random.default_rng(4242)


def states(metric, n, a=0.0):
    out = []
    for _ in range(n):
        pos = RNG.normal(size=3)
        pos = pos / np.linalg.norm(pos) * RNG.uniform(8.0, 30.0)
        d = -pos / np.linalg.norm(pos) + 0.6 * RNG.normal(size=3)
        d /= np.linalg.norm(d)
        x, u = O.init_ray(metric, 1.0, a, pos, d)
        out.append(np.concatenate([x, u]))
    return np.array(out)


def frame():
    pos, al, be = CF.orbit_pose(20.0, 80.0)
    return A.make_frame(pos, al, be, math.radians(50.0), 8, 8, mass=1.0, spin=0.0)


@pytest.mark.parametrize("metric", [A.PRR_SCHW_KS_CART, A.PRR_SCHW_SPH])
@pytest.mark.parametrize("solver", [A.PRR_RKF45, A.PRR_EULER])
def test_host_loop_takes_the_oracles_steps(metric, solver):
    if solver == A.PRR_RKF45:
        T = A.make_tracer(solver=solver, h=0.05, hmin=1e-9, hmax=5.0, atol=1e-10, rtol=1e-10, r_escape=60.0,
                          horizon_eps_rel=1e-3, horizon_eps_abs=0.0, lambda_max=1e9, max_steps=20000,
                          max_attempts=200000)
    else:
        T = A.make_tracer(solver=solver, h=0.01, r_escape=60.0, horizon_eps_rel=1e-3, horizon_eps_abs=0.0,
                          lambda_max=1e9, max_steps=12000, max_attempts=12000)
    x0 = states(metric, 48)
    h = api.trace_rays_host(metric, T, frame(), x0u0=x0)
    o = O.Scene(metric, 1.0, 0.0, T).trace(x0)
    assert np.array_equal(h["cls"], o["cls"])
    # adaptive: the RHS forms differ in rounding (factored KS form on the
    # device, the reference's expanded Phi form in the oracle), so an error
    # estimate at the threshold can flip one decision, by one step
    same_steps = (h["steps"] == o["steps"]) & (h["attempts"] == o["attempts"])
    assert same_steps.mean() >= (1.0 if solver == A.PRR_EULER else 0.8), (h["steps"] - o["steps"]).tolist()
    assert np.abs(h["steps"] - o["steps"]).max() <= (0 if solver == A.PRR_EULER else 2)
    assert np.abs(h["attempts"] - o["attempts"]).max() <= (0 if solver == A.PRR_EULER else 3)
    esc = o["cls"] == A.PRR_ESCAPE
    assert esc.sum() >= 10 and (o["cls"] == A.PRR_HORIZON).sum() >= 5
    err = np.abs(h["xend"][esc] - o["xend"][esc]).max(axis=1) / (np.abs(o["xend"][esc]).max(axis=1) + 1)
    assert np.median(err) < (1e-12 if solver == A.PRR_EULER else 1e-9), np.sort(err)[-5:]


def test_step_log_follows_the_controller():
    """One RKF45 ray's accepted steps (prr_debug_trace_log): one row per
    accepted step, attempts never fall behind, every proposal lies in
    [hmin, hmax], and the 3 M initial step is rejected: the first accepted
    step is shorter than 0.9 h0 (the controller's safety factor times
    errn^(-1/5) < 1)."""
    T = A.make_tracer(solver=A.PRR_RKF45, h=3.0, hmin=1e-9, hmax=3.0, atol=1e-12, rtol=1e-12, r_escape=60.0,
                      horizon_eps_rel=1e-3, horizon_eps_abs=0.0, lambda_max=1e9, max_steps=20000,
                      max_attempts=200000)
    for ray in (10, 27, 36):
        log = api.trace_log(A.PRR_SCHW_SPH, T, frame(), ray=ray)
        assert log.shape[0] > 100
        steps, att, hprop = log[:, 0], log[:, 1], log[:, 2]
        assert np.array_equal(steps, np.arange(1, log.shape[0] + 1))
        assert np.all(np.diff(att) >= 1) and np.all(att >= steps)
        assert np.all((hprop >= T.hmin) & (hprop <= T.hmax * (1 + 1e-12)))
        assert att[0] >= 2 and hprop[0] < 0.9 * T.h
        # the host log's totals are the loop's own
        h = api.trace_rays_host(A.PRR_SCHW_SPH, T, frame(), ids=[ray])
        assert h["steps"][0] == log.shape[0] and h["attempts"][0] >= att[-1]
