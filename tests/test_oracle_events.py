"""Pins of the oracle's located events (CPU, no GPU): the escape endpoint
at the r = r_escape crossing (D21) and sphere hits located on the integrated
path (D22), against closed-form intersections in flat space and against an
independent event search (scipy's DOP853 with root-finding events) on the
oracle's own Schwarzschild right-hand side, which the Christoffel tests pin.

Reference behaviour replaced: the RKF45 loop stops at the first state past
r_escape (metrics/schwarzschild.cuh:66-79) and tests spheres on step chords
(main/Spacetime.h:38-46, scene/sphere.h:34-50).
"""
import math

import numpy as np
from scipy import integrate

import oracle as O
from parrelray_amd import _abi as A

RNG = np.random.default_rng(77)


def tracer(**kw):
    base = dict(solver=A.PRR_RKF45, h=0.01, hmin=1e-9, hmax=2.0, atol=1e-12, rtol=1e-12, r_escape=60.0,
                horizon_eps_rel=1e-3, horizon_eps_abs=0.0, lambda_max=1e9, max_steps=10**7, max_attempts=10**8)
    base.update(kw)
    return A.make_tracer(**base)


def states(kind, M, pos, dirs):
    rows = []
    for d in dirs:
        x, u = O.init_ray(kind, M, 0.0, pos, d)
        rows.append(np.concatenate([x, u]))
    return np.array(rows)


def unit(v):
    return v / np.linalg.norm(v)


def test_flat_escape_is_located_on_the_escape_sphere():
    R = 60.0
    S = O.Scene(A.PRR_FLAT, 0.0, 0.0, tracer(r_escape=R))
    pos = np.array([12.0, -5.0, 3.0])
    dirs = [unit(RNG.normal(size=3)) for _ in range(50)]
    out = S.trace(states(A.PRR_FLAT, 0.0, pos, dirs))
    assert np.all(out["cls"] == A.PRR_ESCAPE)
    for d, xe in zip(dirs, out["xend"]):
        # the exit point of the straight line pos + t d through |x| = R
        b = pos @ d
        t = -b + math.sqrt(b * b - (pos @ pos - R * R))
        assert np.abs(xe[:3] - (pos + t * d)).max() < 1e-9 * R


def test_flat_sphere_hits_are_the_analytic_entry_points():
    c, Rs = np.array([8.0, 1.0, 0.5]), 1.0
    S = O.Scene(A.PRR_FLAT, 0.0, 0.0, tracer())
    S.set_spheres([(c, Rs, 1)])
    pos = np.array([20.0, 3.0, -1.0])
    n_hit = 0
    dirs = []
    for _ in range(200):
        target = c + RNG.normal(size=3) * 0.9
        dirs.append(unit(target - pos))
    out = S.trace(states(A.PRR_FLAT, 0.0, pos, dirs))
    for d, cls, xe in zip(dirs, out["cls"], out["xend"]):
        m = pos - c
        b = m @ d
        disc = b * b - (m @ m - Rs * Rs)
        if abs(disc) < 1e-6:  # grazing: the class is decided within rounding
            continue
        if disc < 0:
            assert cls != A.PRR_SPHERE
            continue
        t = -b - math.sqrt(disc)
        assert cls == A.PRR_SPHERE
        assert np.abs(xe[:3] - (pos + t * d)).max() < 1e-9
        n_hit += 1
    assert n_hit > 50


def ks_rhs(lam, y):
    x, u = y[:4], y[4:]
    return np.concatenate([u, O.geod(A.PRR_SCHW_KS_CART, 1.0, 0.0, x, u)])


def scipy_first_event(y0, g, lam_max=400.0):
    ev = lambda lam, y: g(y)  # noqa: E731
    ev.terminal = True
    ev.direction = -1
    sol = integrate.solve_ivp(ks_rhs, (0.0, lam_max), y0, method="DOP853", rtol=1e-12, atol=1e-12,
                              events=ev, max_step=0.05)
    if sol.t_events[0].size == 0:
        return None
    return sol.y_events[0][0]


def test_schwarzschild_sphere_hits_match_an_independent_event_search():
    c, Rs = np.array([8.0, 0.0, 0.5]), 1.0
    S = O.Scene(A.PRR_SCHW_KS_CART, 1.0, 0.0, tracer())
    S.set_spheres([(c, Rs, 1)])
    pos = np.array([20.0, 3.0, 1.0])
    dirs = [unit(c + RNG.normal(size=3) * 1.0 - pos) for _ in range(40)]
    y0s = states(A.PRR_SCHW_KS_CART, 1.0, pos, dirs)
    out = S.trace(y0s)
    checked = 0
    for y0, cls, xe in zip(y0s, out["cls"], out["xend"]):
        hit = scipy_first_event(y0, lambda y: (y[:3] - c) @ (y[:3] - c) - Rs * Rs)
        if hit is None:
            assert cls != A.PRR_SPHERE
            continue
        # a path that only grazes the sphere is decided within the tolerances
        if abs(np.linalg.norm(hit[:3] - c) - Rs) > 1e-6:
            continue
        assert cls == A.PRR_SPHERE
        assert abs(np.linalg.norm(xe[:3] - c) - Rs) < 1e-9
        assert np.abs(xe[:3] - hit[:3]).max() < 1e-6
        checked += 1
    assert checked >= 8


def test_schwarzschild_escape_matches_an_independent_event_search():
    R = 60.0
    S = O.Scene(A.PRR_SCHW_KS_CART, 1.0, 0.0, tracer(r_escape=R))
    pos = np.array([15.0, 2.0, -1.0])
    # impact parameters above the 3 sqrt(3) M capture threshold: deflected, escaping
    dirs = []
    for _ in range(12):
        d = unit(-pos + np.cross(pos, unit(RNG.normal(size=3))) * RNG.uniform(0.45, 1.5))
        dirs.append(d)
    y0s = states(A.PRR_SCHW_KS_CART, 1.0, pos, dirs)
    out = S.trace(y0s)
    checked = 0
    for y0, cls, xe in zip(y0s, out["cls"], out["xend"]):
        if cls != A.PRR_ESCAPE:
            continue
        ref = scipy_first_event(y0, lambda y: R * R - y[:3] @ y[:3], lam_max=2000.0)
        assert ref is not None
        assert abs(np.linalg.norm(xe[:3]) - R) < 1e-9 * R
        assert np.abs(xe[:3] - ref[:3]).max() < 1e-6 * R
        checked += 1
    assert checked >= 6
