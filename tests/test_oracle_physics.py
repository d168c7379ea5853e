"""Pins the CPU oracle's physics against closed forms and invariants.

Each check is chosen so that a plausible mistake (a dropped Christoffel term,
a sign, a transposed index, the wrong RKF coupling) fails it; none of them
re-types the oracle's own formula.
"""
import json
import math
import os

import numpy as np
import pytest
from scipy import integrate

import oracle as O
from parrelray_amd import _abi as A
from parrelray_amd import configs as CF

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
RNG = np.random.default_rng(20240531)


def fd_christoffel(kind, M, a, x, eps=1e-5):
    """Christoffel symbols from central differences of the oracle's metric."""
    g = O.metric(kind, M, a, x)
    gi = np.linalg.inv(g)
    dg = np.zeros((4, 4, 4))
    for k in range(4):
        xp, xm = np.array(x, float), np.array(x, float)
        step = eps * max(1.0, abs(x[k]))
        xp[k] += step
        xm[k] -= step
        dg[k] = (O.metric(kind, M, a, xp) - O.metric(kind, M, a, xm)) / (2 * step)
    G = np.zeros((4, 4, 4))
    for al in range(4):
        for mu in range(4):
            for nu in range(4):
                G[al, mu, nu] = 0.5 * sum(gi[al, be] * (dg[mu, be, nu] + dg[nu, be, mu] - dg[be, mu, nu])
                                          for be in range(4))
    return G


def random_points(kind, n):
    pts = []
    for _ in range(n):
        if kind in (A.PRR_SCHW_SPH, A.PRR_KERR_BL):
            x = [RNG.uniform(3.0, 40.0), RNG.uniform(0.3, math.pi - 0.3), RNG.uniform(0, 2 * math.pi), 0.0]
        elif kind == A.PRR_ELLIS:  # l on both sides of the throat
            x = [RNG.uniform(-10.0, 10.0), RNG.uniform(0.3, math.pi - 0.3), RNG.uniform(0, 2 * math.pi), 0.0]
        else:
            v = RNG.normal(size=3)
            x = list(v / np.linalg.norm(v) * RNG.uniform(3.0, 40.0)) + [0.0]
        u = list(RNG.normal(size=4))
        pts.append((np.array(x), np.array(u)))
    return pts


@pytest.mark.parametrize("kind,a", [(A.PRR_SCHW_KS_CART, 0.0), (A.PRR_SCHW_SPH, 0.0), (A.PRR_KERR_BL, 0.0),
                                    (A.PRR_KERR_BL, 0.6), (A.PRR_KERR_BL, 0.998), (A.PRR_SCHW_SPH, 0.8),
                                    (A.PRR_ELLIS, 0.5), (A.PRR_ELLIS, 2.0)])
def test_geod_equals_minus_christoffel_of_metric(kind, a):
    """a: Kerr's spin, the charge of the Schwarzschild-coordinate metric
    (Reissner-Nordstrom), or the Ellis wormhole's throat radius."""
    for x, u in random_points(kind, 20):
        G = fd_christoffel(kind, 1.0, a, x)
        expect = -np.einsum("amn,m,n->a", G, u, u)
        got = O.geod(kind, 1.0, a, x, u)
        scale = np.abs(np.einsum("amn,m,n->amn", np.abs(G), np.abs(u), np.abs(u))).sum(axis=(1, 2)) + 1e-12
        # FD noise is absolute in G; compare against each component's own scale
        # plus a small fraction of the largest one.
        assert np.all(np.abs(got - expect) <= 2e-7 * (scale + 1e-3 * scale.max())), (x, u, got, expect)


def test_flat_geod_is_zero():
    for x, u in random_points(A.PRR_FLAT, 5):
        assert np.all(O.geod(A.PRR_FLAT, 1.0, 0.0, x, u) == 0.0)


def test_kerr_metric_derivatives_match_finite_differences():
    for a in (0.3, 0.95):
        for x, _ in random_points(A.PRR_KERR_BL, 10):
            dg = O.dmetric_kerr(1.0, a, x)
            for k in (0, 1):
                e = 1e-6
                xp, xm = x.copy(), x.copy()
                xp[k] += e
                xm[k] -= e
                fd = (O.metric(A.PRR_KERR_BL, 1.0, a, xp) - O.metric(A.PRR_KERR_BL, 1.0, a, xm)) / (2 * e)
                assert np.allclose(dg[k], fd, rtol=1e-6, atol=1e-8)
            assert np.all(dg[2] == 0) and np.all(dg[3] == 0)


def test_kerr_reduces_to_schwarzschild_at_zero_spin():
    for x, u in random_points(A.PRR_SCHW_SPH, 20):
        assert np.allclose(O.geod(A.PRR_KERR_BL, 1.0, 0.0, x, u), O.geod(A.PRR_SCHW_SPH, 1.0, 0.0, x, u),
                           rtol=1e-12, atol=1e-14)


def test_kerr_metric_known_values():
    # Horizon r+ = M + sqrt(M^2 - a^2) is where Delta = 0: g_rr diverges.
    a = 0.95
    rp = 1 + math.sqrt(1 - a * a)
    g = O.metric(A.PRR_KERR_BL, 1.0, a, [rp * (1 + 1e-9), 1.0, 0.0, 0.0])
    assert g[0, 0] > 1e7
    # ergosphere boundary r = M + sqrt(M^2 - a^2 cos^2 th): g_tt = 0
    th = 0.7
    re = 1 + math.sqrt(1 - a * a * math.cos(th) ** 2)
    assert abs(O.metric(A.PRR_KERR_BL, 1.0, a, [re, th, 0.0, 0.0])[3, 3]) < 1e-13


@pytest.mark.parametrize("kind", [A.PRR_FLAT, A.PRR_SCHW_KS_CART, A.PRR_SCHW_SPH, A.PRR_KERR_BL, A.PRR_ELLIS])
def test_initial_state_is_null(kind):
    for _ in range(20):
        pos = RNG.normal(size=3)
        pos = pos / np.linalg.norm(pos) * RNG.uniform(5, 50)
        d = RNG.normal(size=3)
        d /= np.linalg.norm(d)
        a = 0.9 if kind == A.PRR_KERR_BL else (2.0 if kind == A.PRR_ELLIS else 0.0)
        x, u = O.init_ray(kind, 1.0, a, pos, d)
        g = O.metric(kind, 1.0, a, x)
        norm = u @ g @ u
        assert abs(norm) < 1e-12 * (np.abs(g).max() * (u @ u)), norm
        assert u[3] > 0  # future directed


def test_radial_ingoing_ks_has_unit_time_component():
    # In Kerr-Schild coordinates an ingoing radial null ray has dt/dlambda = dr/dlambda magnitude.
    pos = np.array([7.0, 0.0, 0.0])
    x, u = O.init_ray(A.PRR_SCHW_KS_CART, 1.0, 0.0, pos, [-1.0, 0.0, 0.0])
    assert abs(u[3] - 1.0) < 1e-14


def tracer(**kw):
    base = dict(solver=A.PRR_RKF45, h=0.01, hmin=1e-9, hmax=5.0, atol=1e-12, rtol=1e-12, r_escape=200.0,
                horizon_eps_rel=1e-3, horizon_eps_abs=0.0, lambda_max=1e9, max_steps=10**7, max_attempts=10**8)
    base.update(kw)
    return A.make_tracer(**base)


def kerr_state_from_constants(M, a, r, th, E, L, ur, uth):
    """u^mu of a ray at (r, th) with p_t = -E, p_phi = L and given u^r, u^th."""
    g = O.metric(A.PRR_KERR_BL, M, a, [r, th, 0.0, 0.0])
    gi = np.linalg.inv(g)
    ut = gi[3, 3] * (-E) + gi[3, 2] * L
    uph = gi[2, 3] * (-E) + gi[2, 2] * L
    return np.array([r, th, 0.0, 0.0]), np.array([ur, uth, uph, ut])


def test_kerr_conserves_E_L_Q():
    a = 0.9
    S = O.Scene(A.PRR_KERR_BL, 1.0, a, tracer(max_steps=3000, r_escape=60.0))
    for _ in range(6):
        pos = RNG.normal(size=3)
        pos = pos / np.linalg.norm(pos) * 20.0
        d = -pos / 20.0 + 0.25 * RNG.normal(size=3)
        d /= np.linalg.norm(d)
        x, u = O.init_ray(A.PRR_KERR_BL, 1.0, a, pos, d)
        c0 = O.kerr_constants(1.0, a, x, u)
        res = S.trace(np.concatenate([x, u])[None])
        xe, ue = res["xend"][0, :4], res["xend"][0, 4:]
        c1 = O.kerr_constants(1.0, a, xe, ue)
        assert res["cls"][0] in (A.PRR_ESCAPE, A.PRR_HORIZON, A.PRR_STEP_CAP)
        if res["cls"][0] == A.PRR_HORIZON:
            continue  # BL constants lose precision at the horizon
        scale = np.array([abs(c0[0]), abs(c0[1]) + 1.0, abs(c0[2]) + 1.0])
        assert np.all(np.abs(c1 - c0) <= 1e-8 * scale), (c0, c1)


def test_kerr_equatorial_photon_orbits():
    # Circular equatorial photon orbits r = 2M{1 + cos[2/3 arccos(-+a/M)]} (Bardeen, Press, Teukolsky 1972).
    a = 0.7
    for sign in (+1, -1):  # prograde, retrograde
        r = 2 * (1 + math.cos(2.0 / 3.0 * math.acos(-sign * a)))
        xi = -(r ** 3 - 3 * r ** 2 + a * a * r + a * a) / (a * (r - 1))  # L/E
        x, u = kerr_state_from_constants(1.0, a, r, math.pi / 2, 1.0, xi, 0.0, 0.0)
        g = O.metric(A.PRR_KERR_BL, 1.0, a, x)
        assert abs(u @ g @ u) < 1e-12
        du = O.geod(A.PRR_KERR_BL, 1.0, a, x, u)
        assert abs(du[0]) < 1e-12 * (abs(u[3]) ** 2)  # no radial acceleration
        S = O.Scene(A.PRR_KERR_BL, 1.0, a, tracer(max_steps=400, hmax=0.05))
        res = S.trace(np.concatenate([x, u])[None])
        assert abs(res["xend"][0, 0] - r) < 1e-5 * r
        # a slightly smaller radius falls in, a slightly larger one escapes (unstable orbit)
    r_isco_ph = 2 * (1 + math.cos(2.0 / 3.0 * math.acos(-a)))
    assert r_isco_ph < 3.0


def spherical_photon_orbit(a, r, M=1.0):
    """Constants of the Kerr spherical photon orbit at Boyer-Lindquist radius r
    (Teo 2003, with E = 1): xi = L/E, eta = Q/E^2; the orbits exist where
    eta >= 0, between the prograde and retrograde equatorial photon orbits."""
    xi = (r * r * (3 * M - r) - a * a * (r + M)) / (a * (r - M))
    eta = r ** 3 * (4 * a * a * M - r * (r - 3 * M) ** 2) / (a * a * (r - M) ** 2)
    return xi, eta


@pytest.mark.parametrize("a", [0.95, 0.998])
def test_kerr_spherical_photon_orbits_keep_their_radius(a):
    """Non-equatorial spherical photon orbits (the rays that outline the
    shadow, Bardeen's critical curve): started on the equator with u^r = 0 and
    the orbit's E, L, Q, the ray keeps r = r_p while theta oscillates between
    the zeros of Theta = eta + a^2 cos^2 th - xi^2 cot^2 th.  Every term of the
    Kerr geodesic equation enters: a dropped or mis-signed one moves r off r_p
    within an orbit."""
    r_pro = 2 * (1 + math.cos(2.0 / 3.0 * math.acos(-a)))
    r_ret = 2 * (1 + math.cos(2.0 / 3.0 * math.acos(a)))
    checked = 0
    for f in (0.2, 0.4, 0.6, 0.8):
        r = r_pro + f * (r_ret - r_pro)
        xi, eta = spherical_photon_orbit(a, r)
        assert eta > 0
        uth = math.sqrt(eta) / (r * r)  # p_theta = +sqrt(Q) at the equator, u^th = p_th / Sigma
        x, u = kerr_state_from_constants(1.0, a, r, math.pi / 2, 1.0, xi, 0.0, uth)
        g = O.metric(A.PRR_KERR_BL, 1.0, a, x)
        assert abs(u @ g @ u) < 1e-12 * (u[3] ** 2)
        du = O.geod(A.PRR_KERR_BL, 1.0, a, x, u)
        assert abs(du[0]) < 1e-11 * (u[3] ** 2)
        # theta turning points: cos^2 th solves a^2 c^4 + (xi^2 + eta - a^2) c^2 ... (Theta = 0)
        A2, B2, C2 = -a * a, a * a - eta - xi * xi, eta  # in c = cos^2 th: A2 c^2 + B2 c + C2 = 0
        c_max = (-B2 - math.sqrt(B2 * B2 - 4 * A2 * C2)) / (2 * A2)
        th_min = math.acos(math.sqrt(c_max))
        # the orbit is unstable: the tol-1e-12 error grows ~e^lambda near the
        # horizon, to ~3e-8 at lambda = 10 (and ~1e-3 at 25)
        S = O.Scene(A.PRR_KERR_BL, 1.0, a, tracer(lambda_max=10.0, hmax=0.05))
        res = S.trace(np.concatenate([x, u])[None])
        assert res["cls"][0] == A.PRR_STEP_CAP  # still orbiting at the affine cap
        xe = res["xend"][0]
        assert abs(xe[0] - r) < 1e-6 * r, (a, r, xe[0])
        assert th_min - 1e-6 <= xe[1] <= math.pi - th_min + 1e-6
        checked += 1
    assert checked == 4


def test_schwarzschild_photon_sphere_is_circular():
    x = np.array([3.0, math.pi / 2, 0.0, 0.0])
    L = 3.0 * math.sqrt(3.0)  # b_c = 3 sqrt(3) M
    u = np.array([0.0, 0.0, L / 9.0, 1.0 / (1 - 2 / 3.0)])
    g = O.metric(A.PRR_SCHW_SPH, 1.0, 0.0, x)
    assert abs(u @ g @ u) < 1e-12
    assert abs(O.geod(A.PRR_SCHW_SPH, 1.0, 0.0, x, u)[0]) < 1e-14


def tail(b, R, M=1.0):
    val, _ = integrate.quad(lambda w: b / math.sqrt(1 - b * b * w * w * (1 - 2 * M * w)), 0.0, 1.0 / R,
                            epsabs=1e-15, epsrel=1e-13)
    return val


def test_darwin_deflection_schwarzschild_coordinates():
    rows = json.load(open(os.path.join(GOLDEN, "darwin_deflection.json")))["rows"]
    R = 400.0
    # Global RKF45 error grows like tol^(4/5) times the orbit's sensitivity to b,
    # which diverges as r0 -> 3M; tol 1e-13 leaves < 1e-7 rad at r0 = 3.2M.
    S = O.Scene(A.PRR_SCHW_SPH, 1.0, 0.0, tracer(r_escape=R, hmax=1.0, atol=1e-13, rtol=1e-13))
    for row in rows:
        b = row["b"]
        ur = -math.sqrt(1 - (1 - 2 / R) * b * b / R ** 2)
        x = np.array([R, math.pi / 2, 0.0, 0.0])
        u = np.array([ur, 0.0, b / R ** 2, 1 / (1 - 2 / R)])
        res = S.trace(np.concatenate([x, u])[None])
        assert res["cls"][0] == A.PRR_ESCAPE
        r_end, phi_end = res["xend"][0, 0], res["xend"][0, 2]
        expect = row["alpha"] + math.pi - tail(b, R) - tail(b, r_end)
        n = round((expect - phi_end) / (2 * math.pi))
        bound = 1e-7 if row["r0"] <= 4.0 else 5e-9
        assert abs(phi_end + 2 * math.pi * n - expect) < bound, (b, phi_end, expect)


def test_capture_threshold_is_3sqrt3():
    R = 300.0
    S = O.Scene(A.PRR_SCHW_SPH, 1.0, 0.0, tracer(r_escape=R + 1, hmax=20.0, atol=1e-11, rtol=1e-11))
    bc = 3 * math.sqrt(3)
    states = []
    for b in (bc * (1 - 1e-4), bc * (1 + 1e-4)):
        ur = -math.sqrt(1 - (1 - 2 / R) * b * b / R ** 2)
        states.append([R, math.pi / 2, 0.0, 0.0, ur, 0.0, b / R ** 2, 1 / (1 - 2 / R)])
    res = S.trace(np.array(states))
    assert list(res["cls"]) == [A.PRR_HORIZON, A.PRR_ESCAPE]


def test_flat_rays_are_straight():
    S = O.Scene(A.PRR_FLAT, 0.0, 0.0, tracer(r_escape=100.0))
    f = A.make_frame([30.0, 0.0, 5.0], math.pi, 0.1, math.radians(60), 16, 12, mass=0.0,
                     projection=A.PRR_PROJ_SPHERICAL)
    res = S.render_rays(f)
    for i in range(f.nrays):
        uv = O.sample_uv(f, i)
        d = O.camera_dir(f, *uv)
        assert res["cls"][i] == A.PRR_ESCAPE
        assert np.allclose(res["endpoint"][i], d, atol=1e-12)


def test_ks_and_schwarzschild_coordinates_trace_the_same_orbit():
    # Same event and spatial coordinate velocity -> same null geodesic; the spatial
    # coordinates of Kerr-Schild and Schwarzschild charts coincide.
    h, N = 0.01, 1500
    T = tracer(hmin=h, hmax=h, h=h, max_steps=N, r_escape=1e9)
    Sk = O.Scene(A.PRR_SCHW_KS_CART, 1.0, 0.0, T)
    Ss = O.Scene(A.PRR_SCHW_SPH, 1.0, 0.0, T)
    for _ in range(5):
        pos = np.array([12.0, 3.0, 2.0]) + RNG.normal(size=3)
        d = -pos / np.linalg.norm(pos) + 0.3 * RNG.normal(size=3)
        d /= np.linalg.norm(d)
        xk, uk = O.init_ray(A.PRR_SCHW_KS_CART, 1.0, 0.0, pos, d)
        xs, us = O.init_ray(A.PRR_SCHW_SPH, 1.0, 0.0, pos, d)
        rk = Sk.trace(np.concatenate([xk, uk])[None])
        rs = Ss.trace(np.concatenate([xs, us])[None])
        if rk["cls"][0] != A.PRR_STEP_CAP or rs["cls"][0] != A.PRR_STEP_CAP:
            continue
        pk = rk["xend"][0, :3]
        r, th, ph = rs["xend"][0, :3]
        ps = r * np.array([math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th)])
        assert np.allclose(pk, ps, atol=1e-8), (pk, ps)


def test_kerr_escape_direction_at_zero_spin_is_the_ks_velocity():
    """The escape direction of a Kerr ray is the Cartesian direction of its
    coordinate velocity (u^r e_r + r u^th e_th + r sin th u^ph e_ph, in the
    flat-space frame of (r, th, ph)); the Cartesian metrics use their spatial
    velocity (getColorDefault2, main/base.h:119-128).  At a = 0 the Boyer-
    Lindquist and Kerr-Schild charts share r, th, ph, so the same frame gives
    the same classes and, for escaping rays located on r = r_escape (D21), the
    same directions."""
    T = tracer(r_escape=60.0)
    pos, al, be = CF.orbit_pose(15.0, 70.0, 0.4)
    f = A.make_frame(pos, al, be, math.radians(60.0), 32, 24, mass=1.0, spin=0.0)
    rk = O.Scene(A.PRR_SCHW_KS_CART, 1.0, 0.0, T).render_rays(f)
    rb = O.Scene(A.PRR_KERR_BL, 1.0, 0.0, T).render_rays(f)
    assert np.array_equal(rk["cls"], rb["cls"])
    esc = rk["cls"] == A.PRR_ESCAPE
    assert esc.sum() > 300 and (rk["cls"] == A.PRR_HORIZON).sum() > 20
    cosang = np.clip(np.sum(rk["endpoint"][esc] * rb["endpoint"][esc], axis=1), -1.0, 1.0)
    assert np.arccos(cosang).max() < 1e-7


def _fixed_step_end(solver, h, N, x, u):
    T = tracer(solver=solver, h=h, hmin=h, hmax=h, max_steps=N, r_escape=1e9)
    S = O.Scene(A.PRR_SCHW_KS_CART, 1.0, 0.0, T)
    return S.trace(np.concatenate([x, u])[None])["xend"][0]


@pytest.mark.parametrize("solver,order", [(A.PRR_RKF45, 4), (A.PRR_EULER, 1)])
def test_convergence_order(solver, order):
    x, u = O.init_ray(A.PRR_SCHW_KS_CART, 1.0, 0.0, [10.0, 0.0, 1.0], [-0.6, 0.8, 0.0])
    lam = 4.0
    h = 0.2 if solver == A.PRR_RKF45 else 0.01
    y1 = _fixed_step_end(solver, h, int(round(lam / h)), x, u)
    y2 = _fixed_step_end(solver, h / 2, int(round(lam / h * 2)), x, u)
    y4 = _fixed_step_end(solver, h / 4, int(round(lam / h * 4)), x, u)
    e1 = np.abs(y1 - y2).max()
    e2 = np.abs(y2 - y4).max()
    ratio = e1 / e2
    assert 0.75 * 2 ** order < ratio < 1.3 * 2 ** order, ratio


def test_fehlberg_tableau_conditions():
    Am, b4, b5 = O.fehlberg()
    c = Am.sum(axis=1)
    assert np.allclose(c, [0, 1 / 4, 3 / 8, 12 / 13, 1, 1 / 2])
    for b, p in ((b4, 4), (b5, 5)):
        assert abs(b.sum() - 1) < 1e-15
        for q in range(1, p):  # quadrature conditions sum b c^q = 1/(q+1)
            assert abs((b * c ** q).sum() - 1 / (q + 1)) < 1e-14


def test_ray_sphere_chord_brute_force():
    spheres = [([8.0, 1.0, 0.5], 1.0), ([8.5, 1.2, 0.3], 0.7)]
    for _ in range(300):
        P0 = RNG.uniform(-2, 12, size=3)
        P1 = P0 + RNG.normal(size=3) * 3
        k, t = O.chord_hit(spheres, P0, P1)
        ts = np.linspace(0, 1, 20001)
        pts = P0[None] + ts[:, None] * (P1 - P0)[None]
        best = None
        for j, (c, r) in enumerate(spheres):
            inside = np.linalg.norm(pts - np.array(c)[None], axis=1) <= r
            if inside.any():
                tj = ts[np.argmax(inside)]
                if best is None or tj < best[1]:
                    best = (j, tj)
        if best is None:
            assert k == -1
        else:
            assert k == best[0] and abs(t - best[1]) <= 1e-4


def test_camera_centre_ray_is_forward_and_corners_match_fov():
    f = A.make_frame([30.0, 0.0, 0.0], math.pi, math.radians(5), math.radians(60), 101, 51)
    Bm = O.camera_basis(f.alpha, f.beta)
    assert np.allclose(O.camera_dir(f, 0.5, 0.5), Bm[0], atol=1e-15)
    # top-centre ray makes vfov/2 with forward
    d = O.camera_dir(f, 0.5, 0.0)
    assert abs(math.acos(d @ Bm[0]) - math.radians(30)) < 1e-13
    assert d @ Bm[1] > 0  # row 0 is the top of the image
    # right edge: horizontal half-angle atan(tan(vfov/2) W/H)
    d = O.camera_dir(f, 1.0, 0.5)
    assert abs(math.atan2(d @ Bm[2], d @ Bm[0]) - math.atan(math.tan(math.radians(30)) * 101 / 51)) < 1e-13
    assert np.allclose(np.cross(Bm[0], Bm[1]), Bm[2])


def test_sample_positions_regular_grid():
    f = A.make_frame([30, 0, 0], math.pi, 0, 1.0, 4, 3, ssaa=A.PRR_SSAA_REGULAR, samples=9)
    # pixel (1, 2) sample q*3 + p at ((i + (p+1)/4)/W, (j + (q+1)/4)/H)  (kernels/ssaa.cuh:41-46)
    for q in range(3):
        for p in range(3):
            uv = O.sample_uv(f, (2 * 4 + 1) * 9 + q * 3 + p)
            assert np.allclose(uv, [(1 + (p + 1) / 4) / 4, (2 + (q + 1) / 4) / 3], atol=0, rtol=1e-15)
    f1 = A.make_frame([30, 0, 0], math.pi, 0, 1.0, 4, 3)
    assert np.allclose(O.sample_uv(f1, 5), [(1 + 0.5) / 4, (1 + 0.5) / 3])


@pytest.mark.parametrize("ssaa,samples", [(A.PRR_SSAA_RANDOM, 5), (A.PRR_SSAA_JITTERED, 9)])
def test_random_and_jittered_samples_stay_in_their_pixel_and_cell(ssaa, samples):
    f = A.make_frame([30, 0, 0], math.pi, 0, 1.0, 7, 5, ssaa=ssaa, samples=samples, seed=12345)
    n = int(round(math.sqrt(samples)))
    for ray in range(f.nrays):
        pix, k = divmod(ray, samples)
        j, i = divmod(pix, f.width)
        u, v = O.sample_uv(f, ray)
        x, y = u * f.width - i, v * f.height - j  # position inside the pixel
        assert 0.0 <= x < 1.0 and 0.0 <= y < 1.0, (ray, x, y)
        if ssaa == A.PRR_SSAA_JITTERED:
            q, p = divmod(k, n)
            assert p <= x * n < p + 1 and q <= y * n < q + 1, (ray, x * n, y * n)


def test_random_samples_are_uniform_and_keyed_by_the_seed():
    f = A.make_frame([30, 0, 0], math.pi, 0, 1.0, 64, 64, ssaa=A.PRR_SSAA_RANDOM, samples=16, seed=7)
    xy = []
    for ray in range(0, f.nrays, 7):
        pix = ray // 16
        j, i = divmod(pix, f.width)
        u, v = O.sample_uv(f, ray)
        xy.append((u * f.width - i, v * f.height - j))
    xy = np.array(xy)
    # 10 x 10 histogram of the in-pixel offsets: chi-square near its 99 degrees of freedom
    hist, _, _ = np.histogram2d(xy[:, 0], xy[:, 1], bins=10, range=[[0, 1], [0, 1]])
    expected = len(xy) / 100
    chi2 = ((hist - expected) ** 2 / expected).sum()
    assert 60 < chi2 < 150, chi2
    # u and v are not correlated
    assert abs(np.corrcoef(xy[:, 0], xy[:, 1])[0, 1]) < 0.05
    g = A.make_frame([30, 0, 0], math.pi, 0, 1.0, 64, 64, ssaa=A.PRR_SSAA_RANDOM, samples=16, seed=8)
    assert not np.array_equal(O.sample_uv(f, 123), O.sample_uv(g, 123))
    assert np.array_equal(O.sample_uv(f, 123), O.sample_uv(f, 123))


@pytest.mark.parametrize("solver,order", [(A.PRR_DORMAND, 5), (A.PRR_RK4, 4), (A.PRR_RKF45, 4)])
def test_runge_kutta_order_conditions(solver, order):
    """Butcher's order conditions up to the propagated order, for the
    propagated weights; for the adaptive pairs the embedded weights reach the
    other order (Fehlberg: 5, Dormand-Prince: 4)."""
    Am, b, bh = O.tableau(solver)
    c = Am.sum(axis=1)

    def conditions(w, p):
        out = [w.sum() - 1]
        if p >= 2:
            out.append(w @ c - 1 / 2)
        if p >= 3:
            out += [w @ c ** 2 - 1 / 3, w @ Am @ c - 1 / 6]
        if p >= 4:
            out += [w @ c ** 3 - 1 / 4, w @ (c * (Am @ c)) - 1 / 8, w @ Am @ c ** 2 - 1 / 12, w @ Am @ Am @ c - 1 / 24]
        if p >= 5:
            out += [w @ c ** 4 - 1 / 5, w @ (c ** 2 * (Am @ c)) - 1 / 10, w @ ((Am @ c) ** 2) - 1 / 20,
                    w @ (c * (Am @ c ** 2)) - 1 / 15, w @ Am @ c ** 3 - 1 / 20, w @ (c * (Am @ Am @ c)) - 1 / 30,
                    w @ Am @ (c * (Am @ c)) - 1 / 40, w @ Am @ Am @ c ** 2 - 1 / 60, w @ Am @ Am @ Am @ c - 1 / 120]
        return np.abs(out).max()

    assert conditions(b, order) < 1e-14
    if solver != A.PRR_RK4:
        assert conditions(bh, 9 - order) < 1e-14
        if order < 5:  # and not more (conditions above order 5 are not listed here)
            assert conditions(b, order + 1) > 1e-6


@pytest.mark.parametrize("solver,order", [(A.PRR_DORMAND, 5), (A.PRR_RK4, 4)])
def test_fixed_step_convergence_of_the_added_solvers(solver, order):
    """Fixed-step integration of a flat-space... with curvature: a Schwarzschild
    orbit; the error against a much finer run falls as h^order."""
    pos = np.array([10.0, 0.0, 0.0])
    x, u = O.init_ray(A.PRR_SCHW_KS_CART, 1.0, 0.0, pos, [-0.3, 0.9539392014169456, 0.0])
    x0u0 = np.concatenate([x, u])[None]

    def end(h, n):
        S = O.Scene(A.PRR_SCHW_KS_CART, 1.0, 0.0, tracer(solver=solver, h=h, hmin=h, hmax=h, max_steps=n,
                                                          r_escape=1e9, atol=1.0, rtol=0.0))
        return S.trace(x0u0)["xend"][0]

    ref = end(0.01, 2000)
    e1 = np.abs(end(0.4, 50) - ref).max()
    e2 = np.abs(end(0.2, 100) - ref).max()
    rate = np.log2(e1 / e2)
    assert order - 0.4 < rate < order + 0.6, rate


@pytest.mark.parametrize("q", [0.0, 0.5, 0.9, 1.0])
def test_reissner_nordstrom_photon_sphere_and_horizon(q):
    """Charge q in the Schwarzschild-coordinate metric: the circular photon
    orbit at r = (3M + sqrt(9M^2 - 8q^2)) / 2 (geodesic acceleration zero for
    the null tangential state there), and infall stops at r+ = M + sqrt(M^2 - q^2)."""
    M = 1.0
    r = (3 * M + math.sqrt(9 * M * M - 8 * q * q)) / 2
    f = 1 - 2 * M / r + q * q / (r * r)
    x = np.array([r, math.pi / 2, 0.0, 0.0])
    u = np.array([0.0, 0.0, 1.0 / r, 1.0 / math.sqrt(f)])  # f (u^t)^2 = r^2 (u^phi)^2
    g = O.metric(A.PRR_SCHW_SPH, M, q, x)
    assert abs(u @ g @ u) < 1e-12
    assert abs(O.geod(A.PRR_SCHW_SPH, M, q, x, u)[0]) < 1e-13
    # slightly off the photon radius the orbit is not circular (unstable both ways)
    x2 = x.copy()
    x2[0] *= 1.01
    assert abs(O.geod(A.PRR_SCHW_SPH, M, q, x2, u)[0]) > 1e-5
    # radial infall from r = 10 ends at the outer horizon
    S = O.Scene(A.PRR_SCHW_SPH, M, q, tracer(r_escape=100.0, horizon_eps_rel=1e-3, hmax=0.5))
    x0, u0 = O.init_ray(A.PRR_SCHW_SPH, M, q, [10.0, 0.0, 0.0], [-1.0, 0.0, 0.0])
    o = S.trace(np.concatenate([x0, u0])[None])
    rp = M + math.sqrt(M * M - q * q)
    assert o["cls"][0] == A.PRR_HORIZON
    assert rp * (1 - 1e-9) < o["xend"][0][0] <= rp * (1 + 1e-3) + 1e-12
