"""The device plug-ins' formulas, evaluated on the host (include/prr_debug.h),
against the oracle.  No GPU: these run in the CPU suite.

The Kerr check is a cross-formulation test: the kernels integrate the
separated Mino-time equations for (u = 1/r, du/ds, mu = cos th, dmu/ds, phi,
lambda) with E, L and Carter's Q; the oracle uses the second-order geodesic
equation with Christoffel symbols in affine time.  On null states they must
agree on every second derivative, converted between the two time variables
(d/ds = Sigma d/dlambda), which pins every term of both.
"""
import math

import numpy as np
import pytest

import oracle as O
from parrelray_amd import _abi as A
from parrelray_amd import api

RNG = np.random.default_rng(7)


def points(metric, n):
    out = []
    for _ in range(n):
        if metric in (A.PRR_SCHW_SPH, A.PRR_KERR_BL):
            x = np.array([RNG.uniform(2.5, 60.0), RNG.uniform(0.2, math.pi - 0.2), RNG.uniform(-3, 9), 0.0])
        else:
            v = RNG.normal(size=3)
            x = np.concatenate([v / np.linalg.norm(v) * RNG.uniform(2.5, 60.0), [RNG.normal()]])
        out.append((x, RNG.normal(size=4)))
    return out


@pytest.mark.parametrize("metric", [A.PRR_FLAT, A.PRR_SCHW_KS_CART, A.PRR_SCHW_SPH])
def test_second_order_rhs_matches_oracle_getgeod(metric):
    for x, u in points(metric, 50):
        y = np.concatenate([x, u])
        f = api.rhs(metric, y, M=1.0)
        assert np.array_equal(f[:4], u)
        du = O.geod(metric, 1.0, 0.0, x, u)
        scale = np.abs(du).max() + 1e-300
        assert np.allclose(f[4:], du, rtol=1e-12, atol=1e-13 * scale), (x, u, f[4:], du)


@pytest.mark.parametrize("q", [0.3, 0.95])
def test_charged_sph_rhs_matches_oracle(q):
    """Reissner-Nordstrom: the charge rides in the metric's second parameter."""
    for x, u in points(A.PRR_SCHW_SPH, 50):
        f = api.rhs(A.PRR_SCHW_SPH, np.concatenate([x, u]), M=1.0, a=q)
        du = O.geod(A.PRR_SCHW_SPH, 1.0, q, x, u)
        scale = np.abs(du).max() + 1e-300
        assert np.allclose(f[4:], du, rtol=1e-12, atol=1e-13 * scale), (x, u, f[4:], du)


@pytest.mark.parametrize("b", [0.5, 2.0])
def test_ellis_rhs_matches_oracle(b):
    """Ellis wormhole: the throat radius rides in the metric's second
    parameter; l takes both signs."""
    for x, u in points(A.PRR_SCHW_SPH, 50):
        x[0] *= RNG.choice([-1.0, 1.0]) * RNG.uniform(0.0, 0.5)
        f = api.rhs(A.PRR_ELLIS, np.concatenate([x, u]), M=0.0, a=b)
        assert np.array_equal(f[:4], u)
        du = O.geod(A.PRR_ELLIS, 0.0, b, x, u)
        scale = np.abs(du).max() + 1e-300
        assert np.allclose(f[4:], du, rtol=1e-12, atol=1e-13 * scale), (x, u, f[4:], du)


def test_ellis_init_ray_matches_oracle():
    for _ in range(30):
        pos = RNG.normal(size=3)
        pos = pos / np.linalg.norm(pos) * RNG.uniform(4, 50)
        d = RNG.normal(size=3)
        d /= np.linalg.norm(d)
        xo, uo = O.init_ray(A.PRR_ELLIS, 0.0, 1.5, pos, d)
        y, _ = api.init_ray(A.PRR_ELLIS, pos, d, 0.0, 1.5)
        assert np.allclose(y, np.concatenate([xo, uo]), rtol=1e-13, atol=1e-14)


def test_sph_rhs_keeps_the_tan_clamp():
    # |tan th| < 1e-4 -> tan = sin = 1e-4 ("metrics/schwarzschild - sph.cuh":98-101)
    for th in (1e-6, math.pi - 2e-5, -3e-5):
        x = np.array([10.0, th, 0.3, 0.0])
        u = np.array([0.1, 0.02, 0.03, 1.1])
        f = api.rhs(A.PRR_SCHW_SPH, np.concatenate([x, u]))
        assert np.allclose(f[4:], O.geod(A.PRR_SCHW_SPH, 1.0, 0.0, x, u), rtol=1e-9)


def null_kerr_states(a, n, rmin=None, rmax=80.0):
    """Random null Kerr states: (x, u) from the oracle's initRays and the device state."""
    M = 1.0
    rp = M + math.sqrt(M * M - a * a)
    rmin = rmin or max(rp * 1.05, 2.05)  # outside the ergoregion, where the static frame exists
    out = []
    while len(out) < n:
        r = RNG.uniform(rmin, rmax)
        th = RNG.uniform(0.15, math.pi - 0.15)
        ph = RNG.uniform(-3, 9)
        pos = r * np.array([math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th)])
        d = RNG.normal(size=3)
        d /= np.linalg.norm(d)
        x, u = O.init_ray(A.PRR_KERR_BL, M, a, pos, d)
        y, ELQ = api.state_from_xu(A.PRR_KERR_BL, x, u, M, a)
        assert np.all(np.isfinite(y)) and np.all(np.isfinite(ELQ))
        out.append((x, u, y, ELQ))
    return out


@pytest.mark.parametrize("a", [0.0, 0.5, 0.95, 0.998])
def test_kerr_mino_rhs_matches_christoffel_form(a):
    M = 1.0
    for x, u, y, (E, L, Q) in null_kerr_states(a, 60):
        f = api.rhs(A.PRR_KERR_BL, y, M, a, E, L, Q)
        r, th = x[0], x[1]
        s, c = math.sin(th), math.cos(th)
        Sig = r * r + a * a * c * c
        du = O.geod(A.PRR_KERR_BL, M, a, x, u)  # (r, th, phi, t)''
        rd, thd = u[0], u[1]
        Sigd = 2 * r * rd - 2 * a * a * c * s * thd
        # first derivatives: the state's own u' and mu', phi' = Sigma u^phi, lambda' = Sigma
        assert np.isclose(y[1], -Sig * rd / r ** 2, rtol=1e-13)
        assert np.isclose(y[3], -s * Sig * thd, rtol=1e-13, atol=1e-15)
        assert np.isclose(f[4], Sig * u[2], rtol=1e-10, atol=1e-12 * abs(Sig * u[2]) + 1e-12)
        assert np.isclose(f[5], Sig, rtol=1e-13)
        # u'' = Sigma d/dlambda(-Sigma rdot / r^2)
        q_dot = -(Sigd * rd + Sig * du[0]) / r ** 2 + 2 * Sig * rd * rd / r ** 3
        scale_u = Sig * (abs(Sigd * rd) + abs(Sig * du[0]) + 2 * abs(Sig * rd * rd / r)) / r ** 2
        assert abs(f[1] - Sig * q_dot) <= 1e-10 * scale_u, (a, x, u, f[1], Sig * q_dot)
        # mu'' = Sigma d/dlambda(Sigma mudot), mudot = -sin th thdot
        mud = -s * thd
        mudd = -c * thd * thd - s * du[1]
        expect = Sig * (Sigd * mud + Sig * mudd)
        scale_m = Sig * (abs(Sigd * mud) + Sig * (abs(c * thd * thd) + abs(s * du[1])))
        assert abs(f[3] - expect) <= 1e-10 * scale_m + 1e-300, (a, x, u, f[3], expect)


@pytest.mark.parametrize("a", [0.95, 0.998])
def test_kerr_mino_rhs_vanishes_on_spherical_photon_orbits(a):
    """On a spherical photon orbit r_p (Teo 2003) P(u) has a double root at
    u_p = 1/r_p: the state has u' = 0 and the Mino-time RHS u'' = P'(u)/2 = 0,
    which ties c1, c2 and c3 to E, L, Q and a."""
    r_pro = 2 * (1 + math.cos(2.0 / 3.0 * math.acos(-a)))
    r_ret = 2 * (1 + math.cos(2.0 / 3.0 * math.acos(a)))
    for f in (0.15, 0.5, 0.85):
        r = r_pro + f * (r_ret - r_pro)
        xi = (r * r * (3 - r) - a * a * (r + 1)) / (a * (r - 1))
        eta = r ** 3 * (4 * a * a - r * (r - 3) ** 2) / (a * a * (r - 1) ** 2)
        g = O.metric(A.PRR_KERR_BL, 1.0, a, [r, math.pi / 2, 0.0, 0.0])
        gi = np.linalg.inv(g)
        u = np.array([0.0, math.sqrt(eta) / (r * r), gi[2, 3] * -1.0 + gi[2, 2] * xi, gi[3, 3] * -1.0 + gi[3, 2] * xi])
        x = np.array([r, math.pi / 2, 0.0, 0.0])
        y, (E, L, Q) = api.state_from_xu(A.PRR_KERR_BL, x, u, 1.0, a)
        assert abs(L / E - xi) < 1e-12 * abs(xi) and abs(Q / E ** 2 - eta) < 1e-11 * eta
        f_ = api.rhs(A.PRR_KERR_BL, y, 1.0, a, E, L, Q)
        scale = E * E * y[0]
        assert abs(y[1]) < 1e-14 and abs(f_[1]) < 1e-12 * scale, (r, y[1], f_[1])
        # off the orbit radius the radial acceleration is not zero
        y2 = y.copy()
        y2[0] *= 1.01
        assert abs(api.rhs(A.PRR_KERR_BL, y2, 1.0, a, E, L, Q)[1]) > 1e-5 * scale


def test_kerr_first_integrals_hold_at_set_up():
    """u'^2 = P(u) and mu'^2 = Theta_mu(mu) for the E, L, Q of a null ray."""
    for a in (0.0, 0.6, 0.95):
        for _, _, y, (E, L, Q) in null_kerr_states(a, 30):
            uu, w, mu, nu = y[:4]
            K = (L - a * E) ** 2 + Q
            Aa = a * (a * E - L)
            P = (E + Aa * uu * uu) ** 2 - (1 - 2 * uu + a * a * uu * uu) * K * uu * uu
            Th = Q * (1 - mu * mu) - mu * mu * L * L + a * a * E * E * mu * mu * (1 - mu * mu)
            assert abs(w * w - P) <= 1e-11 * (E * E + K * uu * uu), (a, w * w, P)
            assert abs(nu * nu - Th) <= 1e-11 * (abs(Q) + L * L + a * a * E * E + 1), (a, nu * nu, Th)


def test_kerr_delta_form_equals_the_difference_of_rhs():
    """The FP32-bulk stage function g(d) = f(y0 + d) - f(y0), against FP64 differences."""
    for a in (0.0, 0.7, 0.998):
        for _, _, y, (E, L, Q) in null_kerr_states(a, 40, rmin=2.2):
            f0 = api.rhs(A.PRR_KERR_BL, y, 1.0, a, E, L, Q)
            for scale in (1e-1, 1e-3, 1e-6):
                d = RNG.normal(size=6) * scale * np.maximum(np.abs(y), 1e-3)
                d[0] = max(d[0], -0.5 * y[0])
                # stay in the chart: |mu| < 1 (a state past the pole has 1 - mu^2 < 0)
                d[2] = np.clip(y[2] + d[2], -0.999, 0.999) - y[2]
                g = api.kerr_delta_rhs(y, d, a, E, L, Q)
                ref = api.rhs(A.PRR_KERR_BL, y + d, 1.0, a, E, L, Q) - f0
                # FP32 relative rounding of the difference, plus FP64 rounding of the reference
                tol = 3e-6 * np.abs(ref) + 1e-15 * np.abs(f0) + 1e-30
                # 1 - mu^2 is formed in FP32 as s20 - (mu^2 - mu0^2): its relative error
                # grows by (s20 + |mu^2 - mu0^2|) / s2 near the axis, and L / s2 with it
                s20, s2 = 1 - y[2] ** 2, 1 - (y[2] + d[2]) ** 2
                tol[4] += 1.2e-7 * (s20 + abs(s2 - s20)) / s2 * abs(L) / s2
                assert np.all(np.abs(g - ref) <= tol + 2e-7 * np.abs(g)), (a, scale, g, ref)


def test_kerr_fp32_bulk_attempt_tracks_the_fp64_attempt():
    """The FP32-bulk attempt's increment and error vector equal the FP64 attempt's up to
    FP32 rounding of the nonlinear part d4 - h f(y), which is small for an accurate step."""
    for a in (0.0, 0.95):
        for _, _, y, (E, L, Q) in null_kerr_states(a, 40, rmin=2.5):
            F = api.rhs(A.PRR_KERR_BL, y, 1.0, a, E, L, Q)
            h = RNG.uniform(0.01, 0.2) * min(y[0] / abs(F[0]), 1.0 / np.abs(F[1:4]).max())
            d64, e64 = api.kerr_stages(y, h, a, E, L, Q)
            d32, e32 = api.kerr_stages(y, h, a, E, L, Q, fp32=True)
            nl = np.abs(d64 - h * F)
            assert np.all(np.abs(d32 - d64) <= 1e-5 * nl + 1e-15 * np.abs(d64) + 1e-300), (h, d32 - d64, nl)
            assert np.all(np.abs(e32 - e64) <= 1e-4 * np.abs(e64) + 1e-6 * nl + 1e-300), (h, e32, e64)


def test_kerr_state_round_trip_and_constants():
    a = 0.9
    for _ in range(30):
        pos = RNG.normal(size=3)
        pos = pos / np.linalg.norm(pos) * RNG.uniform(4, 50)
        d = RNG.normal(size=3)
        d /= np.linalg.norm(d)
        xo, uo = O.init_ray(A.PRR_KERR_BL, 1.0, a, pos, d)
        y, ELQ = api.init_ray(A.PRR_KERR_BL, pos, d, 1.0, a)
        E, L, Q = O.kerr_constants(1.0, a, xo, uo)
        assert np.allclose(ELQ, [E, L, Q], rtol=1e-11, atol=1e-12)
        x, u = api.xu_from_state(A.PRR_KERR_BL, y, 1.0, a, *ELQ)
        assert np.allclose(x[:3], xo[:3], rtol=1e-14)
        assert np.allclose(u, uo, rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("metric", [A.PRR_FLAT, A.PRR_SCHW_KS_CART, A.PRR_SCHW_SPH])
def test_init_ray_matches_oracle(metric):
    for _ in range(30):
        pos = RNG.normal(size=3)
        pos = pos / np.linalg.norm(pos) * RNG.uniform(4, 50)
        d = RNG.normal(size=3)
        d /= np.linalg.norm(d)
        xo, uo = O.init_ray(metric, 1.0, 0.0, pos, d)
        y, _ = api.init_ray(metric, pos, d, 1.0, 0.0)
        assert np.allclose(y, np.concatenate([xo, uo]), rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize("proj,ssaa,s", [(A.PRR_PROJ_PINHOLE, A.PRR_SSAA_NONE, 1),
                                         (A.PRR_PROJ_PINHOLE, A.PRR_SSAA_REGULAR, 4),
                                         (A.PRR_PROJ_SPHERICAL, A.PRR_SSAA_REGULAR, 9)])
def test_camera_rays_match_oracle(proj, ssaa, s):
    f = A.make_frame([20.0, 3.0, 4.0], 3.0, 0.2, math.radians(50), 37, 23, ssaa=ssaa, samples=s, projection=proj)
    for ray in RNG.integers(0, f.nrays, size=200):
        d = api.camera_ray(f, int(ray))
        uv = O.sample_uv(f, int(ray))
        assert np.allclose(d, O.camera_dir(f, *uv), rtol=0, atol=2e-16)


@pytest.mark.parametrize("metric", [A.PRR_SCHW_KS_CART, A.PRR_SCHW_SPH, A.PRR_KERR_BL])
def test_fp32_rhs_close_to_fp64(metric):
    if metric == A.PRR_KERR_BL:
        cases = [(y, ELQ) for _, _, y, ELQ in null_kerr_states(0.9, 20)]
    else:
        cases = [(np.concatenate([x, u]), (0.0, 0.0, 0.0)) for x, u in points(metric, 20)]
    for y, ELQ in cases:
        f64 = api.rhs(metric, y, 1.0, 0.9 if metric == A.PRR_KERR_BL else 0.0, *ELQ)
        f32 = api.rhs(metric, y, 1.0, 0.9 if metric == A.PRR_KERR_BL else 0.0, *ELQ, fp32=True)
        assert np.allclose(f32, f64, rtol=2e-4, atol=2e-5 * np.abs(f64).max())


@pytest.mark.parametrize("a", [0.0, 0.7, 0.95])
def test_kerr_rhs_vjp_matches_central_differences(a):
    """The adjoint kernel's hand-written reverse mode of the Kerr RHS."""
    for _, _, y, (E, L, Q) in null_kerr_states(a, 40, rmin=2.0):
        fb = RNG.normal(size=6)
        yb, cb = api.rhs_vjp(y, fb, a, E, L, Q)

        def F(yy, c):
            return fb @ api.rhs(A.PRR_KERR_BL, yy, 1.0, c[3], c[0], c[1], c[2])

        c0 = np.array([E, L, Q, a])
        scaleF = np.abs(fb) @ np.abs(api.rhs(A.PRR_KERR_BL, y, 1.0, a, E, L, Q))
        for i in range(6):
            h = 1e-6 * max(1e-2, abs(y[i]))
            yp, ym = y.copy(), y.copy()
            yp[i] += h
            ym[i] -= h
            fd = (F(yp, c0) - F(ym, c0)) / (2 * h)
            tol = 1e-6 * (abs(fd) + 1e-3 * np.abs(yb).max()) + 1e-14 * scaleF / h
            assert abs(yb[i] - fd) <= tol, (i, yb[i], fd)
        for j in range(4):
            h = 1e-6 * max(1.0, abs(c0[j]))
            cp, cm = c0.copy(), c0.copy()
            cp[j] += h
            cm[j] -= h
            fd = (F(y, cp) - F(y, cm)) / (2 * h)
            tol = 1e-6 * (abs(fd) + 1e-3 * np.abs(cb).max()) + 1e-14 * scaleF / h
            assert abs(cb[j] - fd) <= tol, (j, cb[j], fd)
