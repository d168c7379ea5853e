"""The device plug-ins' formulas, evaluated on the host (include/prr_debug.h),
against the oracle.  No GPU: these run in the CPU suite.

The Kerr check is a cross-formulation test: the kernels integrate Hamilton's
equations for (r, cos th, phi, p_r, p_mu) with conserved E and L; the oracle uses
the second-order geodesic equation with Christoffel symbols.  They must agree
on dx/dlambda and on dp/dlambda = d(g u)/dlambda for arbitrary (not only
null) u, which pins every term of both.
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


def test_sph_rhs_keeps_the_tan_clamp():
    # |tan th| < 1e-4 -> tan = sin = 1e-4 ("metrics/schwarzschild - sph.cuh":98-101)
    for th in (1e-6, math.pi - 2e-5, -3e-5):
        x = np.array([10.0, th, 0.3, 0.0])
        u = np.array([0.1, 0.02, 0.03, 1.1])
        f = api.rhs(A.PRR_SCHW_SPH, np.concatenate([x, u]))
        assert np.allclose(f[4:], O.geod(A.PRR_SCHW_SPH, 1.0, 0.0, x, u), rtol=1e-9)


@pytest.mark.parametrize("a", [0.0, 0.5, 0.95, 0.998])
def test_kerr_hamiltonian_rhs_matches_christoffel_form(a):
    M = 1.0
    rp = M + math.sqrt(M * M - a * a)
    for _ in range(60):
        x = np.array([RNG.uniform(rp * 1.05, 80.0), RNG.uniform(0.15, math.pi - 0.15), RNG.uniform(-3, 9), 0.0])
        u = RNG.normal(size=4)
        y, EL = api.state_from_xu(A.PRR_KERR_BL, x, u, M, a)
        f = api.rhs(A.PRR_KERR_BL, y, M, a, EL[0], EL[1])
        s, c = math.sin(x[1]), math.cos(x[1])
        # positions: y = (r, mu = cos th, phi): y' = (u^r, -sin th u^th, u^phi)
        assert np.allclose(f[:3], [u[0], -s * u[1], u[2]], rtol=1e-11, atol=1e-13 * np.abs(u).max())
        # momenta: dp_i/dlambda = d_k g_{i nu} u^k u^nu + g_{i nu} du^nu for i = r, th
        g = O.metric(A.PRR_KERR_BL, M, a, x)
        dg = O.dmetric_kerr(M, a, x)
        du = O.geod(A.PRR_KERR_BL, M, a, x, u)
        pdot, pscale = [], []
        for i in (0, 1):
            pdot.append(sum(dg[k][i] @ u * u[k] for k in range(4)) + g[i] @ du)
            pscale.append(sum(np.abs(dg[k][i]) @ np.abs(u) * abs(u[k]) for k in range(4)) + np.abs(g[i]) @ np.abs(du))
        assert abs(f[3] - pdot[0]) <= 1e-11 * pscale[0], (a, x, u, f[3], pdot[0])
        # p_mu = -p_th / sin th  ->  p_mu' = -p_th'/sin th + p_th cos th th' / sin^2 th
        pth = g[1, 1] * u[1]
        expect = -pdot[1] / s + pth * c * u[1] / (s * s)
        scale = pscale[1] / s + abs(pth * c * u[1]) / (s * s)
        assert abs(f[4] - expect) <= 1e-11 * scale, (a, x, u, f[4], expect)


def test_kerr_state_round_trip_and_constants():
    a = 0.9
    for _ in range(30):
        pos = RNG.normal(size=3)
        pos = pos / np.linalg.norm(pos) * RNG.uniform(4, 50)
        d = RNG.normal(size=3)
        d /= np.linalg.norm(d)
        xo, uo = O.init_ray(A.PRR_KERR_BL, 1.0, a, pos, d)
        y, EL = api.init_ray(A.PRR_KERR_BL, pos, d, 1.0, a)
        E, L, _ = O.kerr_constants(1.0, a, xo, uo)
        assert np.allclose(EL, [E, L], rtol=1e-12, atol=1e-13)
        x, u = api.xu_from_state(A.PRR_KERR_BL, y, 1.0, a, EL[0], EL[1])
        assert np.allclose(x[:3], xo[:3], rtol=1e-14)
        assert np.allclose(u, uo, rtol=1e-11, atol=1e-13)


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
    for x, u in points(metric, 20):
        if metric == A.PRR_KERR_BL:
            y, EL = api.state_from_xu(metric, x, u, 1.0, 0.9)
            f64 = api.rhs(metric, y, 1.0, 0.9, EL[0], EL[1])
            f32 = api.rhs(metric, y, 1.0, 0.9, EL[0], EL[1], fp32=True)
        else:
            y = np.concatenate([x, u])
            f64, f32 = api.rhs(metric, y), api.rhs(metric, y, fp32=True)
        assert np.allclose(f32, f64, rtol=2e-4, atol=2e-5 * np.abs(f64).max())


@pytest.mark.parametrize("a", [0.0, 0.7, 0.95])
def test_kerr_rhs_vjp_matches_central_differences(a):
    """The adjoint kernel's hand-written reverse mode of the Kerr RHS."""
    for _ in range(40):
        x = np.array([RNG.uniform(2.0, 50.0), RNG.uniform(0.3, math.pi - 0.3), RNG.uniform(0, 6), 0.0])
        u = RNG.normal(size=4)
        y, EL = api.state_from_xu(A.PRR_KERR_BL, x, u, 1.0, a)
        fb = RNG.normal(size=5)
        yb, cb = api.rhs_vjp(A.PRR_KERR_BL, y, fb, 1.0, a, EL[0], EL[1])

        def F(yy, E, L, aa):
            return fb @ api.rhs(A.PRR_KERR_BL, yy, 1.0, aa, E, L)

        # central differences: truncation ~h^2, rounding ~1e-16 |F| / h
        scaleF = np.abs(fb) @ np.abs(api.rhs(A.PRR_KERR_BL, y, 1.0, a, EL[0], EL[1]))
        for i in range(5):
            h = 1e-6 * max(1.0, abs(y[i]))
            yp, ym = y.copy(), y.copy()
            yp[i] += h
            ym[i] -= h
            fd = (F(yp, *EL, a) - F(ym, *EL, a)) / (2 * h)
            tol = 1e-6 * (abs(fd) + 1e-3 * np.abs(yb).max()) + 1e-14 * scaleF / h
            assert abs(yb[i] - fd) <= tol, (i, yb[i], fd)
        for j, (dE, dL, da) in enumerate(((1, 0, 0), (0, 1, 0), (0, 0, 1))):
            h = 1e-6
            fd = (F(y, EL[0] + h * dE, EL[1] + h * dL, a + h * da) -
                  F(y, EL[0] - h * dE, EL[1] - h * dL, a - h * da)) / (2 * h)
            tol = 1e-6 * (abs(fd) + 1e-3 * np.abs(cb).max()) + 1e-14 * scaleF / h
            assert abs(cb[j] - fd) <= tol, (j, cb[j], fd)
