"""Pins of the oracle's Ellis wormhole (CPU, no GPU): the metric that makes
the reference's second sky reachable (bimage2, "l < 0 if wormhole",
main/base.h:168; the reference has the texture slot but no metric that
crosses to l < 0).

ds^2 = -dt^2 + dl^2 + (b^2 + l^2) dOmega^2.  A null ray with impact
parameter beta = L / E has (dl/dlambda)^2 = E^2 (1 - beta^2 / R^2),
R^2 = b^2 + l^2, so an ingoing ray crosses the throat exactly when beta < b, and its
orbital angle is a closed form in elliptic integrals of the first kind
(with s = beta / R, resp. b / R, along the path):
  beta > b:  int_R^inf beta dR' / sqrt((R'^2 - b^2)(R'^2 - beta^2)) = F(asin(beta / R) | (b / beta)^2)
  beta < b:  the same integral = (beta / b) F(asin(b / R) | (beta / b)^2).
These are checked against the oracle's second-order geodesic integration
(RKF45 at tol 1e-12), which the finite-difference Christoffel test pins term
by term (test_oracle_physics).
"""
import math

import numpy as np
from scipy import special

import oracle as O
from parrelray_amd import _abi as A
from parrelray_amd import configs as CF

RNG = np.random.default_rng(1973)
B = 2.0        # throat radius
R_ESC = 60.0   # |l| at escape


def tracer():
    return A.make_tracer(solver=A.PRR_RKF45, h=0.01, hmin=1e-9, hmax=2.0, atol=1e-12, rtol=1e-12, r_escape=R_ESC,
                         horizon_eps_rel=1e-3, horizon_eps_abs=0.0, lambda_max=1e9, max_steps=10**7,
                         max_attempts=10**8)


def unit(v):
    return v / np.linalg.norm(v)


def impact_and_energy(x, u):
    """beta = L / E from a state: E = u^t (g_tt = -1), L = R^2 |angular rate|."""
    R2 = B * B + x[0] * x[0]
    L = R2 * math.sqrt(u[1] ** 2 + math.sin(x[1]) ** 2 * u[2] ** 2)
    return L / u[3], u[3], L


def swept_angle(beta, r_cam, ingoing):
    """Orbital angle from the camera (areal radius r_cam, l > 0) to escape at |l| = R_ESC."""
    R_end = math.sqrt(B * B + R_ESC * R_ESC)
    if beta > B:  # turns at R = beta, escapes on the near side
        m = (B / beta) ** 2
        phi = lambda R: special.ellipkinc(math.asin(min(beta / R, 1.0)), m)  # noqa: E731
        full = special.ellipk(m)
        return (2 * full - phi(r_cam) - phi(R_end)) if ingoing else (phi(r_cam) - phi(R_end))
    # no turning point: ingoing rays cross the throat R = b and escape on
    # the far side, outgoing ones escape directly
    m = (beta / B) ** 2
    psi = lambda R: (beta / B) * special.ellipkinc(math.asin(min(B / R, 1.0)), m)  # noqa: E731
    if not ingoing:
        return psi(r_cam) - psi(R_end)
    return 2 * (beta / B) * special.ellipk(m) - psi(r_cam) - psi(R_end)


def rays(n, r_cam):
    """Cameras at areal radius r_cam, directions in orbital planes of any
    inclination: polar ones cross the chart's axis (theta = 0, pi), where the
    state is mapped to the same point and direction in the other half-chart."""
    out = []
    for k in range(n):
        ph0 = RNG.uniform(0, 2 * math.pi)
        p = r_cam * np.array([math.cos(ph0), math.sin(ph0), 0.0])
        t = np.array([-math.sin(ph0), math.cos(ph0), 0.0])
        inc = RNG.uniform(-0.5 * math.pi, 0.5 * math.pi) if k % 4 else 0.5 * math.pi  # some exactly polar
        t = math.cos(inc) * t + math.sin(inc) * np.array([0.0, 0.0, 1.0])  # in-plane tangent
        # angle from the inward radial direction; every other ray within the
        # narrow cone of the throat (beta < b needs sin psi < ~b / r_cam)
        psi = RNG.uniform(0.0, math.pi) if k % 2 else RNG.uniform(0.0, 0.2)
        d = -math.cos(psi) * unit(p) + math.sin(psi) * t
        out.append((p, d, t))
    return out


def test_throat_crossing_and_orbital_angle_follow_the_closed_forms():
    S = O.Scene(A.PRR_ELLIS, 0.0, B, tracer())
    r_cam = 15.0
    cases = rays(160, r_cam)
    y0 = []
    for p, d, _ in cases:
        x, u = O.init_ray(A.PRR_ELLIS, 0.0, B, p, d)
        y0.append(np.concatenate([x, u]))
    y0 = np.array(y0)
    out = S.trace(y0)
    crossed = checked = 0
    for (p, d, t), y, cls, ye in zip(cases, y0, out["cls"], out["xend"]):
        beta, E, L = impact_and_energy(y[:4], y[4:])
        if abs(beta / B - 1.0) < 0.02:  # near-critical rays wind; not a closed-form case here
            continue
        assert cls == A.PRR_ESCAPE
        # the side of escape is fixed by the impact parameter and the sense of l'
        ingoing = (d @ p) < 0
        assert (ye[0] < 0) == (beta < B and ingoing), (beta, ingoing, ye[0])
        assert abs(abs(ye[0]) - R_ESC) < 1e-9 * R_ESC
        crossed += ye[0] < 0
        # E and L are conserved, the end state is null
        b_end, E_end, L_end = impact_and_energy(ye[:4], ye[4:])
        assert abs(E_end - E) < 1e-12 * E
        # (an absolute error of ~atol in the angular rate at R^2 ~ 3600 is ~4e-9 in L)
        assert abs(L_end - L) < 1e-8 * (L + E * B)
        g = O.metric(A.PRR_ELLIS, 0.0, B, ye[:4])
        assert abs(ye[4:] @ g @ ye[4:]) < 1e-9 * E * E
        # the end direction in the orbital plane (e1 = camera direction, e2 = the
        # direction's tangential part) is at the closed-form swept angle
        e1 = unit(p)
        e2 = unit(d - (d @ e1) * e1)
        ang = swept_angle(beta, r_cam, ingoing)
        th, ph = ye[1], ye[2]
        n = np.array([math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th)])
        expect = math.cos(ang) * e1 + math.sin(ang) * e2
        assert np.abs(n - expect).max() < 1e-7, (beta, ang, n, expect)
        checked += 1
    assert checked >= 120 and crossed >= 40


def test_radial_ray_passes_straight_through_to_the_far_sky():
    """beta = 0: the ray keeps its angles and reaches l = -r_escape."""
    S = O.Scene(A.PRR_ELLIS, 0.0, B, tracer())
    for _ in range(8):
        p = unit(RNG.normal(size=3) * np.array([1.0, 1.0, 0.5])) * RNG.uniform(5.0, 40.0)
        x, u = O.init_ray(A.PRR_ELLIS, 0.0, B, p, -unit(p))
        o = S.trace(np.concatenate([x, u])[None])
        assert o["cls"][0] == A.PRR_ESCAPE
        ye = o["xend"][0]
        assert abs(ye[0] + R_ESC) < 1e-9 * R_ESC
        dph = (ye[2] - x[2] + math.pi) % (2 * math.pi) - math.pi
        assert abs(ye[1] - x[1]) < 1e-12 and abs(dph) < 1e-12


def test_far_side_has_its_own_escape_radius():
    """rout2: rays through the throat escape at l = -r_esc2, the others at
    l = r_escape; located on the sphere in both cases."""
    S = O.Scene(A.PRR_ELLIS, 0.0, B, tracer())
    S.set_far_escape(35.0)
    p = np.array([11.0, 4.0, -3.0])
    through = [-unit(p), unit(-unit(p) + 0.05 * unit(np.cross(p, [0.0, 0.0, 1.0])))]
    back = [unit(p), unit(-unit(p) + 0.6 * unit(np.cross(p, [0.0, 0.0, 1.0])))]
    y0 = np.array([np.concatenate(O.init_ray(A.PRR_ELLIS, 0.0, B, p, d)) for d in through + back])
    o = S.trace(y0)
    assert np.all(o["cls"] == A.PRR_ESCAPE)
    assert np.allclose(o["xend"][:2, 0], -35.0, rtol=1e-12)
    assert np.allclose(o["xend"][2:, 0], R_ESC, rtol=1e-12)
    S.set_far_escape(0.0)  # 0: both sides at r_escape
    assert np.allclose(S.trace(y0)["xend"][:2, 0], -R_ESC, rtol=1e-12)


def test_init_ray_converts_areal_radius_to_proper_distance():
    """A camera at areal radius r sits at l = sqrt(r^2 - b^2) > 0, and the
    state is null with u^t > 0 (the flat-space coordinate velocity in the
    areal-radius chart, as for the Schwarzschild-coordinate metric)."""
    for _ in range(20):
        p = unit(RNG.normal(size=3)) * RNG.uniform(2.5, 50.0)
        d = unit(RNG.normal(size=3))
        x, u = O.init_ray(A.PRR_ELLIS, 0.0, B, p, d)
        r = np.linalg.norm(p)
        assert abs(x[0] - math.sqrt(r * r - B * B)) < 1e-12 * r
        # dl = dr r / l with dr the radial component of d
        assert abs(u[0] - (d @ p / r) * r / x[0]) < 1e-12 * (abs(u[0]) + 1)
        g = O.metric(A.PRR_ELLIS, 0.0, B, x)
        assert abs(u @ g @ u) < 1e-12 * (u @ u) and u[3] > 0


def test_far_side_rays_take_the_second_sky():
    """Rendered colour: texture slot 0 for rays escaping at l > 0, the second
    sky (slot PRR_SKY2_SLOT) for rays through the throat; no horizon."""
    S = O.Scene(A.PRR_ELLIS, 0.0, B, tracer())
    red = np.zeros((4, 8, 3), np.uint8)
    red[..., 0] = 200
    blue = np.zeros((4, 8, 3), np.uint8)
    blue[..., 2] = 180
    S.set_texture(0, red)
    S.set_texture(A.PRR_SKY2_SLOT, blue)
    p, al, be = CF.orbit_pose(12.5, 80.0, 0.3)  # looking at the throat
    f = A.make_frame(p, al, be, math.radians(60.0), 24, 16, mass=0.0)
    r = S.render_rays(f, want_state=True)
    assert np.all(r["cls"] == A.PRR_ESCAPE)
    far = r["state8"][:, 0] < 0
    assert far.sum() >= 10 and (~far).sum() >= 10
    assert np.allclose(r["color"][far], [0.0, 0.0, 180 / 255], atol=1e-6)
    assert np.allclose(r["color"][~far], [200 / 255, 0.0, 0.0], atol=1e-6)
