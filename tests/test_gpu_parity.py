"""GPU parity: the HIP integrator and AA passes against the CPU oracle.

All tests call libprr.so through the C ABI (parrelray_amd.api) on cuda:0 and
compare with oracle/ on the same seeded inputs.
"""
import math

import numpy as np
import pytest

import oracle as O
from parrelray_amd import _abi as A
from parrelray_amd import api
from parrelray_amd import configs as CF

from parity import angle, boundary_excused, compare_rays, gpu_context, oracle_scene, rgb8_agreement

pytestmark = pytest.mark.gpu
RNG = np.random.default_rng(11)


def random_states(metric, n, a=0.0):
    out = []
    for _ in range(n):
        pos = RNG.normal(size=3)
        pos = pos / np.linalg.norm(pos) * RNG.uniform(8.0, 30.0)
        d = -pos / np.linalg.norm(pos) + 0.6 * RNG.normal(size=3)
        d /= np.linalg.norm(d)
        x, u = O.init_ray(metric, 1.0, a, pos, d)
        out.append(np.concatenate([x, u]))
    return np.array(out)


def trace_tracer(solver):
    if solver == A.PRR_EULER:
        return A.make_tracer(solver=solver, h=0.01, r_escape=60.0, horizon_eps_rel=1e-3, horizon_eps_abs=0.0,
                             lambda_max=1e9, max_steps=3000, max_attempts=3000)
    if solver == A.PRR_RK4:
        return A.make_tracer(solver=solver, h=0.02, r_escape=60.0, horizon_eps_rel=1e-3, horizon_eps_abs=0.0,
                             lambda_max=1e9, max_steps=20000, max_attempts=20000)
    return A.make_tracer(solver=solver, h=0.05, hmin=1e-9, hmax=5.0, atol=1e-10, rtol=1e-10, r_escape=60.0,
                         horizon_eps_rel=1e-3, horizon_eps_abs=0.0, lambda_max=1e9, max_steps=20000,
                         max_attempts=200000)


@pytest.mark.parametrize("metric", [A.PRR_FLAT, A.PRR_SCHW_KS_CART, A.PRR_SCHW_SPH])
@pytest.mark.parametrize("solver", [A.PRR_EULER, A.PRR_RKF45, A.PRR_DORMAND, A.PRR_RK4])
def test_trace_second_order_metrics_match_oracle(metric, solver):
    """Same algorithm, same state: classes agree; fixed-step (Euler, RK4)
    final states agree to rounding; adaptive (RKF45, Dormand-Prince) escape
    states (located at r = r_escape) agree to the tolerance.  Adaptive step
    sequences may differ by rounding of the error estimate, so their other
    end states (first state inside the horizon, step cap) are not compared."""
    T = trace_tracer(solver)
    x0 = random_states(metric, 512)
    ctx = api.Context(0)
    ctx.set_spacetime(metric, 1.0, 0.0)
    ctx.set_tracer(T)
    g = ctx.trace(x0)
    S = O.Scene(metric, 1.0, 0.0, T)
    o = S.trace(x0)
    agree = g["cls"] == o["cls"]
    assert agree.mean() >= 0.99, np.nonzero(~agree)
    fixed = solver in (A.PRR_EULER, A.PRR_RK4)
    m = agree if fixed else agree & (o["cls"] == A.PRR_ESCAPE)
    assert m.sum() >= 50
    err = np.abs(g["xend"][m] - o["xend"][m]).max(axis=1) / (np.abs(o["xend"][m]).max(axis=1) + 1)
    if fixed:
        assert np.median(err) < 1e-12
        assert (err < 1e-6).mean() >= 0.99, np.sort(err)[-10:]
        assert (g["steps"][agree] == o["steps"][agree]).mean() >= 0.97
    else:
        assert np.median(err) < 1e-8, np.sort(err)[-10:]
        assert (err < 1e-6).mean() >= 0.98, np.sort(err)[-10:]


@pytest.mark.parametrize("solver", [A.PRR_EULER, A.PRR_RKF45])
def test_trace_reissner_nordstrom_matches_oracle(solver):
    """The Schwarzschild-coordinate metric with charge q = 0.7 (the
    spacetime's q): classes and end states as for the uncharged metric."""
    q = 0.7
    T = trace_tracer(solver)
    x0 = random_states_charged(256, q)
    ctx = api.Context(0)
    ctx.set_spacetime(A.PRR_SCHW_SPH, 1.0, 0.0, q)
    ctx.set_tracer(T)
    g = ctx.trace(x0)
    o = O.Scene(A.PRR_SCHW_SPH, 1.0, q, T).trace(x0)
    agree = g["cls"] == o["cls"]
    assert agree.mean() >= 0.99
    assert (o["cls"] == A.PRR_HORIZON).sum() >= 10
    if solver != A.PRR_EULER:  # (3000 Euler steps of 0.01 M do not reach r_escape = 60 M)
        assert (o["cls"] == A.PRR_ESCAPE).sum() >= 10
    m = agree if solver == A.PRR_EULER else agree & (o["cls"] == A.PRR_ESCAPE)
    err = np.abs(g["xend"][m] - o["xend"][m]).max(axis=1) / (np.abs(o["xend"][m]).max(axis=1) + 1)
    assert np.median(err) < (1e-12 if solver == A.PRR_EULER else 1e-8)
    # captured rays stop at the charged horizon r+ = M + sqrt(M^2 - q^2), not at 2M
    hz = agree & (o["cls"] == A.PRR_HORIZON)
    rp = 1.0 + math.sqrt(1.0 - q * q)
    assert np.all(g["xend"][hz, 0] <= rp * (1 + 1e-3) + 1e-9) and np.all(g["xend"][hz, 0] > rp * 0.99)


def random_states_charged(n, q):
    out = []
    for _ in range(n):
        pos = RNG.normal(size=3)
        pos = pos / np.linalg.norm(pos) * RNG.uniform(8.0, 30.0)
        d = -pos / np.linalg.norm(pos) + 0.6 * RNG.normal(size=3)
        d /= np.linalg.norm(d)
        x, u = O.init_ray(A.PRR_SCHW_SPH, 1.0, q, pos, d)
        out.append(np.concatenate([x, u]))
    return np.array(out)


ELLIS_B = 2.0  # throat radius


def wormhole_states(n):
    """Cameras at areal radius 8-30 looking roughly at the throat: a mix of
    rays through it (impact parameter < b) and rays deflected back."""
    out = []
    for _ in range(n):
        pos = RNG.normal(size=3)
        pos = pos / np.linalg.norm(pos) * RNG.uniform(8.0, 30.0)
        d = -pos / np.linalg.norm(pos) + 0.15 * RNG.normal(size=3)
        d /= np.linalg.norm(d)
        x, u = O.init_ray(A.PRR_ELLIS, 0.0, ELLIS_B, pos, d)
        out.append(np.concatenate([x, u]))
    return np.array(out)


def wormhole_context(T, r_esc2=0.0):
    ctx = api.Context(0)
    ctx.set_spacetime(A.PRR_ELLIS, 0.0, 0.0, 0.0, p=[ELLIS_B, r_esc2] + [0.0] * 14)
    ctx.set_tracer(T)
    return ctx


@pytest.mark.parametrize("solver", [A.PRR_EULER, A.PRR_RKF45])
def test_trace_wormhole_matches_oracle(solver):
    """Ellis wormhole (throat radius p[0] = 2): classes (escape only, on
    either side of the throat) and end states as for the other metrics."""
    T = trace_tracer(solver)
    if solver == A.PRR_EULER:  # long enough to reach |l| = 60 from l <= 30
        T.max_steps = T.max_attempts = 12000
    x0 = wormhole_states(256)
    g = wormhole_context(T, r_esc2=45.0).trace(x0)  # beyond the throat: escape at |l| = 45 (rout2)
    S = O.Scene(A.PRR_ELLIS, 0.0, ELLIS_B, T)
    S.set_far_escape(45.0)
    o = S.trace(x0)
    agree = g["cls"] == o["cls"]
    assert agree.mean() >= 0.99
    esc = agree & (o["cls"] == A.PRR_ESCAPE)
    assert esc.mean() >= 0.95
    far = esc & (o["xend"][:, 0] < 0)
    assert far.sum() >= 30 and (esc & ~far).sum() >= 30
    # the same side of the throat, and each side's escape radius
    assert np.all((g["xend"][esc, 0] < 0) == (o["xend"][esc, 0] < 0))
    if solver != A.PRR_EULER:  # located escapes
        assert np.allclose(g["xend"][far, 0], -45.0, rtol=1e-12)
        assert np.allclose(g["xend"][esc & ~far, 0], 60.0, rtol=1e-12)
    err = np.abs(g["xend"][esc] - o["xend"][esc]).max(axis=1) / (np.abs(o["xend"][esc]).max(axis=1) + 1)
    assert np.median(err) < (1e-12 if solver == A.PRR_EULER else 1e-8), np.sort(err)[-10:]
    assert (err < 1e-6).mean() >= 0.98, np.sort(err)[-10:]


def test_wormhole_render_shows_the_second_sky():
    """A frame looking into the throat, 2x2 SSAA + FXAA, RKF45 tol 1e-8: rays
    escaping at l < 0 take the second sky (slot PRR_SKY2_SLOT); per-ray
    classes and endpoints against the oracle, RGB8 within 1 LSB."""
    pos, al, be = CF.orbit_pose(12.0, 75.0, 0.4)
    f = A.make_frame(pos, al, be, math.radians(30.0), 96, 64, ssaa=A.PRR_SSAA_REGULAR, samples=4, fxaa=1,
                     mass=0.0)
    T = CF.tracer_kerr()
    T.r_escape = 60.0
    ctx = wormhole_context(T)
    ctx.set_texture(0, CF.sky_texture(), A.PRR_FILTER_SQAVG)
    ctx.set_texture(A.PRR_SKY2_SLOT, CF.sphere_texture(), A.PRR_FILTER_SQAVG)
    out = ctx.render(f, rays=True)[0]
    # the same tracer (tol 1e-8): nothing re-imposes first integrals here, so
    # the GPU's endpoints are compared with the oracle's at the same tolerance
    S = O.Scene(A.PRR_ELLIS, 0.0, ELLIS_B, T)
    S.set_texture(0, CF.sky_texture(), A.PRR_FILTER_SQAVG)
    S.set_texture(A.PRR_SKY2_SLOT, CF.sphere_texture(), A.PRR_FILTER_SQAVG)
    orc = S.render_rays(f, want_state=True)
    assert np.all(orc["cls"] == A.PRR_ESCAPE)
    far = orc["state8"][:, 0] < 0
    assert 0.05 < far.mean() < 0.95
    cmp = compare_rays(S, f, out, orc)
    assert cmp["unexcused"] == 0, cmp["unexcused_ids"][:20]
    ang = cmp["angles"]
    print("endpoint angle: median %.2e p99 %.2e max %.2e" % (np.median(ang), np.quantile(ang, 0.99), ang.max()))
    assert np.median(ang) < 1e-9 and np.quantile(ang, 0.99) < 1e-7
    # the colour of each ray comes from the sky of its own side
    assert np.abs(out["color"] - orc["color"]).max(axis=1).mean() < 1e-3
    _, img, rgb8 = S.render_frame(f)
    frac, _ = rgb8_agreement(out["rgb8"], rgb8)
    assert frac >= 0.999
    # without the second texture the far side is black
    ctx.set_texture(A.PRR_SKY2_SLOT, None)
    dark = ctx.render(f, rays=True)[0]
    assert np.all(dark["color"][far] == 0) and np.array_equal(dark["color"][~far], out["color"][~far])


@pytest.mark.parametrize("a,solver", [(0.0, A.PRR_RKF45), (0.6, A.PRR_RKF45), (0.95, A.PRR_RKF45),
                                      (0.95, A.PRR_DORMAND)])
def test_trace_kerr_conserves_constants_and_matches_oracle(a, solver):
    T = trace_tracer(solver)
    x0 = random_states(A.PRR_KERR_BL, 256, a)
    ctx = api.Context(0)
    ctx.set_spacetime(A.PRR_KERR_BL, 1.0, a)
    ctx.set_tracer(T)
    g = ctx.trace(x0)
    S = O.Scene(A.PRR_KERR_BL, 1.0, a, A.make_tracer(**{**{k: getattr(T, k) for k, _ in A.Tracer._fields_},
                                                         "atol": 1e-12, "rtol": 1e-12}))
    o = S.trace(x0)
    agree = g["cls"] == o["cls"]
    assert agree.mean() >= 0.98
    esc = agree & (o["cls"] == A.PRR_ESCAPE)
    # Carter constant drift of the GPU's final state
    for i in np.nonzero(esc)[0][:50]:
        c0 = O.kerr_constants(1.0, a, x0[i, :4], x0[i, 4:])
        c1 = O.kerr_constants(1.0, a, g["xend"][i, :4], g["xend"][i, 4:])
        assert abs(c1[0] - c0[0]) < 1e-12 * abs(c0[0])  # E is exact by construction
        assert abs(c1[1] - c0[1]) < 1e-12 * (abs(c0[1]) + 1)
        assert abs(c1[2] - c0[2]) < 1e-6 * (abs(c0[2]) + 1)


@pytest.mark.parametrize("a", [0.95, 0.998])
def test_kerr_spherical_photon_orbits_keep_their_radius_on_the_gpu(a):
    """The shadow-defining spherical photon orbits (Teo 2003) through the
    GPU's Mino-time integrator: started on the equator with u^r = 0 and the
    orbit's constants, the ray is still at r_p when it reaches the affine cap
    lambda = 10 (the orbit is unstable; the oracle at tol 1e-12 drifts ~3e-8),
    with theta between the analytic turning points of Theta."""
    T = A.make_tracer(solver=A.PRR_RKF45, h=0.01, hmin=1e-12, hmax=0.05, atol=1e-12, rtol=1e-12, r_escape=200.0,
                      horizon_eps_rel=1e-3, horizon_eps_abs=0.0, lambda_max=10.0, max_steps=10**6,
                      max_attempts=10**7)
    r_pro = 2 * (1 + math.cos(2.0 / 3.0 * math.acos(-a)))
    r_ret = 2 * (1 + math.cos(2.0 / 3.0 * math.acos(a)))
    rows, radii, th_min = [], [], []
    for f in (0.2, 0.4, 0.6, 0.8):
        r = r_pro + f * (r_ret - r_pro)
        xi = (r * r * (3 - r) - a * a * (r + 1)) / (a * (r - 1))
        eta = r ** 3 * (4 * a * a - r * (r - 3) ** 2) / (a * a * (r - 1) ** 2)
        # theta turning points: Theta = eta + a^2 c^2 - xi^2 c^2 / (1 - c^2) = 0, c = cos th
        B2 = a * a - eta - xi * xi
        c2 = (-B2 - math.sqrt(B2 * B2 + 4 * a * a * eta)) / (-2 * a * a)
        th_min.append(math.acos(math.sqrt(c2)))
        g = O.metric(A.PRR_KERR_BL, 1.0, a, [r, math.pi / 2, 0.0, 0.0])
        gi = np.linalg.inv(g)
        u = [0.0, math.sqrt(eta) / (r * r), gi[2, 3] * -1.0 + gi[2, 2] * xi, gi[3, 3] * -1.0 + gi[3, 2] * xi]
        rows.append([r, math.pi / 2, 0.0, 0.0] + u)
        radii.append(r)
    ctx = api.Context(0)
    ctx.set_spacetime(A.PRR_KERR_BL, 1.0, a)
    ctx.set_tracer(T)
    g = ctx.trace(np.array(rows))
    assert np.all(g["cls"] == A.PRR_STEP_CAP)
    dr = np.abs(g["xend"][:, 0] - np.array(radii)) / np.array(radii)
    print("a", a, "relative radius drift", dr, "theta", g["xend"][:, 1], "theta_min", th_min)
    assert dr.max() < 1e-6
    # theta stays between its turning points (and has left the equator)
    th = g["xend"][:, 1]
    assert np.all(th >= np.array(th_min) - 1e-9) and np.all(th <= math.pi - np.array(th_min) + 1e-9)
    assert np.all(np.abs(th - math.pi / 2) > 1e-3)


def test_c0_parity_full_frame():
    T = CF.tracer_c0()
    f = CF.frame_c0()
    ctx = gpu_context(A.PRR_SCHW_SPH, T)
    out = ctx.render(f, rays=True)[0]
    S = oracle_scene(A.PRR_SCHW_SPH, T)
    orc = S.render_rays(f)
    cmp = compare_rays(S, f, out, orc)
    assert cmp["unexcused"] == 0, cmp["unexcused_ids"][:20]
    # endpoints of rays stopped by the step cap: final positions
    same = cmp["same"]
    d = np.linalg.norm(out["endpoint"][same] - orc["endpoint"][same], axis=1)
    r = np.linalg.norm(orc["endpoint"][same], axis=1)
    assert np.quantile(d / r, 0.999) < 1e-6
    _, img, rgb8 = S.render_frame(f)
    frac, _ = rgb8_agreement(out["rgb8"], rgb8)
    assert frac >= 0.999


@pytest.mark.parametrize("ssaa,samples", [(A.PRR_SSAA_RANDOM, 3), (A.PRR_SSAA_JITTERED, 4)])
def test_random_and_jittered_ssaa_match_oracle(ssaa, samples):
    """The counter-based sample stream gives the same sample positions on the
    GPU as in the oracle: per-ray classes and endpoints agree as for the
    regular grid (C1 scene at 96x54, FP64 RKF45 tol 1e-8)."""
    f = CF.frame_c1(96, 54)
    f.ssaa, f.samples, f.seed = ssaa, samples, 2024
    ctx = gpu_context(A.PRR_KERR_BL, CF.tracer_kerr())
    out = ctx.render(f, rays=True)[0]
    S = oracle_scene(A.PRR_KERR_BL, CF.tracer_kerr(tol=1e-12))
    orc = S.render_rays(f)
    cmp = compare_rays(S, f, out, orc)
    assert cmp["unexcused"] <= 1, cmp["unexcused_ids"][:20]
    assert np.median(cmp["angles"]) < 1e-6
    _, img, rgb8 = S.render_frame(f)
    frac, _ = rgb8_agreement(out["rgb8"], rgb8)
    assert frac >= 0.999
    # another seed moves the samples, and so the image
    f.seed = 2025
    other = ctx.render(f, rays=False)[0]["rgb8"]
    assert not np.array_equal(other, out["rgb8"])


@pytest.mark.parametrize("precision", [A.PRR_FP64, A.PRR_FP32_BULK])
def test_c4_reduced_parity(precision):
    """The precision-stress scene (Kerr a = 0.998, camera r = 6M edge-on at
    89.9 deg, tol 1e-10, photon-ring rays of ~10^5 steps) at 96x96 without
    SSAA, against the FP64 tol-1e-12 oracle: classes agree outside the
    excused boundary rays; endpoint angle median < 1e-7 rad and 99th
    percentile < 1e-5 rad; RGB8 within 1 LSB on >= 99.9 % of pixels."""
    f = CF.frame_c4(96, 96, ssaa=1)
    ctx = gpu_context(A.PRR_KERR_BL, CF.tracer_c4(precision=precision))
    out = ctx.render(f, rays=True)[0]  # the frame carries the spin
    S = oracle_scene(A.PRR_KERR_BL, CF.tracer_kerr(tol=1e-12, max_steps=400000))
    orc = S.render_rays(f)
    cmp = compare_rays(S, f, out, orc)
    assert cmp["unexcused"] <= 2, cmp["unexcused_ids"][:20]
    ang = cmp["angles"]
    print("endpoint angle: median %.2e p99 %.2e max %.2e" % (np.median(ang), np.quantile(ang, 0.99), ang.max()))
    assert np.median(ang) < 1e-7
    assert np.quantile(ang, 0.99) < 1e-5
    _, img, rgb8 = S.render_frame(f)
    frac, _ = rgb8_agreement(out["rgb8"], rgb8)
    assert frac >= 0.999


@pytest.mark.parametrize("precision,solver", [(A.PRR_FP64, A.PRR_RKF45), (A.PRR_FP32_BULK, A.PRR_RKF45),
                                              (A.PRR_FP32_BULK, A.PRR_DORMAND)])
def test_c1_reduced_parity(precision, solver):
    """C1 at 192x108, 2x2 SSAA + FXAA, tol 1e-8, against the FP64 tol-1e-12 oracle.

    Endpoint tolerance at this integration tolerance: median < 1e-6 rad,
    99th percentile < 1e-5 rad (both forms re-impose Kerr's first integrals
    in FP64 after every step).  Dormand-Prince 5(4) with the same controller
    meets the same bounds."""
    f = CF.frame_c1(192, 108)
    T = CF.tracer_kerr(precision=precision)
    T.solver = solver
    ctx = gpu_context(A.PRR_KERR_BL, T)
    out = ctx.render(f, rays=True)[0]
    S = oracle_scene(A.PRR_KERR_BL, CF.tracer_kerr(tol=1e-12))
    col, img, rgb8 = S.render_frame(f)
    orc = S.render_rays(f)
    cmp = compare_rays(S, f, out, orc)
    assert cmp["unexcused"] <= 2, cmp["unexcused_ids"][:20]
    ang = cmp["angles"]
    print("endpoint angle: median %.2e p99 %.2e max %.2e" % (np.median(ang), np.quantile(ang, 0.99), ang.max()))
    assert np.median(ang) < 1e-6
    assert np.quantile(ang, 0.99) < 1e-5
    frac, _ = rgb8_agreement(out["rgb8"], rgb8)
    assert frac >= 0.999


def multi_sphere_frame(W=128, H=72, spin=0.9):
    """Four spheres: two overlapping on the sky (one behind the other), an
    untextured one, one close to the hole; camera r = 20M, inclination 80 deg."""
    pos, al, be = CF.orbit_pose(20.0, 80.0)
    spheres = [CF.equatorial_sphere(8.0, 0.3, 1.0, 1), CF.equatorial_sphere(11.0, 0.22, 1.2, 1),
               CF.equatorial_sphere(6.0, -0.45, 0.8, 0), ([-2.0, 4.5, 0.8], 0.7, 1)]
    return A.make_frame(pos, al, be, math.radians(50.0), W, H, mass=1.0, spin=spin, spheres=spheres)


@pytest.mark.parametrize("precision", [A.PRR_FP64, A.PRR_FP32_BULK])
def test_multiple_spheres_match_oracle(precision):
    """Several spheres in one scene (the reference's sphereN blocks): the
    nearest located hit along the path wins, per-sphere textures (slot 0: white);
    classes, endpoints and RGB8 against the oracle (Kerr a = 0.9, RKF45)."""
    f = multi_sphere_frame()
    T = CF.tracer_kerr(precision=precision)
    ctx = gpu_context(A.PRR_KERR_BL, T)
    out = ctx.render(f, rays=True)[0]
    S = oracle_scene(A.PRR_KERR_BL, CF.tracer_kerr(tol=1e-12))
    orc = S.render_rays(f)
    hits = orc["cls"] == A.PRR_SPHERE
    assert hits.sum() >= 300
    cmp = compare_rays(S, f, out, orc)
    assert cmp["unexcused"] <= 1, cmp["unexcused_ids"][:20]
    ang = cmp["angles"]
    print("endpoint angle: median %.2e p99 %.2e max %.2e" % (np.median(ang), np.quantile(ang, 0.99), ang.max()))
    assert np.median(ang) < 1e-6 and np.quantile(ang, 0.99) < 1e-5
    # every sphere is seen, by the colour of its texture or its white
    same = cmp["same"] & hits
    assert np.abs(out["color"][same] - orc["color"][same]).max(axis=1).mean() < 1e-3
    _, img, rgb8 = S.render_frame(f)
    frac, _ = rgb8_agreement(out["rgb8"], rgb8)
    assert frac >= 0.999


def test_multiple_spheres_euler_chords_match_oracle():
    """The Euler solver's chord test over several spheres (the first sphere
    that a step chord enters, scene/sphere.h:34-50 with D11): Schwarzschild
    coordinates, fixed steps, so classes and end points agree to rounding."""
    pos, al, be = CF.orbit_pose(20.0, 80.0)
    spheres = [CF.equatorial_sphere(8.0, 0.3, 1.0, 1), CF.equatorial_sphere(11.0, 0.22, 1.2, 1),
               CF.equatorial_sphere(6.0, -0.45, 0.8, 0)]
    f = A.make_frame(pos, al, be, math.radians(50.0), 96, 54, mass=1.0, spin=0.0, spheres=spheres)
    T = A.make_tracer(solver=A.PRR_EULER, h=0.01, r_escape=60.0, horizon_eps_rel=1e-3, horizon_eps_abs=0.0,
                      lambda_max=1e9, max_steps=12000, max_attempts=12000)
    ctx = gpu_context(A.PRR_SCHW_SPH, T)
    out = ctx.render(f, rays=True)[0]
    S = oracle_scene(A.PRR_SCHW_SPH, T)
    orc = S.render_rays(f)
    assert (orc["cls"] == A.PRR_SPHERE).sum() >= 200
    cmp = compare_rays(S, f, out, orc)
    assert cmp["unexcused"] == 0, cmp["unexcused_ids"][:20]
    assert np.median(cmp["angles"]) < 1e-9
    _, img, rgb8 = S.render_frame(f)
    frac, _ = rgb8_agreement(out["rgb8"], rgb8)
    assert frac >= 0.999


RENDER_VARIANTS = {
    # metric, solver, projection, samples per pixel, fxaa, spin
    "ks_cart_rkf45_ssaa2": (A.PRR_SCHW_KS_CART, A.PRR_RKF45, A.PRR_PROJ_PINHOLE, 4, 0, 0.0),
    "sph_euler_spherical_projection": (A.PRR_SCHW_SPH, A.PRR_EULER, A.PRR_PROJ_SPHERICAL, 1, 0, 0.0),
    "kerr_rkf45_ssaa3_fxaa": (A.PRR_KERR_BL, A.PRR_RKF45, A.PRR_PROJ_PINHOLE, 9, 1, 0.6),
    "flat_dormand_spherical_projection": (A.PRR_FLAT, A.PRR_DORMAND, A.PRR_PROJ_SPHERICAL, 4, 1, 0.0),
}


@pytest.mark.parametrize("name", sorted(RENDER_VARIANTS))
def test_render_variants_match_oracle(name):
    """Frames over the other metrics, solvers, projections and SSAA factors
    (camera r = 20M, inclination 80 deg, the C1 sphere): classes, endpoints
    and RGB8 against the oracle on the same frame."""
    metric, solver, proj, s, fxaa, spin = RENDER_VARIANTS[name]
    pos, al, be = CF.orbit_pose(20.0, 80.0)
    n = int(round(math.sqrt(s)))
    f = A.make_frame(pos, al, be, math.radians(50.0), 64, 40, ssaa=A.PRR_SSAA_REGULAR if s > 1 else A.PRR_SSAA_NONE,
                     samples=s, projection=proj, fxaa=fxaa, mass=1.0, spin=spin,
                     spheres=[CF.equatorial_sphere(8.0, 0.3)])
    assert n * n == s
    T = trace_tracer(solver)
    if solver == A.PRR_EULER:
        T.max_steps = T.max_attempts = 12000
    ctx = gpu_context(metric, T)
    out = ctx.render(f, rays=True)[0]
    Tq = trace_tracer(solver)
    if solver == A.PRR_EULER:
        Tq.max_steps = Tq.max_attempts = 12000
    else:
        Tq.atol = Tq.rtol = 1e-12
    S = oracle_scene(metric, Tq)
    orc = S.render_rays(f)
    cmp = compare_rays(S, f, out, orc)
    assert cmp["unexcused"] <= 1, cmp["unexcused_ids"][:20]
    ang = cmp["angles"]
    print("%s: classes %s, endpoint angle median %.2e p99 %.2e" % (name, np.bincount(orc["cls"], minlength=5).tolist(),
                                                                    np.median(ang), np.quantile(ang, 0.99)))
    assert np.median(ang) < (1e-9 if solver == A.PRR_EULER else 1e-6)
    _, img, rgb8 = S.render_frame(f)
    frac, _ = rgb8_agreement(out["rgb8"], rgb8)
    assert frac >= 0.999


def test_aa_passes_bit_exact():
    """SSAA combine, FXAA and RGB8 on the GPU's own ray colours equal the oracle's passes bit for bit."""
    f = CF.frame_c1(160, 90)
    ctx = gpu_context(A.PRR_KERR_BL, CF.tracer_kerr())
    out = ctx.render(f, rays=True)[0]
    ss = O.ssaa(out["color"], f.width, f.height, f.samples)
    fx = O.fxaa(ss)
    assert np.array_equal(out["rgbf"], fx)
    assert np.array_equal(out["rgb8"], O.rgb8(fx))
    f.fxaa = 0
    out = ctx.render(f, rays=True)[0]
    ss = O.ssaa(out["color"], f.width, f.height, f.samples)
    assert np.array_equal(out["rgbf"], ss)
    assert np.array_equal(out["rgb8"], O.rgb8(ss))


def test_aa_passes_bit_exact_every_sample_count():
    """The FXAA pass combines each frame's samples itself: one batch of FXAA
    frames with 1, 4, 9 and 16 samples per pixel, the 4-sample frame placed
    after a frame of an odd ray count (its ray colours start off a 16-byte
    boundary), and a frame without FXAA; SSAA, FXAA and RGB8 of the GPU's own
    ray colours equal the oracle's passes bit for bit in every frame."""
    frames = [CF.frame_c1(29, 17, ssaa=1, fxaa=1), CF.frame_c1(40, 24), CF.frame_c1(36, 20, ssaa=3, fxaa=1),
              CF.frame_c1(24, 16, ssaa=4, fxaa=1), CF.frame_c1(21, 13, ssaa=2, fxaa=0)]
    assert [f.samples for f in frames] == [1, 4, 9, 16, 4] and frames[0].nrays % 4 == 1
    ctx = gpu_context(A.PRR_KERR_BL, CF.tracer_kerr())
    outs = ctx.render(frames, rays=True)
    for f, out in zip(frames, outs):
        ss = O.ssaa(out["color"], f.width, f.height, f.samples)
        img = O.fxaa(ss) if f.fxaa else ss
        assert np.array_equal(out["rgbf"], img), f.samples
        assert np.array_equal(out["rgb8"], O.rgb8(img)), f.samples


@pytest.mark.parametrize("width,height", [(1, 1), (1, 7), (7, 1), (2, 2), (33, 17), (65, 34)])
def test_tiny_and_ragged_frames(width, height):
    """Frames of one pixel, one row or column, and sizes that do not divide
    the integrator's 4 x 4-pixel tiles or FXAA's 32 x 32 tiles, all in one
    batch (2 x 2 SSAA + FXAA): every ray's class against the oracle (rays
    within 1e-6 rad of a class boundary excused), and SSAA, FXAA and RGB8 of
    the GPU's own ray colours equal the oracle's passes bit for bit (FXAA's
    taps clamp to the edge of images as small as one pixel)."""
    f = CF.frame_c1(width, height)
    ctx = gpu_context(A.PRR_KERR_BL, CF.tracer_kerr())
    out = ctx.render([f, CF.frame_c1(16, 9)], rays=True)[0]
    assert ctx.stats()["rays"] == f.nrays + 16 * 9 * 4
    ss = O.ssaa(out["color"], f.width, f.height, f.samples)
    fx = O.fxaa(ss)
    assert np.array_equal(out["rgbf"], fx)
    assert np.array_equal(out["rgb8"], O.rgb8(fx))
    S = oracle_scene(A.PRR_KERR_BL, CF.tracer_kerr(tol=1e-12))
    ids = np.arange(f.nrays, dtype=np.int64)
    orc = S.render_rays(f, ids)
    bad = np.nonzero(orc["cls"] != out["cls"])[0]
    assert np.all(boundary_excused(S, f, ids[bad], orc["cls"][bad])), bad


def test_every_ray_processed_once_and_batch_equals_single():
    frames = CF.frames_c2(nframes=10, scale=0.08)
    ctx = gpu_context(A.PRR_KERR_BL, CF.tracer_kerr(precision=A.PRR_FP32_BULK))
    batch = ctx.render(frames, rays=True)
    st = ctx.stats()
    assert st["rays"] == sum(f.nrays for f in frames)
    for f, b in zip(frames, batch):
        single = ctx.render(f, rays=True)[0]
        assert np.array_equal(single["cls"], b["cls"])
        assert np.array_equal(single["rgb8"], b["rgb8"])
        assert np.array_equal(single["rgbf"], b["rgbf"])


@pytest.mark.parametrize("refill,block", [(1, 64), (32, 256), (64, 128)])
def test_launch_shape_does_not_change_results(refill, block):
    f = CF.frame_c1(96, 54)
    ctx = gpu_context(A.PRR_KERR_BL, CF.tracer_kerr(precision=A.PRR_FP32_BULK))
    ref = ctx.render(f, rays=True)[0]
    ctx.set_launch(block=block, blocks_per_cu=1, refill=refill)
    out = ctx.render(f, rays=True)[0]
    assert np.array_equal(ref["rgbf"], out["rgbf"])
    assert np.array_equal(ref["steps"], out["steps"])


@pytest.mark.parametrize("refill", [1, 32])
def test_fp64_partial_refills_batch_equals_single(refill):
    """Plain FP64 with partial refills (a wave takes new rays once `refill`
    lanes are idle: ballot + mbcnt ranks, padding positions between frames):
    a batch holding a frame whose ray count is not a multiple of 64 equals
    the frames rendered alone, bit for bit, and every ray is integrated once."""
    odd = CF.frame_c1(37, 23)  # 851 pixels x 4 samples: 3404 rays
    frames = [odd, CF.frame_c1(64, 36), CF.frame_c1(29, 17, ssaa=1, fxaa=0)]
    ctx = gpu_context(A.PRR_KERR_BL, CF.tracer_kerr(precision=A.PRR_FP64))
    ctx.set_launch(block=128, blocks_per_cu=1, refill=refill)
    batch = ctx.render(frames, rays=True)
    assert ctx.stats()["rays"] == sum(f.nrays for f in frames)
    ctx.set_launch()
    for f, b in zip(frames, batch):
        single = ctx.render(f, rays=True)[0]
        for key in ("cls", "steps", "attempts", "rgb8", "rgbf"):
            assert np.array_equal(single[key], b[key]), (refill, key)


@pytest.mark.parametrize("metric", [A.PRR_SCHW_KS_CART, A.PRR_SCHW_SPH, A.PRR_ELLIS])
def test_fp32_bulk_second_order_metrics_track_fp64(metric):
    """The 8-component metrics in FP32-bulk: FP32 stages of the full
    right-hand side (no difference form), FP64 state, error norm and
    controller.  At RKF45 tol 1e-8 (FP32 rounding of the stages, ~6e-8 of
    h f, stays below the tolerance for steps up to ~0.1 of the state's
    scale), classes agree with the FP64 trace and escape end states agree to
    the tolerance's scale."""
    T64 = trace_tracer(A.PRR_RKF45)
    T64.atol = T64.rtol = 1e-8
    T32 = trace_tracer(A.PRR_RKF45)
    T32.atol = T32.rtol = 1e-8
    T32.precision = A.PRR_FP32_BULK
    T32.r_fp64 = 0.0  # FP32 stages everywhere
    if metric == A.PRR_ELLIS:
        x0 = wormhole_states(256)
        g64, g32 = (wormhole_context(T).trace(x0) for T in (T64, T32))
    else:
        x0 = random_states(metric, 256)
        out = []
        for T in (T64, T32):
            ctx = api.Context(0)
            ctx.set_spacetime(metric, 1.0, 0.0)
            ctx.set_tracer(T)
            out.append(ctx.trace(x0))
        g64, g32 = out
    agree = g64["cls"] == g32["cls"]
    assert agree.mean() >= 0.98
    esc = agree & (g64["cls"] == A.PRR_ESCAPE)
    assert esc.sum() >= 50
    err = np.abs(g32["xend"][esc] - g64["xend"][esc]).max(axis=1) / (np.abs(g64["xend"][esc]).max(axis=1) + 1)
    print("fp32-bulk vs fp64 escape states: median %.2e p99 %.2e" % (np.median(err), np.quantile(err, 0.99)))
    assert np.median(err) < 1e-6 and np.quantile(err, 0.9) < 1e-5
    # FP32 stage rounding costs rejected attempts, not a collapse of the step size
    assert g32["attempts"][esc].sum() < 3 * g64["attempts"][esc].sum()


def test_fp32_bulk_close_to_fp64():
    f = CF.frame_c1(192, 108, ssaa=1, fxaa=0)
    c64 = gpu_context(A.PRR_KERR_BL, CF.tracer_kerr(precision=A.PRR_FP64)).render(f, rays=True)[0]
    c32 = gpu_context(A.PRR_KERR_BL, CF.tracer_kerr(precision=A.PRR_FP32_BULK)).render(f, rays=True)[0]
    same = c64["cls"] == c32["cls"]
    assert same.mean() > 0.995
    m = same & ((c64["cls"] == A.PRR_ESCAPE) | (c64["cls"] == A.PRR_SPHERE))
    ang = angle(c64["endpoint"][m], c32["endpoint"][m])
    assert np.median(ang) < 1e-6


@pytest.mark.parametrize("metric", [A.PRR_SCHW_KS_CART, A.PRR_KERR_BL])
def test_fp32_bulk_euler_tracks_fp64_and_does_not_depend_on_launch_shape(metric):
    """Euler steps in FP32-bulk: FP32 right-hand side outside r_fp64 (the
    whole wave in FP64 when any of its lanes is inside), FP64 state.  Classes
    agree with the FP64 Euler render; images and step counts do not depend on
    the launch shape, and a frame renders the same alone or after a frame
    whose ray count is not a multiple of 64."""
    def ctx_for(precision):
        T = A.make_tracer(solver=A.PRR_EULER, h=0.02, r_escape=40.0, horizon_eps_rel=1e-3, horizon_eps_abs=0.0,
                          lambda_max=1e9, max_steps=6000, max_attempts=6000)
        T.precision = precision
        ctx = api.Context(0)
        ctx.set_spacetime(metric, 1.0, 0.95 if metric == A.PRR_KERR_BL else 0.0)
        ctx.set_tracer(T)
        ctx.set_texture(0, CF.sky_texture())
        ctx.set_texture(1, CF.sphere_texture(), A.PRR_FILTER_BILINEAR)
        return ctx
    f = CF.frame_c1(96, 54)
    odd = CF.frame_c1(37, 23)  # 851 pixels x 4 samples: 3404 rays, not a multiple of 64
    c64 = ctx_for(A.PRR_FP64).render(f, rays=True)[0]
    ctx = ctx_for(A.PRR_FP32_BULK)
    ref = ctx.render(f, rays=True)[0]
    assert (ref["cls"] == c64["cls"]).mean() >= 0.99
    assert (ref["cls"] == A.PRR_HORIZON).sum() >= 100 and (ref["cls"] == A.PRR_ESCAPE).sum() >= 100
    st = ctx.stats()
    assert 0 < st["attempts_fp64"] < st["attempts"]
    ctx.set_launch(block=64, blocks_per_cu=1, refill=1)
    out = ctx.render(f, rays=True)[0]
    assert np.array_equal(ref["rgbf"], out["rgbf"]) and np.array_equal(ref["steps"], out["steps"])
    ctx.set_launch()
    both = ctx.render([odd, f], rays=True)
    assert np.array_equal(both[1]["rgbf"], ref["rgbf"]) and np.array_equal(both[1]["steps"], ref["steps"])
