"""The benchmark workloads, as frames and tracer settings.

C0  Schwarzschild (Schwarzschild coordinates), Euler h = 0.01, FP64, 256x256,
    60 deg FOV, camera r = 30, inclination 85 deg, checker sky, escape r > 100,
    horizon r < 2(1 + 1e-3), 10,000-step cap, no SSAA / FXAA.
C1  Kerr a = 0.95 (Boyer-Lindquist), RKF45 rtol = atol = 1e-8, 1920x1080,
    50 deg FOV, camera r = 20, inclination 80 deg, textured sphere R = 1 on the
    equator at r = 8, 2x2 SSAA + FXAA, 50,000-step cap.
C2  64 frames along a camera trajectory (r 40 -> 12, inclination 60 -> 89 deg),
    spin cycling {0, .3, .6, .9, .998}, resolution cycling {512^2, 1024x768,
    1920x1080, 2048x1024}, SSAA cycling {1, 2x2, 3x3}; one persistent launch.
C3  512x512 Kerr a = 0.7 render for the adjoint (camera r = 15, i = 75 deg,
    sphere phi = 0.3) and its target (a = 0.75, r = 15.5, i = 77 deg, phi = 0.35).
C4  Kerr a = 0.998, camera r = 6 edge-on (i = 89.9 deg), 2048x2048, 4x4 SSAA
    + FXAA, tol 1e-10.

Camera convention (main/Camera.cpp:19-28): a camera at inclination i on the
x-z plane looking at the hole has pos = r (sin i, 0, cos i), alpha = 180 deg,
beta = 90 deg - i.
"""
import math

import numpy as np

from . import _abi as A

SKY_SEED = 0x5EED0001
SPHERE_SEED = 0x5EED0002
C2_SEED = 0x5EED0C02


def orbit_pose(r, incl_deg, cam_phi=0.0):
    i = math.radians(incl_deg)
    pos = [r * math.sin(i) * math.cos(cam_phi), r * math.sin(i) * math.sin(cam_phi), r * math.cos(i)]
    return pos, cam_phi + math.pi, math.pi / 2 - i


def equatorial_sphere(r, phi, radius=1.0, tex=1):
    return ([r * math.cos(phi), r * math.sin(phi), 0.0], radius, tex)


def sky_texture():
    from .api import checker_texture
    return checker_texture(512, 256, 16, 8, SKY_SEED)


def sphere_texture():
    from .api import noise_texture
    return noise_texture(1024, 512, 4, SPHERE_SEED)


def tracer_c0():
    return A.make_tracer(solver=A.PRR_EULER, precision=A.PRR_FP64, h=0.01, hmin=0.01, hmax=0.01, atol=1.0,
                         rtol=0.0, r_escape=100.0, horizon_eps_rel=1e-3, horizon_eps_abs=0.0, lambda_max=5000.0,
                         max_steps=10000, max_attempts=10000, r_fp64=3.0)


def frame_c0(width=256, height=256):
    pos, al, be = orbit_pose(30.0, 85.0)
    return A.make_frame(pos, al, be, math.radians(60.0), width, height, mass=1.0, spin=0.0,
                        projection=A.PRR_PROJ_PINHOLE)


def tracer_kerr(precision=A.PRR_FP64, tol=1e-8, max_steps=50000):
    return A.make_tracer(solver=A.PRR_RKF45, precision=precision, h=0.1, hmin=1e-8, hmax=10.0, atol=tol, rtol=tol,
                         r_escape=100.0, horizon_eps_rel=1e-3, horizon_eps_abs=0.0, lambda_max=1000.0,
                         max_steps=max_steps, max_attempts=10 * max_steps, r_fp64=3.0)


def frame_c1(width=1920, height=1080, ssaa=2, fxaa=1):
    pos, al, be = orbit_pose(20.0, 80.0)
    return A.make_frame(pos, al, be, math.radians(50.0), width, height,
                        ssaa=A.PRR_SSAA_REGULAR if ssaa > 1 else A.PRR_SSAA_NONE, samples=ssaa * ssaa,
                        fxaa=fxaa, mass=1.0, spin=0.95, spheres=[equatorial_sphere(8.0, 0.3)])


C2_SPINS = [0.0, 0.3, 0.6, 0.9, 0.998]
C2_RES = [(512, 512), (1024, 768), (1920, 1080), (2048, 1024)]
C2_SSAA = [1, 2, 3]


def frames_c2(nframes=64, scale=1.0):
    """The 64-frame trajectory; `scale` shrinks every resolution (tests)."""
    rng = np.random.default_rng(C2_SEED)
    jitter = rng.uniform(-0.5, 0.5, size=nframes)
    frames = []
    for k in range(nframes):
        r = 40.0 - 28.0 * k / 63.0
        inc = 60.0 + 29.0 * k / 63.0 + jitter[k]
        pos, al, be = orbit_pose(r, inc)
        w, h = C2_RES[k % 4]
        w, h = max(4, int(w * scale)), max(4, int(h * scale))
        n = C2_SSAA[k % 3]
        frames.append(A.make_frame(pos, al, be, math.radians(50.0), w, h,
                                   ssaa=A.PRR_SSAA_REGULAR if n > 1 else A.PRR_SSAA_NONE, samples=n * n, fxaa=1,
                                   mass=1.0, spin=C2_SPINS[k % 5],
                                   spheres=[equatorial_sphere(8.0, 0.3 + 0.1 * k)]))
    return frames


def c3_orbit(target=False):
    o = A.Orbit()
    o.spin, o.r_cam, o.incl, o.sphere_phi = (0.75, 15.5, math.radians(77.0), 0.35) if target else \
        (0.7, 15.0, math.radians(75.0), 0.3)
    o.cam_phi, o.sphere_r, o.sphere_radius, o.sphere_tex, o.has_sphere = 0.0, 8.0, 1.0, 1, 1
    return o


def frame_c3(width=512, height=512, target=False):
    o = c3_orbit(target)
    pos, al, be = orbit_pose(o.r_cam, math.degrees(o.incl))
    return A.make_frame(pos, al, be, math.radians(50.0), width, height, mass=1.0, spin=o.spin,
                        spheres=[equatorial_sphere(o.sphere_r, o.sphere_phi)])


def frame_c4(width=2048, height=2048, ssaa=4):
    pos, al, be = orbit_pose(6.0, 89.9)
    return A.make_frame(pos, al, be, math.radians(50.0), width, height,
                        ssaa=A.PRR_SSAA_REGULAR if ssaa > 1 else A.PRR_SSAA_NONE, samples=ssaa * ssaa, fxaa=1,
                        mass=1.0, spin=0.998)


def tracer_c4(precision=A.PRR_FP32_BULK):
    return tracer_kerr(precision=precision, tol=1e-10, max_steps=200000)
