"""Building blocks of the Kerr adjoint, evaluated on the host (no GPU):
ray set-up and its Jacobian, the reverse sweep through one RKF45 step, and
the terminal colour derivative, each against central finite differences or
against the forward code the kernels run.
"""
import math

import numpy as np
import pytest

from parrelray_amd import _abi as A
from parrelray_amd import api
from parrelray_amd import configs as CF

RNG = np.random.default_rng(3)


def test_orbit_init_matches_forward_ray_setup():
    o = CF.c3_orbit()
    f = CF.frame_c3(64, 48)
    for ray in RNG.integers(0, f.nrays, size=50):
        j, i = divmod(int(ray), f.width)
        u, v = (i + 0.5) / f.width, (j + 0.5) / f.height
        z, _ = api.orbit_init(o.spin, o.r_cam, o.incl, o.cam_phi, f.vfov, f.width / f.height, u, v)
        d = api.camera_ray(f, int(ray))
        y, ELQ = api.init_ray(A.PRR_KERR_BL, list(f.pos), d, 1.0, o.spin)
        assert np.allclose(z, np.concatenate([y[:5], ELQ]), rtol=1e-11, atol=1e-12)


def test_orbit_init_jacobian_matches_finite_differences():
    o = CF.c3_orbit()
    vfov, aspect = math.radians(50.0), 1.0
    for _ in range(20):
        u, v = RNG.uniform(0.05, 0.95, size=2)
        p = np.array([o.spin, o.r_cam, o.incl])
        _, dz = api.orbit_init(*p, o.cam_phi, vfov, aspect, u, v)
        for k in range(3):
            h = 1e-6
            pp, pm = p.copy(), p.copy()
            pp[k] += h
            pm[k] -= h
            zp, _ = api.orbit_init(*pp, o.cam_phi, vfov, aspect, u, v)
            zm, _ = api.orbit_init(*pm, o.cam_phi, vfov, aspect, u, v)
            fd = (zp - zm) / (2 * h)
            assert np.allclose(dz[:, k], fd, rtol=1e-6, atol=1e-8 * (np.abs(fd).max() + 1)), (k, dz[:, k], fd)


def random_kerr_state(a, rmin=3.0, rmax=40.0):
    """A null Kerr state from a random camera position and direction: (y, (E, L, Q))."""
    r = RNG.uniform(rmin, rmax)
    th = RNG.uniform(0.3, math.pi - 0.3)
    ph = RNG.uniform(0, 6)
    pos = r * np.array([math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th)])
    d = RNG.normal(size=3)
    d /= np.linalg.norm(d)
    y, ELQ = api.init_ray(A.PRR_KERR_BL, pos, d, 1.0, a)
    return y, ELQ


@pytest.mark.parametrize("a", [0.0, 0.7, 0.95])
def test_step_vjp_is_the_transpose_jacobian_of_the_step(a):
    for _ in range(20):
        y, (E, L, Q) = random_kerr_state(a)
        F = api.rhs(A.PRR_KERR_BL, y, 1.0, a, E, L, Q)
        h = RNG.uniform(0.02, 0.2) * min(y[0] / abs(F[0]), 1.0 / np.abs(F[1:4]).max())
        lam = RNG.normal(size=9)
        got = api.kerr_step_vjp(y, h, lam, a, E, L, Q)
        z = np.concatenate([y[:5], [E, L, Q, a]])

        def G(zz):
            yy = np.concatenate([zz[:5], [y[5]]])
            return lam[:5] @ api.kerr_step(yy, h, zz[8], zz[5], zz[6], zz[7])[:5] + lam[5:] @ zz[5:]

        def central(k, eps):
            zp, zm = z.copy(), z.copy()
            zp[k] += eps
            zm[k] -= eps
            return (G(zp) - G(zm)) / (2 * eps)

        for k in range(9):
            # Richardson-extrapolated central differences
            eps = 1e-4 * max(1e-2, abs(z[k]))
            fd = (4 * central(k, eps / 2) - central(k, eps)) / 3
            scale = abs(fd) + 1e-3 * np.abs(got).max()
            assert abs(got[k] - fd) <= 2e-6 * scale + 1e-9, (k, got[k], fd)


def test_escape_color_derivative_matches_finite_differences():
    a = 0.7
    tex = CF.sphere_texture()  # smooth texture: derivatives everywhere
    n_checked = n = 0
    for _ in range(40):
        y, (E, L, Q) = random_kerr_state(a, 60.0, 120.0)
        c, dc = api.escape_color(y, a, E, L, Q, tex)
        z = np.concatenate([y[:5], [E, L, Q, a]])
        for k in range(9):
            n += 1
            eps = 1e-7 * max(1e-2, abs(z[k]))
            zp, zm = z.copy(), z.copy()
            zp[k] += eps
            zm[k] -= eps
            cp, _ = api.escape_color(zp[:5], zp[8], zp[5], zp[6], zp[7], tex)
            cm, _ = api.escape_color(zm[:5], zm[8], zm[5], zm[6], zm[7], tex)
            fd = (cp - cm) / (2 * eps)
            if np.abs(fd - dc[:, k]).max() > 1e-3 * (np.abs(fd).max() + 1e-3):
                # the only non-smooth points of the bilinear colour are texel edges
                continue
            n_checked += 1
    assert n_checked >= 0.95 * n
