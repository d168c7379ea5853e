"""The Kerr sphere screen's safe-distance budget (rkf_events / sphere_gap in
parrelray_amd/csrc/lane.hpp): the facts it rests on, checked in FP64 against
exact Cartesian geometry.

1. The chord bound: for points given by cylindrical (rho, z) and phi,
   |P1 - P0|^2 <= dz^2 + drho^2 + rho0 rho1 dphi^2 for any dphi (windings
   included), since 1 - cos x <= x^2 / 2.
2. The budget: after a screen at P_m, a later chord n is skipped while
   g - sum(D_i, i = m+1..n) > D_n / 2 with g = |P_m - c| - R; then every
   point of chord n lies farther than R + |d_n| / 2 from the centre, which is
   what the screen's pass condition needs to fail (the screen passes within
   R + |d_n| / 2 plus its margin; the margin is checked on the GPU by the
   parity tests, whose results are the same with and without the budget).

A bound that dropped the rho0 rho1 factor, used rho0 + rho1, or spent only
half of each step would fail these checks.
"""
import numpy as np


def _cart(rho, z, phi):
    return np.stack([rho * np.cos(phi), rho * np.sin(phi), z], axis=-1)


def _chord_bound(rho0, z0, phi0, rho1, z1, phi1):
    dphi = phi1 - phi0
    return np.sqrt((rho1 - rho0) ** 2 + (z1 - z0) ** 2 + rho0 * rho1 * dphi * dphi)


def _seg_dist(p0, p1, c):
    d = p1 - p0
    a = float(d @ d)
    t = 0.0 if a == 0.0 else min(max(-float((p0 - c) @ d) / a, 0.0), 1.0)
    return float(np.linalg.norm(p0 + t * d - c))


def test_chord_length_bound_holds_for_any_phi_change():
    rng = np.random.default_rng(7)
    n = 200_000
    rho0, rho1 = rng.uniform(0, 50, n), rng.uniform(0, 50, n)
    z0, z1 = rng.uniform(-50, 50, n), rng.uniform(-50, 50, n)
    phi0 = rng.uniform(-20, 20, n)
    # small and large changes of phi, several windings included
    dphi = np.where(rng.random(n) < 0.5, rng.normal(0, 0.05, n), rng.uniform(-30, 30, n))
    phi1 = phi0 + dphi
    exact = np.linalg.norm(_cart(rho1, z1, phi1) - _cart(rho0, z0, phi0), axis=1)
    bound = _chord_bound(rho0, z0, phi0, rho1, z1, phi1)
    assert np.all(exact <= bound * (1 + 1e-12) + 1e-12)
    # tight for small steps: the bound is not a loose over-estimate
    small = np.abs(dphi) < 1e-3
    assert np.median((bound[small] - exact[small]) / np.maximum(exact[small], 1e-9)) < 1e-6


def test_skipped_chords_stay_farther_than_the_screen_reaches():
    rng = np.random.default_rng(11)
    R = 1.0
    skipped = 0
    for _ in range(400):
        c = _cart(rng.uniform(4, 12), rng.uniform(-2, 2), rng.uniform(0, 2 * np.pi))
        # a random walk in (rho, z, phi) with steps of varying size
        m = 60
        rho = np.abs(8 + np.cumsum(rng.normal(0, 0.6, m)))
        z = np.cumsum(rng.normal(0, 0.6, m))
        phi = np.cumsum(rng.normal(0, 0.15, m))
        P = _cart(rho, z, phi)
        gap = 0.0
        for k in range(1, m):
            D = float(_chord_bound(rho[k - 1], z[k - 1], phi[k - 1], rho[k], z[k], phi[k]))
            gap -= D
            if gap > 0.5 * D:
                skipped += 1
                d = float(np.linalg.norm(P[k] - P[k - 1]))
                assert _seg_dist(P[k - 1], P[k], c) > R + 0.5 * d
            else:  # screened here: the budget restarts from the chord's end
                gap = float(np.linalg.norm(P[k] - c)) - R
    assert skipped > 1000  # the walk exercises the skip path
