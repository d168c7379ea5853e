"""Generate darwin_deflection.json: exact light deflection by a Schwarzschild mass.

Darwin (1959), closed form with incomplete elliptic integrals of the first kind:
for closest approach r0 (> 3M), impact parameter b^2 = r0^3 / (r0 - 2M),
    Q^2 = (r0 - 2M)(r0 + 6M),  k^2 = (Q - r0 + 6M) / (2Q),
    sin^2(psi) = (Q - r0 + 2M) / (Q - r0 + 6M),
    alpha = 4 sqrt(r0/Q) [K(k) - F(psi, k)] - pi.
Each value is cross-checked against direct quadrature of the orbit integral
    alpha = 2 int_0^{1/r0} b dw / sqrt(1 - b^2 w^2 (1 - 2Mw)) - pi
(written as an integral over w = 1/r with the endpoint singularity removed by
w = (1 - t^2)/r0).  scipy's ellipkinc takes the parameter m = k^2.
Run from the repo root: python tests/golden/make_darwin.py
"""
import json
import os

import numpy as np
from scipy import integrate, special

M = 1.0


def darwin(r0):
    Q = np.sqrt((r0 - 2 * M) * (r0 + 6 * M))
    m = (Q - r0 + 6 * M) / (2 * Q)
    psi = np.arcsin(np.sqrt((Q - r0 + 2 * M) / (Q - r0 + 6 * M)))
    return 4 * np.sqrt(r0 / Q) * (special.ellipk(m) - special.ellipkinc(psi, m)) - np.pi


def quadrature(r0):
    b = np.sqrt(r0 ** 3 / (r0 - 2 * M))
    w0 = 1.0 / r0

    def f(t):  # w = w0 (1 - t^2), dw = -2 w0 t dt
        w = w0 * (1 - t * t)
        g = 1 - b * b * w * w * (1 - 2 * M * w)
        return 2 * w0 * t * b / np.sqrt(max(g, 1e-300))

    val, _ = integrate.quad(f, 0.0, 1.0, epsabs=1e-14, epsrel=1e-13, limit=500)
    return 2 * val - np.pi


rows = []
for r0 in [3.2, 3.5, 4.0, 5.0, 6.0, 8.0, 10.0, 15.0, 20.0, 30.0, 50.0, 100.0]:
    b = float(np.sqrt(r0 ** 3 / (r0 - 2 * M)))
    a_d, a_q = float(darwin(r0)), float(quadrature(r0))
    assert abs(a_d - a_q) < 1e-9 * max(1.0, abs(a_d)), (r0, a_d, a_q)
    rows.append({"r0": r0, "b": b, "alpha": a_d, "alpha_quadrature": a_q})

out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "darwin_deflection.json")
with open(out, "w") as fh:
    json.dump({"M": M, "source": "Darwin 1959 elliptic-integral form, cross-checked by quadrature",
               "rows": rows}, fh, indent=1)
print(out, len(rows))
