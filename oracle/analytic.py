"""Closed-form references used by the tests (test oracle; see oracle/__init__.py).

Restates upstream ``src/caustica/analytic/oneill.py:33-80`` (O'Neil on-axis pressure of
a focused bowl) and ``analytic/planewave.py:18-57`` (exponential attenuation, shock
distance, Fubini harmonic amplitudes). These are textbook formulas; the tests pin them
by their limits (focal gain k h, B1 -> 1 and B2 -> sigma/2 as sigma -> 0, Bessel series
sum) before using them to pin the solvers.
"""

from __future__ import annotations

import numpy as np
from scipy.special import jv


def oneill_axial(z, aperture_radius, roc, f0, c0=1500.0, rho0=1000.0, u0=1.0):
    """p(z) = rho c u0 / (1 - z/roc) (exp(ikz) - exp(ik d_rim)), focus value -i k h rho c u0 e^{ik roc}."""
    z = np.atleast_1d(np.asarray(z, np.float64))
    k = 2.0 * np.pi * f0 / c0
    a, f = float(aperture_radius), float(roc)
    h = f - np.sqrt(f * f - a * a)
    d_rim = np.sqrt(a * a + (z - h) ** 2)
    den = 1.0 - z / f
    near = np.abs(den) < 1e-9
    p = rho0 * c0 * u0 / np.where(near, 1.0, den) * (np.exp(1j * k * z) - np.exp(1j * k * d_rim))
    return np.where(near, -1j * k * h * rho0 * c0 * u0 * np.exp(1j * k * f), p)


def shock_distance(p0, f0, c0=1500.0, rho0=1000.0, beta=3.5):
    return rho0 * c0**3 / (beta * 2.0 * np.pi * f0 * p0)


def fubini(n: int, sigma):
    """B_n(sigma) = 2 J_n(n sigma) / (n sigma), sigma in [0, 1]."""
    s = np.asarray(sigma, np.float64)
    with np.errstate(invalid="ignore", divide="ignore"):
        b = 2.0 * jv(n, n * s) / (n * s)
    return np.where(s == 0, 1.0 if n == 1 else 0.0, b)
