"""Closed-form references used by the tests (test oracle; see oracle/__init__.py).

Restates upstream ``src/caustica/analytic/oneill.py:33-80`` (O'Neil on-axis pressure of
a focused bowl), ``analytic/planewave.py:18-57`` (exponential attenuation, shock
distance, Fubini harmonic amplitudes) and ``analytic/rayleigh.py:30-84`` (Rayleigh-
Sommerfeld integral of a sampled surface). These are textbook formulas; the tests pin
them by their limits (focal gain k h, B1 -> 1 and B2 -> sigma/2 as sigma -> 0, Bessel
series sum; the Rayleigh sum of a finely sampled cap reproduces O'Neil on the axis, and
one point source gives rho c k v dS / (2 pi R)) before using them to pin the solvers.
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


def rayleigh_pressure(points, areas, v_n, targets, k: float, rho: float = 1000.0, c: float = 1500.0,
                      chunk: int = 4096):
    """Rayleigh-Sommerfeld CW pressure of a sampled source surface (e^{-i omega t} convention):
    p(x) = -i rho c k / (2 pi) sum_j v_j dS_j exp(i k R_j) / R_j, R_j = |x - x_j|
    (upstream analytic/rayleigh.py:30-84)."""
    src = np.asarray(points, np.float64)
    fld = np.atleast_2d(np.asarray(targets, np.float64))
    strength = np.broadcast_to(np.asarray(v_n, np.complex128), (src.shape[0],)) * np.asarray(areas, np.float64)
    out = np.empty(fld.shape[0], np.complex128)
    for a in range(0, fld.shape[0], chunk):
        r = np.sqrt(((fld[a:a + chunk, None, :] - src[None, :, :]) ** 2).sum(axis=-1))
        out[a:a + chunk] = (np.exp(1j * k * r) / r) @ strength
    return (-1j * rho * c * k / (2.0 * np.pi)) * out


def powerlaw_wavenumber(f, alpha_f0, f0, y, c0=1500.0, iters=60):
    """Complex wavenumber k = k_r + i alpha of a CW plane wave in the fractional-Laplacian
    power-law medium (Treeby and Cox, JASA 127(5) 2010): omega^2 = c0^2 k^2 (1 + A),
    A = tau i omega |k|^(y-2) - eta |k|^(y-1), tau = -2 a0 c0^(y-1),
    eta = 2 a0 c0^y tan(pi y / 2), a0 = alpha_f0 / omega0^y; |k| taken as the real part,
    solved by fixed-point iteration. To first order in a0: alpha = a0 omega^y and
    1 / c_phase = 1 / c0 + a0 tan(pi y / 2) omega^(y - 1)."""
    w = 2.0 * np.pi * float(f)
    a0 = float(alpha_f0) / (2.0 * np.pi * float(f0)) ** y
    tau = -2.0 * a0 * c0 ** (y - 1.0)
    eta = 2.0 * a0 * c0**y * np.tan(np.pi * y / 2.0)
    k = complex(w / c0)
    for _ in range(iters):
        kr = k.real
        a = tau * 1j * w * kr ** (y - 2.0) - eta * kr ** (y - 1.0)
        k = (w / c0) / np.sqrt(1.0 + a)
    return k
