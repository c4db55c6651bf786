"""Pennes bioheat + CEM43 dose restated on the CPU (test oracle; see oracle/__init__.py).

Restates upstream ``src/caustica/thermal/pennes.py`` (stable_dt 376-402,
coefficient assembly 404-447, solve 451-689, hot loop 608-654) and
``thermal/dose.py:74-89`` (cem43_rate).

Pinning (tests/test_oracle_golden.py): fp32 mode against the upstream fixtures
pennes_*.npz (insulated and Dirichlet walls, sub-stepping, perfusion, heat
source, chained cooling phase with dose0); fp64 mode against the same fixtures
at float32 round-off and against closed forms (Gaussian spreading, perfusion
relaxation, CEM43 unit definitions) in tests/test_oracle_physics.py.
"""

from __future__ import annotations

from dataclasses import dataclass
from math import ceil

import numpy as np

BLOOD_DENSITY = 1050.0
BLOOD_SPECIFIC_HEAT = 3617.0
ARTERIAL_TEMPERATURE_C = 37.0

_ALL = slice(None)
_LOWER = slice(None, -1)
_UPPER = slice(1, None)


def _ax(arr, axis, sl):
    idx = [_ALL] * arr.ndim
    idx[axis] = sl
    return arr[tuple(idx)]


def cem43_rate(t):
    """R^(43 - T), R = 0.5 above 43 C and 0.25 at or below (dose.py:74-89)."""
    t = np.asarray(t)
    s = t.dtype.type
    r = np.where(t > 43.0, s(0.5), s(0.25))
    return r ** (s(43.0) - t)


@dataclass
class ThermalOut:
    temperature: np.ndarray
    temperature_max: np.ndarray
    dose: np.ndarray | None
    substeps: int
    dt_stable: float


class PennesOracle:
    """Explicit conservative FD Pennes solve in "fp32" (upstream order) or "fp64"."""

    def __init__(
        self,
        k,
        rho,
        specific_heat,
        perfusion,
        dx,
        *,
        boundary="insulated",
        on_unstable="refuse",
        precision="fp64",
        blood_density=BLOOD_DENSITY,
        blood_specific_heat=BLOOD_SPECIFIC_HEAT,
        arterial_temperature_c=ARTERIAL_TEMPERATURE_C,
    ):
        self.fp32 = precision == "fp32"
        self.rdt = np.float32 if self.fp32 else np.float64
        self.k = np.ascontiguousarray(k, np.float32)
        self.rho = np.ascontiguousarray(rho, np.float32)
        self.cp = np.ascontiguousarray(specific_heat, np.float32)
        self.w = np.ascontiguousarray(perfusion, np.float32)
        self.dx = float(dx)
        self.boundary = boundary
        self.on_unstable = on_unstable
        self.rho_b = float(blood_density)
        self.c_b = float(blood_specific_heat)
        self.t_a = float(arterial_temperature_c)
        self.t_ref = self.t_a
        self.t_wall = self.t_a

    # --------------------------------------------------------- coefficients
    def faces(self, k):
        inv_dx2 = 1.0 / (self.dx * self.dx)
        out = []
        for ax in range(k.ndim):
            lo = _ax(k, ax, _LOWER)
            hi = _ax(k, ax, _UPPER)
            out.append(((2.0 * lo * hi / (lo + hi)) * inv_dx2).astype(self.rdt))
        return out

    def walls(self, k):
        inv_dx2 = 1.0 / (self.dx * self.dx)
        out = []
        for ax in range(k.ndim):
            lo = (2.0 * _ax(k, ax, slice(0, 1)) * inv_dx2).astype(self.rdt)
            hi = (2.0 * _ax(k, ax, slice(-1, None)) * inv_dx2).astype(self.rdt)
            out.append((lo, hi))
        return out

    def perf_coeff(self):
        if self.fp32:
            return (self.w * np.float32(self.rho_b * self.c_b)).astype(np.float32)
        return self.w.astype(np.float64) * (self.rho_b * self.c_b)

    def inv_rho_c(self):
        rc = (self.rho.astype(np.float64) * self.cp).astype(np.float32)
        if self.fp32:
            return (1.0 / rc.astype(np.float64)).astype(np.float32)
        return 1.0 / (self.rho.astype(np.float64) * self.cp.astype(np.float64))

    def stable_dt(self):
        k = self.k.astype(self.rdt)
        fc = self.faces(k)
        diag = np.zeros(k.shape, self.rdt)
        for ax, kf in enumerate(fc):
            _ax(diag, ax, _LOWER)[...] += kf
            _ax(diag, ax, _UPPER)[...] += kf
        if self.boundary == "dirichlet":
            for ax, (lo, hi) in enumerate(self.walls(k)):
                _ax(diag, ax, slice(0, 1))[...] += lo
                _ax(diag, ax, slice(-1, None))[...] += hi
        diag += self.perf_coeff()
        diag *= self.inv_rho_c()
        peak = float(diag.max())
        return float("inf") if peak <= 0.0 else 1.0 / peak

    # ---------------------------------------------------------------- solve
    def solve(self, temperature0, q, dt, n_steps, *, dose=False, dose0=None) -> ThermalOut:
        rdt = self.rdt
        t_ref = rdt(self.t_ref)
        theta = np.asarray(temperature0, np.float32).astype(rdt) - t_ref
        q_dev = None if q is None else np.asarray(q, np.float32).astype(rdt)
        k = self.k.astype(rdt)
        fc = self.faces(k)
        wl = self.walls(k) if self.boundary == "dirichlet" else None
        perf = self.perf_coeff()
        perfused = bool((self.w > 0).any())
        irc = self.inv_rho_c()
        dt_stable = self.stable_dt()
        if dt <= dt_stable:
            n_sub = 1
        elif self.on_unstable == "substep":
            n_sub = int(ceil(dt / dt_stable))
        else:
            raise ValueError(f"dt {dt} above stability bound {dt_stable}")
        dt_sub = dt / n_sub
        coeff = (irc * rdt(dt_sub)).astype(rdt)
        theta_wall = rdt(self.t_wall) - t_ref
        shape = theta.shape
        work = np.zeros(shape, rdt)
        theta_max = theta.copy()
        dose_map = None
        if dose:
            dose_map = (
                np.zeros(shape, rdt) if dose0 is None else np.asarray(dose0, np.float32).astype(rdt)
            )
            rate = cem43_rate(theta + t_ref)
        half_dt_min = rdt(0.5 * dt_sub / 60.0)
        for _step in range(n_steps):
            for _ in range(n_sub):
                work.fill(0)
                for ax, kf in enumerate(fc):
                    buf = (_ax(theta, ax, _UPPER) - _ax(theta, ax, _LOWER)).astype(rdt)
                    buf *= kf
                    _ax(work, ax, _LOWER)[...] += buf
                    _ax(work, ax, _UPPER)[...] -= buf
                if wl is not None:
                    for ax, (lo, hi) in enumerate(wl):
                        _ax(work, ax, slice(0, 1))[...] += lo * (
                            theta_wall - _ax(theta, ax, slice(0, 1))
                        )
                        _ax(work, ax, slice(-1, None))[...] += hi * (
                            theta_wall - _ax(theta, ax, slice(-1, None))
                        )
                if q_dev is not None:
                    work += q_dev
                if perfused:
                    work -= perf * theta
                work *= coeff
                theta += work
                np.maximum(theta_max, theta, out=theta_max)
                if dose_map is not None:
                    new_rate = cem43_rate(theta + t_ref)
                    inc = new_rate + rate
                    inc *= half_dt_min
                    dose_map += inc
                    rate = new_rate
            if not np.isfinite(theta).all():
                raise FloatingPointError("thermal state diverged")
        return ThermalOut(
            temperature=theta + t_ref,
            temperature_max=theta_max + t_ref,
            dose=dose_map,
            substeps=n_sub,
            dt_stable=dt_stable,
        )
