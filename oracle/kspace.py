"""CW k-space PSTD engine restated on the CPU (test oracle; see oracle/__init__.py).

Restates upstream ``src/caustica/solvers/kspace/engine.py`` (run loop 167-586,
step 458-482), ``solvers/kspace/operators.py`` (23-157), ``core/pml.py:53-73``
and ``sources.py:98-102``.

Pinning (tests/test_oracle_golden.py):
  * fp32 mode against the upstream fixtures cw_*.npz (phasors, p_max, spp,
    step counts, tof, convergence history) and ops_*.npz (kappa, derivative
    factors, sponge) -- bit-exact or float32 round-off;
  * fp64 mode against the same fixtures at the float32 noise floor, and against
    closed forms (plane-wave phase speed, attenuation slope, beta=0 identity,
    linearity in the drive amplitude) in tests/test_oracle_physics.py.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from itertools import product
from math import ceil, floor, isfinite

import numpy as np
import scipy.fft as spfft


# ----------------------------------------------------------------- operators
def optimal_fft_size(n: int) -> int:
    """Smallest 2/3/5-smooth length >= n (operators.py:23-34)."""
    while True:
        m = n
        for p in (2, 3, 5):
            while m % p == 0:
                m //= p
        if m == 1:
            return n
        n += 1


def pad_shape(shape):
    return tuple(optimal_fft_size(int(n)) for n in shape)


def pad_volume(vol: np.ndarray, padded) -> np.ndarray:
    """High-end edge replication (operators.py:42-57)."""
    pad = [(0, w - h) for h, w in zip(vol.shape, padded)]
    if all(p == (0, 0) for p in pad):
        return vol
    return np.pad(vol, pad, mode="edge")


def sponge_profile_1d(n: int, width: int, edge: float, dtype) -> np.ndarray:
    """Gaussian sponge ramp mirrored at both ends (core/pml.py:53-73)."""
    s = np.ones(n, dtype=dtype)
    if width > 0:
        r = np.arange(width, dtype=dtype)
        ramp = np.exp(-edge * ((width - r) / width) ** 2)
        s[:width] = ramp
        s[-width:] = ramp[::-1]
    return s


def k_vectors(shape, dx: float, dtype):
    """Angular k per axis, rfft layout on the last axis (operators.py:60-73)."""
    ks = []
    nd = len(shape)
    for ax, n in enumerate(shape):
        f = np.fft.rfftfreq(n, d=dx) if ax == nd - 1 else np.fft.fftfreq(n, d=dx)
        k = (2.0 * np.pi * f).astype(dtype)
        sh = [1] * nd
        sh[ax] = k.size
        ks.append(k.reshape(sh))
    return ks


def kappa_sinc(ks, c_ref: float, dt: float, dtype):
    """sinc(c_ref |k| dt / 2) (operators.py:76-87)."""
    k2 = ks[0] ** 2
    for k in ks[1:]:
        k2 = k2 + k**2
    return np.sinc(c_ref * np.sqrt(k2) * (dt / (2.0 * np.pi))).astype(dtype)


def derivative_factors(ks, kappa, shape, cdtype):
    """i k_a kappa with k_a's Nyquist bin zeroed on even axes (operators.py:90-142)."""
    out = []
    for ax, k in enumerate(ks):
        if shape[ax] % 2 == 0:
            k = k.copy()
            idx = [0] * k.ndim
            idx[ax] = shape[ax] // 2
            k[tuple(idx)] = 0.0
        out.append((1j * k * kappa).astype(cdtype))
    return out


def sponge_volume(shape, width: int, edge: float, dtype):
    """Separable sponge product (operators.py:145-157)."""
    out = np.ones(shape, dtype=dtype)
    for ax, n in enumerate(shape):
        prof = sponge_profile_1d(n, width, edge, dtype)
        sh = [1] * len(shape)
        sh[ax] = n
        out = out * prof.reshape(sh)
    return out


# ------------------------------------------------------------------ schedule
@dataclass
class RunSpec:
    """Mirror of CWRunSpec defaults (solvers/base.py:55-83)."""

    cfl: float = 0.48
    cfl_hard_max: float = 0.50
    min_settle_periods: int = 8
    max_settle_periods: int = 96
    convergence_tol: float = 0.01
    n_record_periods: int = 2
    t_end_min_us: float | None = None


def cw_discretization(dx: float, c_max: float, spec: RunSpec, f0: float, harmonics=(1,)):
    """Exact-period (spp, dt) under the CFL limit (engine.py:86-113)."""
    period = 1.0 / f0
    dt_cfl = spec.cfl * dx / c_max
    spp = max(2, int(floor(period / dt_cfl)))
    dt = period / spp
    while dt * c_max / dx > spec.cfl_hard_max + 1e-12 and spp < 1024:
        spp += 1
        dt = period / spp
    if 2 * max(harmonics) >= spp:
        raise ValueError(f"harmonic {max(harmonics)} at/above temporal Nyquist for spp={spp}")
    return spp, dt


def cw_tof_periods(shape, dx: float, c_min: float, src_idx: np.ndarray, f0: float, ref=None):
    """Periods of flight to the farthest reference (engine.py:116-134)."""
    period = 1.0 / f0
    src_pos = src_idx.astype(np.float64) * dx
    if ref is not None:
        refs = np.asarray(ref, np.float64)[None, :] * dx
    else:
        refs = np.array(list(product(*((0, n - 1) for n in shape))), np.float64) * dx
    dmax = max(float(np.sqrt(((src_pos - r) ** 2).sum(axis=1)).max()) for r in refs)
    return max(1, int(ceil(dmax / c_min / period)))


def ramp_envelope(t: float, period: float, ramp_periods: float) -> float:
    """Cosine taper 0 -> 1 (sources.py:98-102)."""
    x = min(t / (ramp_periods * period), 1.0)
    return 0.5 * (1.0 - np.cos(np.pi * x))


class Diverged(RuntimeError):
    pass


@dataclass
class OracleResult:
    phasors: dict
    p_max: np.ndarray
    region: tuple
    dt: float
    spp: int
    steps_total: int
    tof_periods: int
    converged_period: int
    settle_capped: bool
    history: list = field(default_factory=list)
    c_ref: float = 0.0
    padded_shape: tuple = ()
    nonlinear_active: bool = False

    @property
    def phasor(self):
        return self.phasors[1]


class KSpaceOracle:
    """One CW solve on the CPU, step by step.

    Parameters mirror ``run_cw_kspace_pstd`` with plain arrays in place of
    Grid/Medium/CWSource. ``precision`` is "fp32" (upstream dtypes and order)
    or "fp64" (float64 everywhere, exact phase tables).
    """

    def __init__(
        self,
        shape,
        dx,
        pml_vox,
        pml_edge,
        alpha,
        rho,
        c,
        beta,
        src_idx,
        src_phase,
        amplitude,
        f0,
        ramp_periods=3.0,
        spec: RunSpec | None = None,
        harmonics=(1,),
        record_region=None,
        reference_point=None,
        nonlinear=False,
        precision="fp64",
        workers=1,
    ):
        if precision not in ("fp32", "fp64"):
            raise ValueError(precision)
        self.precision = precision
        self.fp32 = precision == "fp32"
        rdt = np.float32 if self.fp32 else np.float64
        cdt = np.complex64 if self.fp32 else np.complex128
        self.rdt, self.cdt = rdt, cdt
        self.workers = workers
        spec = spec or RunSpec()
        self.spec = spec
        self.shape = tuple(int(s) for s in shape)
        nd = len(self.shape)
        self.nd = nd
        self.dx = float(dx)
        self.harmonics = tuple(dict.fromkeys(int(h) for h in harmonics))
        alpha = np.ascontiguousarray(alpha, np.float32)
        rho = np.ascontiguousarray(rho, np.float32)
        c = np.ascontiguousarray(c, np.float32)
        beta = np.ascontiguousarray(beta, np.float32)
        self.f0 = float(f0)
        self.period = 1.0 / self.f0
        self.omega = 2.0 * np.pi * self.f0
        self.c_max = float(c.max())
        self.c_min = float(c.min())
        spp, dt = cw_discretization(self.dx, self.c_max, spec, self.f0, self.harmonics)
        self.spp, self.dt = spp, dt
        padded = pad_shape(self.shape)
        self.padded = padded

        if self.fp32:
            # engine.py:239-249 under NumPy-2 scalar promotion
            self.dt_over_rho = np.asarray(dt / pad_volume(rho, padded), dtype=np.float32)
            self.rhoc2_dt = np.asarray(
                pad_volume(rho * c.astype(np.float64) ** 2, padded) * dt, dtype=np.float32
            )
            self.absorb = np.asarray(np.exp(-pad_volume(alpha * c, padded) * dt), np.float32)
            beta_map = pad_volume(beta, padded) * (2.0 * dt)
        else:
            r64, c64, a64 = (v.astype(np.float64) for v in (rho, c, alpha))
            self.dt_over_rho = dt / pad_volume(r64, padded)
            self.rhoc2_dt = pad_volume(r64 * c64**2, padded) * dt
            self.absorb = np.exp(-pad_volume(a64 * c64, padded) * dt)
            beta_map = pad_volume(beta.astype(np.float64), padded) * (2.0 * dt)
        self.nonlinear = bool(nonlinear and not bool((beta == 0).all()))
        self.beta2_dt = np.asarray(beta_map, rdt) if self.nonlinear else None

        ks = k_vectors(padded, self.dx, rdt)
        kappa = kappa_sinc(ks, self.c_max, dt, rdt)
        self.deriv = derivative_factors(ks, kappa, padded, cdt)
        self.sponge = sponge_volume(padded, int(pml_vox), float(pml_edge), rdt)
        self.pml_vox = int(pml_vox)

        self.p = np.zeros(padded, rdt)
        self.u = [np.zeros(padded, rdt) for _ in range(nd)]

        src_idx = np.asarray(src_idx, np.int64)
        self.src_idx = tuple(src_idx[:, d] for d in range(nd))
        self.src_idx_arr = src_idx
        self.src_ph = np.asarray(src_phase, np.float32)
        self.amp = float(amplitude)
        self.ramp_periods = float(ramp_periods)
        c_src = c[self.src_idx]
        scale64 = 2.0 * c_src.astype(np.float64) * dt / self.dx
        self.src_scale = scale64.astype(np.float32) if self.fp32 else scale64

        self.tof_periods = cw_tof_periods(
            self.shape, self.dx, self.c_min, src_idx, self.f0, reference_point
        )
        self.eff_min = self.tof_periods + max(
            spec.min_settle_periods, int(ceil(self.ramp_periods)) + 1
        )
        self.eff_max = max(self.tof_periods + spec.max_settle_periods, self.eff_min)
        margin = self.pml_vox + 2
        if any(n <= 2 * margin for n in self.shape):
            margin = 0
        self.conv = tuple(slice(margin, n - margin) for n in self.shape)
        if record_region is None:
            self.rec = tuple(slice(0, n) for n in self.shape)
        else:
            self.rec = tuple(slice(*sl.indices(n)[:2]) for sl, n in zip(record_region, self.shape))
        self.axes = tuple(range(nd))

    # ----------------------------------------------------------------- step
    def _rfftn(self, a):
        return spfft.rfftn(a, workers=self.workers)

    def _irfftn(self, a):
        return spfft.irfftn(a, s=self.padded, axes=self.axes, workers=self.workers)

    def source_values(self, n: int):
        """Drive added at the source voxels after step n (engine.py:478-482)."""
        t = n * self.dt
        env = ramp_envelope(t, self.period, self.ramp_periods)
        if self.fp32:
            return (np.float32(self.amp * env) * self.src_scale) * np.sin(
                np.float32(self.omega * t) - self.src_ph
            )
        theta = 2.0 * np.pi * (n % self.spp) / self.spp
        return (self.amp * env) * self.src_scale * np.sin(theta - self.src_ph.astype(np.float64))

    def step(self, n: int) -> None:
        """One leapfrog step (engine.py:458-482)."""
        p, u, deriv = self.p, self.u, self.deriv
        pk = self._rfftn(p)
        for i in range(self.nd):
            grad_i = self._irfftn(deriv[i] * pk)
            u[i] -= self.dt_over_rho * grad_i
            u[i] *= self.absorb
            u[i] *= self.sponge
        acc = None
        for i in range(self.nd):
            term = deriv[i] * self._rfftn(u[i])
            acc = term if acc is None else acc + term
        divu = self._irfftn(acc)
        if self.beta2_dt is None:
            p -= self.rhoc2_dt * divu
        else:
            p -= (self.rhoc2_dt + self.beta2_dt * p) * divu
        p *= self.absorb
        p *= self.sponge
        p[self.src_idx] += self.source_values(n)

    def phase_factor(self, h: int, n: int):
        """exp(+i h omega t) of the record kernel (engine.py:548)."""
        if self.fp32:
            return np.exp(np.complex64(+1j * h * self.omega * (n * self.dt)))
        return np.exp(1j * 2.0 * np.pi * ((h * n) % self.spp) / self.spp)

    # ------------------------------------------------------------------ run
    def run(self) -> OracleResult:
        """Settle / record control loop (engine.py:484-586)."""
        spec, spp = self.spec, self.spp
        n = 0
        prev_peak = None
        converged = False
        converged_period = self.eff_max
        history = []
        periods_done = 0
        period_idx = 0
        while not converged and period_idx < self.eff_max:
            period_idx += 1
            period_peak = self.rdt(0.0)
            for _ in range(spp):
                self.step(n)
                n += 1
                period_peak = np.maximum(period_peak, np.abs(self.p[self.conv]).max())
            peak = float(period_peak)
            if not isfinite(peak):
                raise Diverged(f"settling period {period_idx} (step {n})")
            periods_done = period_idx
            rel = abs(peak - prev_peak) / max(prev_peak, 1e-9) if prev_peak is not None else None
            if period_idx >= self.tof_periods:
                history.append((period_idx, peak, float("nan") if rel is None else rel))
            if period_idx >= self.eff_min and rel is not None:
                if peak > 0.0 and rel < spec.convergence_tol:
                    converged_period = period_idx
                    converged = True
            if not converged:
                prev_peak = peak
        if spec.t_end_min_us is not None:
            need = ceil(spec.t_end_min_us * 1e-6 / self.period) - spec.n_record_periods
            while periods_done < need:
                for _ in range(spp):
                    self.step(n)
                    n += 1
                periods_done += 1
        rec_steps = spec.n_record_periods * spp
        rshape = self.p[self.rec].shape
        buffers = {h: np.zeros(rshape, self.cdt) for h in self.harmonics}
        pmax = np.zeros(rshape, self.rdt)
        for _ in range(rec_steps):
            self.step(n)
            pa = self.p[self.rec]
            for h in self.harmonics:
                buffers[h] += pa * self.phase_factor(h, n)
            np.maximum(pmax, np.abs(pa), out=pmax)
            n += 1
        scale = np.complex64(2.0 / rec_steps) if self.fp32 else 2.0 / rec_steps
        for h in self.harmonics:
            buffers[h] *= scale
        if not isfinite(float(pmax.max())):
            raise Diverged(f"record window (step {n})")
        return OracleResult(
            phasors=buffers,
            p_max=pmax,
            region=self.rec,
            dt=self.dt,
            spp=spp,
            steps_total=n,
            tof_periods=self.tof_periods,
            converged_period=converged_period,
            settle_capped=not converged,
            history=history,
            c_ref=self.c_max,
            padded_shape=self.padded,
            nonlinear_active=self.beta2_dt is not None,
        )


def run_cw(**kwargs) -> OracleResult:
    return KSpaceOracle(**kwargs).run()


def rel_l2(a, b) -> float:
    a = np.asarray(a, np.complex128)
    b = np.asarray(b, np.complex128)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))
