"""Staggered-grid k-space scheme with power-law absorption and a split-field PML, on the CPU
(test oracle; see oracle/__init__.py).

The upstream solver has no such scheme: its step (``engine.py:458-482``) is collocated,
with frequency-independent ``exp(-alpha c dt)`` absorption and a multiplicative Gaussian
sponge, and the power-law scheme was planned but never built (upstream
``MILESTONES.md:1201``: alpha(f) = alpha0 f^y with Kramers-Kronig dispersion, "flag off ->
bit-unchanged"). This module is the float64 statement of the opt-in scheme the hip
backend runs when a solve asks for ``scheme="staggered"``; it is pinned by closed forms
(tests/test_oracle_staggered.py: shifted spectral derivatives of sinusoids, lossless
phase speed, the absorption slope y and magnitude alpha0 omega^y, the Kramers-Kronig
phase-speed change, the PML reflection, beta = 0 identity), and the GPU path is checked
against it (tests/test_gpu_staggered.py).

Equations (first-order fluid form; Tabei, Mast and Waag, JASA 111(1) 2002 for the
staggered k-space derivatives; Treeby and Cox, JASA 127(5) 2010 for the fractional-
Laplacian absorption; Berenger's split field for the PML). Per step n, for every active
axis a with grid spacing dx:

  d_a^+ f = IFFT(i k_a kappa exp(+i k_a dx / 2) FFT f)        (value at x_a + dx/2)
  d_a^- f = IFFT(i k_a kappa exp(-i k_a dx / 2) FFT f)        (value at x_a - dx/2)
  u_a   <- s'_a (s'_a u_a - dt / rho0'_a d_a^+ p)             (u_a lives at x_a + dx/2)
  rho_a <- s_a (s_a rho_a - dt (rho0 + 2 m rho) d_a^- u_a)      (split density)
  rho_a[src] += q(n) / (nd c^2)                                (mass source, all axes)
  rho   = sum_a rho_a
  p     = c^2 (rho + tau L1 - eta L2 + m (beta - 1) rho^2 / rho0)
  L1    = IFFT(|k|^(y-2) FFT(rho0 sum_a d_a^- u_a)),  L2 = IFFT(|k|^(y-1) FFT(rho))
  tau   = -2 alpha0 c^(y-1),  eta = 2 alpha0 c^y tan(pi y / 2),  alpha0 = alpha / omega0^y

kappa = sinc(c_ref dt |k| / 2) with c_ref = c_max; |k|^(y-2) and |k|^(y-1) are zero at
k = 0; rho0'_a is the mean of rho0 at x_a and x_a + dx (rho0 at the last cell of the
axis); s_a, s'_a = exp(-sigma dt / 2) with sigma(d) = pml_alpha (c_max / dx) (d / w)^4,
d the depth into a w-cell layer at grid or staggered positions (split: u_a and rho_a are
damped along axis a only); m = 1 where beta > 0 on nonlinear solves, else 0 (beta = 1 +
B / 2A); q(n) is the upstream mass-source increment amp env(t) (2 c dt / dx)
sin(omega t - phi). The schedule (exact-period dt, settle / record loop, phasor and
p_max record, period peaks) is the upstream one, inherited from oracle.kspace.
"""

from __future__ import annotations

import numpy as np

from oracle.kspace import KSpaceOracle, k_vectors, kappa_sinc, pad_volume


def pml_profile(n: int, width: int, sigma_max: float, dt: float, staggered: bool) -> np.ndarray:
    """exp(-sigma(d) dt / 2) at grid points (q = i) or staggered points (q = i + 1/2) of an
    axis of n cells, d = depth into the layer of ``width`` cells at either end."""
    q = np.arange(n, dtype=np.float64) + (0.5 if staggered else 0.0)
    if width <= 0:
        return np.ones(n)
    d = np.maximum(np.maximum(width - q, width - (n - 1 - q)), 0.0)
    d = np.minimum(d, float(width))
    return np.exp(-sigma_max * (d / width) ** 4 * dt / 2.0)


def staggered_density(rho: np.ndarray, axis: int) -> np.ndarray:
    """rho0 at x_a + dx/2: mean of the cell and its successor along ``axis`` (the last cell
    keeps its own value)."""
    nxt = np.concatenate([np.take(rho, np.arange(1, rho.shape[axis]), axis=axis),
                          np.take(rho, [rho.shape[axis] - 1], axis=axis)], axis=axis)
    return 0.5 * (rho + nxt)


def shift_derivative(ks, kappa, axis: int, dx: float, sign: int):
    """i k_a kappa exp(sign i k_a dx / 2) broadcast over the half spectrum."""
    k = ks[axis].astype(np.float64)
    return 1j * k * np.exp(sign * 0.5j * k * dx) * kappa


class StaggeredOracle(KSpaceOracle):
    """One CW solve of the staggered scheme in float64 (parameters as KSpaceOracle, plus
    ``alpha_power`` y and ``pml_alpha``; the medium's alpha is the absorption at f0 in
    Np/m). A medium with alpha = 0 everywhere skips the fractional-Laplacian terms."""

    def __init__(self, *, shape, dx, pml_vox, alpha, rho, c, beta, src_idx, src_phase,
                 amplitude, f0, ramp_periods=3.0, spec=None, harmonics=(1,), record_region=None,
                 reference_point=None, nonlinear=False, alpha_power: float = 1.5,
                 pml_alpha: float = 2.0, workers: int = 1):
        super().__init__(shape=shape, dx=dx, pml_vox=pml_vox, pml_edge=2.0, alpha=alpha, rho=rho,
                         c=c, beta=beta, src_idx=src_idx, src_phase=src_phase,
                         amplitude=amplitude, f0=f0, ramp_periods=ramp_periods, spec=spec,
                         harmonics=harmonics, record_region=record_region,
                         reference_point=reference_point, nonlinear=nonlinear,
                         precision="fp64", workers=workers)
        y = float(alpha_power)
        if not (0.0 < y < 3.0) or abs(y - 1.0) < 1e-6:
            raise ValueError(f"alpha_power must lie in (0, 3) and differ from 1, got {y}")
        self.y = y
        padded, nd, dx, dt = self.padded, self.nd, self.dx, self.dt

        def vol(v):
            return pad_volume(np.asarray(v, np.float32).astype(np.float64), padded)

        rho, c, alpha, beta = vol(rho), vol(c), vol(alpha), vol(beta)
        self.rho0 = rho
        self.c2 = c * c
        self.dtrho_sg = [dt / staggered_density(rho, a) for a in range(nd)]
        a0 = alpha / self.omega**y
        self.absorbing = bool((alpha > 0).any())
        self.tau = -2.0 * a0 * c ** (y - 1.0)
        self.eta = 2.0 * a0 * c**y * np.tan(np.pi * y / 2.0)
        m = (beta > 0).astype(np.float64) if self.nonlinear else np.zeros_like(beta)
        self.nl_m = m
        self.bh = m * (beta - 1.0)
        ks = k_vectors(padded, dx, np.float64)
        kappa = kappa_sinc(ks, self.c_max, dt, np.float64)
        self.dplus = [shift_derivative(ks, kappa, a, dx, +1) for a in range(nd)]
        self.dminus = [shift_derivative(ks, kappa, a, dx, -1) for a in range(nd)]
        kk = sum(k.astype(np.float64) ** 2 for k in ks)
        kmag = np.sqrt(kk)
        with np.errstate(divide="ignore"):
            self.nabla1 = np.where(kmag > 0, kmag ** (y - 2.0), 0.0)
            self.nabla2 = np.where(kmag > 0, kmag ** (y - 1.0), 0.0)
        w = int(self.pml_vox)
        smax = float(pml_alpha) * self.c_max / dx
        self.s = []
        self.s_sg = []
        for a in range(nd):
            sh = [1] * nd
            sh[a] = padded[a]
            self.s.append(pml_profile(padded[a], w, smax, dt, False).reshape(sh))
            self.s_sg.append(pml_profile(padded[a], w, smax, dt, True).reshape(sh))
        self.rho_a = [np.zeros(padded) for _ in range(nd)]
        c_src = c[self.src_idx]
        self.src_rho = 1.0 / (nd * c_src * c_src)  # density per unit pressure, per axis
        self.src_rho_total = 1.0 / (c_src * c_src)

    def step(self, n: int) -> None:
        nd = self.nd
        pk = self._rfftn(self.p)
        for a in range(nd):
            g = self._irfftn(self.dplus[a] * pk)
            s = self.s_sg[a]
            self.u[a] = s * (s * self.u[a] - self.dtrho_sg[a] * g)
        d = [self._irfftn(self.dminus[a] * self._rfftn(self.u[a])) for a in range(nd)]
        if self.nonlinear:
            fac = self.dt * (self.rho0 + 2.0 * self.nl_m * sum(self.rho_a))
        else:
            fac = self.dt * self.rho0
        for a in range(nd):
            s = self.s[a]
            self.rho_a[a] = s * (s * self.rho_a[a] - fac * d[a])
        q = self.source_values(n)
        for a in range(nd):
            self.rho_a[a][self.src_idx] += q * self.src_rho
        rho = sum(self.rho_a)
        total = rho
        if self.absorbing:
            l1 = self._irfftn(self.nabla1 * self._rfftn(self.rho0 * sum(d)))
            l2 = self._irfftn(self.nabla2 * self._rfftn(rho))
            total = rho + self.tau * l1 - self.eta * l2
        if self.nonlinear:
            total = total + self.bh * rho * rho / self.rho0
        self.p = self.c2 * total
