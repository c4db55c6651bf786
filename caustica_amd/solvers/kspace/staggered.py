"""Host setup of a solve of the opt-in staggered scheme (session scheme CST_SCHEME_STAGGERED).

The upstream solver has no such scheme: its step (``engine.py:458-482``) is collocated,
with frequency-independent ``exp(-alpha c dt)`` absorption and a multiplicative Gaussian
sponge, and power-law absorption is its unstarted milestone M16
(``MILESTONES.md:1201``: alpha(f) = alpha0 f^y with Kramers-Kronig dispersion). The
device scheme (``csrc/staggered_kernels.h``) steps the first-order equations on a
staggered grid:

  u_a   <- s'_a (s'_a u_a - dt / rho0'_a d+_a p)
  rho_a <- s_a (s_a rho_a - dt (rho0 + 2 m rho) d-_a u_a),   rho_a[src] += q(n) / (nd c^2)
  p = c^2 (rho + tau L1 - eta L2 + m (beta - 1) rho^2 / rho0),  rho = sum_a rho_a

with d+-_a = IFFT(i k_a kappa exp(+-i k_a dx / 2) FFT) (Tabei, Mast and Waag, JASA 111(1)
2002), L1 = IFFT(|k|^(y-2) FFT(rho0 sum_a d-_a u_a)), L2 = IFFT(|k|^(y-1) FFT(rho)),
tau = -2 alpha0 c^(y-1), eta = 2 alpha0 c^y tan(pi y / 2), alpha0 = alpha / omega0^y
(Treeby and Cox, JASA 127(5) 2010; the medium's alpha is the absorption at f0), rho0'_a
the mean of rho0 at x_a and x_a + dx (the last cell keeps its own), s_a, s'_a =
exp(-sigma dt / 2) with sigma(d) = pml_alpha (c_max / dx) (d / w)^4 at depth d into the
w-cell layer at grid or staggered points (Berenger's split field: rho_a and u_a damped
along axis a only), m = 1 where beta > 0 on nonlinear solves. Everything else -- the
exact-period dt, the settle / record schedule, the source drive q(n) = amp env(t)
(2 c dt / dx) sin(omega t - phi), the period peaks and phasor record -- is the
reference scheme's (``plan_member``).
"""

from __future__ import annotations

import numpy as np

from caustica_amd.solvers.kspace.operators import MemberPlan, pad_volume, plan_member

SCHEME = "cw-kspace-pstd-staggered/1"


def pml_profile(n: int, width: int, sigma_max: float, dt: float, staggered: bool) -> np.ndarray:
    """float64 exp(-sigma(d) dt / 2) at grid points (q = i) or staggered points (q = i + 1/2)
    of an axis of n cells, d the depth into the ``width``-cell layer at either end."""
    q = np.arange(n, dtype=np.float64) + (0.5 if staggered else 0.0)
    if width <= 0:
        return np.ones(n)
    d = np.clip(np.maximum(width - q, width - (n - 1 - q)), 0.0, float(width))
    return np.exp(-sigma_max * (d / width) ** 4 * dt / 2.0)


def staggered_mean(v: np.ndarray, axis: int) -> np.ndarray:
    """Mean of each cell and its successor along ``axis`` (the last cell keeps its own)."""
    nxt = np.concatenate([np.take(v, np.arange(1, v.shape[axis]), axis=axis),
                          np.take(v, [v.shape[axis] - 1], axis=axis)], axis=axis)
    return 0.5 * (v + nxt)


def shift_factors(padded3, dx: float) -> np.ndarray:
    """complex64 k_a exp(+i k_a dx / 2) for a = x, y, z (fftfreq on x, y; rfftfreq on z)."""
    out = []
    for ax, n in enumerate(padded3):
        f = np.fft.rfftfreq(n, d=dx) if ax == 2 else np.fft.fftfreq(n, d=dx)
        k = 2.0 * np.pi * f
        out.append(k * np.exp(0.5j * k * dx))
    return np.ascontiguousarray(np.concatenate(out), dtype=np.complex64)


def pml_tables(padded3, nd: int, width: int, sigma_max: float, dt: float) -> np.ndarray:
    """float32 [4][nz + nx + ny]: table 0 at grid points on every axis, table 1 + a at
    staggered points along embedded axis a (ones along the others); axes that only embed
    a rank-1/2 grid in 3-D are all ones. Each table is laid out sz, sx, sy."""
    nx, ny, nz = padded3

    def prof(ax, staggered):
        if ax < 3 - nd:
            return np.ones(padded3[ax])
        return pml_profile(padded3[ax], width, sigma_max, dt, staggered)

    def table(fx, fy, fz):
        return np.concatenate([fz, fx, fy])

    ones = [np.ones(n) for n in padded3]
    t0 = table(prof(0, False), prof(1, False), prof(2, False))
    tx = table(prof(0, True), ones[1], ones[2])
    ty = table(ones[0], prof(1, True), ones[2])
    tz = table(ones[0], ones[1], prof(2, True))
    return np.ascontiguousarray(np.stack([t0, tx, ty, tz]), dtype=np.float32)


def plan_staggered(grid, medium, source, spec, *, harmonics=(1,), record_region=None,
                   reference_point=None, nonlinear=False, alpha_power: float = 1.5,
                   pml_alpha: float = 2.0) -> MemberPlan:
    """MemberPlan of the staggered scheme: the reference schedule, source and boxes, with
    the scheme's maps and tables in ``plan.stg`` (the reference maps are dropped)."""
    y = float(alpha_power)
    if not (0.0 < y < 3.0) or abs(y - 1.0) < 1e-6:
        raise ValueError(f"alpha_power must lie in (0, 3) and differ from 1, got {y}")
    if not pml_alpha >= 0.0:
        raise ValueError(f"pml_alpha must be >= 0, got {pml_alpha}")
    plan = plan_member(grid, medium, source, spec, harmonics=harmonics, record_region=record_region,
                       reference_point=reference_point, nonlinear=nonlinear)
    nd = grid.ndim
    padded, padded3, dt, dx = plan.padded, plan.padded3, plan.dt, plan.dx

    def vol(v):
        return pad_volume(np.asarray(v, np.float32).astype(np.float64), padded).reshape(padded3)

    rho, c = vol(medium.rho), vol(medium.c)
    alpha, beta = vol(medium.alpha), vol(medium.beta)
    omega = 2.0 * np.pi * plan.f0
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)  # noqa: E731
    maps: list = [f32(dt / staggered_mean(rho, a)) for a in range(3)]
    maps += [f32(dt * rho), f32(c * c)]
    absorbing = bool((alpha > 0).any())
    if absorbing:
        a0 = alpha / omega**y
        tau = -2.0 * a0 * c ** (y - 1.0)
        eta = 2.0 * a0 * c**y * np.tan(np.pi * y / 2.0)
        maps += [f32(c * c * tau / dt), f32(c * c * eta)]
    else:
        maps += [None, None]
    nl_on = bool(nonlinear and not medium.is_linear)
    if nl_on:
        m = (beta > 0).astype(np.float64)
        maps += [f32(2.0 * m * dt), f32(c * c * m * (beta - 1.0) / rho)]
    else:
        maps += [None, None]
    sigma_max = float(pml_alpha) * plan.c_ref / dx
    idx = source.indices
    c_src = np.asarray(medium.c, np.float32)[tuple(idx[:, d] for d in range(nd))].astype(np.float64)
    plan.src_scale = f32(2.0 * c_src * dt / dx / (nd * c_src * c_src))
    plan.dt_over_rho = plan.rhoc2_dt = plan.absorb = plan.beta2_dt = None
    plan.sponge = None
    plan.scheme = "staggered"
    plan.stg = {
        "maps": maps,
        "kfac": shift_factors(padded3, dx),
        "pml": pml_tables(padded3, nd, grid.pml_vox, sigma_max, dt),
        "alpha_power": y,
        "pml_alpha": float(pml_alpha),
        "absorbing": absorbing,
        "nonlinear": nl_on,
        "src_axes": sum(1 << a for a in range(3 - nd, 3)),
    }
    return plan
