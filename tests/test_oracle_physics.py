"""Pin the oracle's fp64 mode against what the mathematics fixes (no fixture needed).

Each check is chosen so that a plausible slip in the oracle (a wrong sign of i k, a
dropped kappa, a transposed axis, a missing factor 2 in the source scale or in the
phasor normalization, alpha/2 instead of alpha) fails one of them.
"""

import numpy as np
import pytest

from oracle.analytic import fubini, oneill_axial, rayleigh_pressure, shock_distance
from oracle.heating import heat_source
from oracle.kspace import KSpaceOracle, RunSpec, derivative_factors, k_vectors, kappa_sinc
from oracle.pennes import PennesOracle, cem43_rate
from oracle.sources import plane_indices, spherical_cap_points

F0, C0 = 1.0e6, 1500.0


def _plane_1d(n=256, pml=24, src=40, alpha=0.0, beta=0.0, amp=1e5, ppw=4.0, settle=(40, 90),
              tol=0.005, harmonics=(1,), nonlinear=False):
    dx = C0 / (F0 * ppw)
    shape = (n,)
    full = np.ones(shape, np.float32)
    spec = RunSpec(min_settle_periods=settle[0], max_settle_periods=settle[1], convergence_tol=tol)
    return KSpaceOracle(shape=shape, dx=dx, pml_vox=pml, pml_edge=2.0, alpha=full * alpha,
                        rho=full * 1000.0, c=full * C0, beta=full * beta,
                        src_idx=plane_indices(shape, 0, src), src_phase=np.zeros(1, np.float32),
                        amplitude=amp, f0=F0, spec=spec, harmonics=harmonics,
                        nonlinear=nonlinear, precision="fp64").run(), dx


@pytest.fixture(scope="module")
def lossless():
    return _plane_1d()


def test_exact_period_policy(lossless):
    res, dx = lossless
    assert res.spp == 8 and res.dt == pytest.approx(1.0 / F0 / 8, rel=1e-12)


def test_plane_wave_phase_speed(lossless):
    res, dx = lossless
    ph = res.phasor[80:200]
    k_num = np.angle(ph[1:] * np.conj(ph[:-1])).mean() / dx
    # outgoing wave is exp(+ikx) in the p = Re{P e^{-i w t}} convention
    assert abs(k_num - 2 * np.pi * F0 / C0) / (2 * np.pi * F0 / C0) < 1e-3


def test_mass_source_amplitude_and_pml_ripple(lossless):
    res, dx = lossless
    amp = np.abs(res.phasor[80:200])
    assert 0.9 < amp.mean() / 1e5 < 1.12  # 2 c dt / dx mass-source normalization
    assert (amp.max() - amp.min()) / (amp.max() + amp.min()) < 0.03


def test_phasor_against_time_peak(lossless):
    res, dx = lossless
    ratio = np.abs(res.phasor[80:200]) / res.p_max[80:200]
    assert ratio.max() <= 1 / np.cos(np.pi / res.spp) * (1 + 1e-3) and ratio.min() > 0.9


def test_absorption_slope():
    res, dx = _plane_1d(n=288, pml=32, alpha=30.0, settle=(45, 95))
    x = np.arange(80, 200) * dx
    slope = np.polyfit(x, np.log(np.abs(res.phasor[80:200])), 1)[0]
    assert abs(-slope - 30.0) / 30.0 < 0.01


def test_linearity_in_drive_amplitude():
    a, _ = _plane_1d(n=96, pml=12, src=20, settle=(10, 10), tol=0.0, amp=1e5)
    b, _ = _plane_1d(n=96, pml=12, src=20, settle=(10, 10), tol=0.0, amp=3e5)
    np.testing.assert_allclose(b.phasor, 3 * a.phasor, rtol=1e-10, atol=1e-6)


def test_beta_zero_westervelt_equals_linear():
    a, _ = _plane_1d(n=96, pml=12, src=20, settle=(6, 6), tol=0.0)
    b, _ = _plane_1d(n=96, pml=12, src=20, settle=(6, 6), tol=0.0, nonlinear=True)
    np.testing.assert_array_equal(a.phasor, b.phasor)


@pytest.mark.slow
def test_fubini_second_harmonic():
    beta = 3.5
    res, dx = _plane_1d(n=600, pml=40, src=60, beta=beta, amp=2e6, ppw=16.0, settle=(45, 100),
                        tol=0.003, harmonics=(1, 2), nonlinear=True)
    a1, a2 = np.abs(res.phasors[1]), np.abs(res.phasors[2])
    i0 = 90
    sig0 = ((i0 - 60) * dx) / shock_distance(a1[i0], F0, beta=beta)
    x_sh = shock_distance(a1[i0] / fubini(1, sig0), F0, beta=beta)
    checked = 0
    for iv in (150, 250, 350, 450, 540):
        sigma = ((iv - 60) * dx) / x_sh
        if 0.05 <= sigma <= 0.95:
            rel = abs(a2[iv] / a1[iv] - fubini(2, sigma) / fubini(1, sigma)) / (fubini(2, sigma) / fubini(1, sigma))
            assert rel < 0.05
            checked += 1
    assert checked >= 4


def test_derivative_factor_differentiates_resolved_modes():
    """Re(ifft(i k kappa fft f)) with kappa -> 1 (dt -> 0) is the exact derivative."""
    n, dx = 64, 1e-3
    ks = k_vectors((n,), dx, np.float64)
    kap = kappa_sinc(ks, 1500.0, 1e-15, np.float64)
    d = derivative_factors(ks, kap, (n,), np.complex128)[0]
    x = np.arange(n) * dx
    for m in (1, 5, 17, 31):
        f = np.sin(2 * np.pi * m * x / (n * dx))
        df = np.fft.irfft(d * np.fft.rfft(f), n=n)
        np.testing.assert_allclose(df, 2 * np.pi * m / (n * dx) * np.cos(2 * np.pi * m * x / (n * dx)),
                                   atol=1e-9 * 2 * np.pi * m / (n * dx))
    # the Nyquist mode cannot be differentiated on a collocated grid: it maps to 0
    f = np.cos(np.pi * np.arange(n))
    assert np.abs(np.fft.irfft(d * np.fft.rfft(f), n=n)).max() < 1e-9


def test_kappa_is_sinc_of_half_step_phase():
    ks = k_vectors((16,), 1e-3, np.float64)
    kap = kappa_sinc(ks, 1500.0, 2e-7, np.float64)
    arg = 1500 * np.abs(ks[0]) * 1e-7
    expect = np.ones_like(arg)
    np.divide(np.sin(arg), arg, out=expect, where=arg != 0)
    np.testing.assert_allclose(kap, expect, rtol=1e-12)


def test_oneill_limits():
    a, roc = 7.5e-3, 18e-3
    k = 2 * np.pi * F0 / C0
    h = roc - np.sqrt(roc**2 - a**2)
    assert abs(oneill_axial(roc, a, roc, F0)[0]) == pytest.approx(1000 * C0 * k * h, rel=1e-12)
    z = roc * (1 - 1e-7)
    assert abs(oneill_axial(z, a, roc, F0)[0]) == pytest.approx(1000 * C0 * k * h, rel=1e-5)


def test_rayleigh_point_source_and_cap_axis():
    """One point source: |p| = rho c k v dS / (2 pi R) with phase k R - pi/2; a finely
    sampled spherical cap reproduces O'Neil's on-axis closed form (its exact Rayleigh
    integral) to sampling error."""
    from oracle.sources import spherical_cap_points

    k = 2 * np.pi * F0 / C0
    r = 0.031
    p = rayleigh_pressure(np.zeros((1, 3)), np.array([1e-6]), 0.5, np.array([[0.0, 0.0, r]]), k)[0]
    assert abs(p) == pytest.approx(1000 * C0 * k * 0.5e-6 / (2 * np.pi * r), rel=1e-12)
    assert np.angle(p * np.exp(-1j * k * r)) == pytest.approx(-np.pi / 2, abs=1e-12)
    a, roc = 7.5e-3, 18e-3
    pts, areas = spherical_cap_points(a, roc, C0 / F0 / 20)
    z = np.linspace(0.3 * roc, 1.6 * roc, 40)
    tg = np.column_stack([np.zeros_like(z), np.zeros_like(z), z])
    ratio = np.abs(rayleigh_pressure(pts, areas, 1.0, tg, k)) / np.abs(oneill_axial(z, a, roc, F0))
    assert np.abs(ratio - 1).max() < 1e-5


def test_fubini_small_sigma_and_energy():
    assert fubini(1, 1e-6) == pytest.approx(1.0, abs=1e-10)
    assert fubini(2, 1e-3) == pytest.approx(0.5e-3, rel=1e-5)


def test_spherical_cap_points_on_sphere():
    pts, areas = spherical_cap_points(8e-3, 20e-3, 0.1875e-3)
    r = np.linalg.norm(pts - np.array([0, 0, 20e-3]), axis=1)
    np.testing.assert_allclose(r, 20e-3, rtol=1e-12)
    assert areas.sum() == pytest.approx(2 * np.pi * 20e-3**2 * (1 - np.sqrt(1 - 0.16)), rel=1e-12)


def test_heat_source_is_two_alpha_intensity():
    p = np.full((3, 3), 2e5 + 0j)
    q = heat_source({1: p}, alpha=5.0, rho=1000.0, c=1500.0)
    intensity = 2e5**2 / (2 * 1000 * 1500)
    np.testing.assert_allclose(q, 2 * 5.0 * intensity, rtol=1e-14)


def test_cem43_unit_definitions():
    assert float(cem43_rate(np.float64(43.0))) == 1.0
    assert float(cem43_rate(np.float64(44.0))) == 2.0
    assert float(cem43_rate(np.float64(42.0))) == 0.25


def test_pennes_gaussian_spreading():
    """Insulated homogeneous block: a Gaussian spreads as sigma^2 = sigma0^2 + 2 D t per axis."""
    n, dx = 48, 0.5e-3
    k, rc = 0.5, 3.6e6
    D = k / rc
    shape = (n, n, n)
    x = (np.arange(n) - (n - 1) / 2) * dx
    s0 = 2.5e-3
    r2 = x[:, None, None] ** 2 + x[None, :, None] ** 2 + x[None, None, :] ** 2
    t0 = (37.0 + np.exp(-r2 / (2 * s0**2))).astype(np.float32)
    o = PennesOracle(np.full(shape, k), np.full(shape, 1000.0), np.full(shape, 3600.0),
                     np.zeros(shape), dx, precision="fp64")
    dt = 0.5 * o.stable_dt()
    steps = 200
    out = o.solve(t0, None, dt, steps)
    theta = out.temperature - 37.0
    prof = theta[:, n // 2, n // 2]
    var = (prof * x**2).sum() / prof.sum()
    assert var == pytest.approx(s0**2 + 2 * D * dt * steps, rel=0.02)
    # energy conservation with insulated walls
    assert theta.sum() == pytest.approx((t0.astype(np.float64) - 37.0).sum(), rel=1e-9)


def test_pennes_perfusion_relaxation():
    """Uniform excess temperature decays as exp(-w rho_b C_b t / (rho C))."""
    shape = (6, 6, 6)
    w = 0.5 / 1050.0
    o = PennesOracle(np.full(shape, 0.5), np.full(shape, 1000.0), np.full(shape, 3600.0),
                     np.full(shape, w), 1e-3, precision="fp64")
    dt = 0.05
    out = o.solve(np.full(shape, 39.0, np.float32), None, dt, 400)
    lam = w * 1050.0 * 3617.0 / 3.6e6
    expect = 2.0 * (1 - lam * dt) ** 400  # forward Euler of the exponential
    np.testing.assert_allclose(out.temperature - 37.0, expect, rtol=1e-9)
