"""The opt-in staggered scheme's float64 oracle (oracle/staggered.py) against closed forms.

The upstream solver has no staggered / power-law / split-PML scheme (its power-law
milestone, upstream MILESTONES.md:1201, names the criteria: alpha(f) = alpha0 f^y and
Kramers-Kronig dispersion), so nothing upstream-written pins it; these do:
  * the shifted spectral derivatives of a sinusoid equal its derivative half a cell
    forward (+) or back (-), on even and odd lengths (a wrong sign of the shift fails);
  * a lossless 1-D plane wave runs at c0 (the kappa correction makes it exact) and the
    split-field PML reflects less than 1e-4 of it (standing-wave ripple);
  * with power-law absorption the amplitude decays as the exact dispersion relation of
    the fractional-Laplacian equations predicts (oracle.analytic.powerlaw_wavenumber:
    alpha0 omega^y to first order) within 2 %, the log-log slope over an octave is y
    within 0.03, and the phase speed carries the Kramers-Kronig shift (1.5 % at
    y = 1.1) within 0.06 % -- dropping or flipping the tau or eta term fails these;
  * the nonlinear term generates the Fubini second and third harmonics (A2/A1 within
    3 %, A3/A1 within 5 %), and beta = 0 is bitwise the linear solve.
"""

import numpy as np
import pytest

from oracle.analytic import fubini, powerlaw_wavenumber, shock_distance
from oracle.kspace import RunSpec, k_vectors
from oracle.staggered import StaggeredOracle, shift_derivative

C0 = 1500.0


def _run_1d(f0=1e6, alpha=0.0, y=1.5, n=512, dx=0.1e-3, pml=20, src=40, amp=1e5, beta=0.0,
            nonlinear=False, harmonics=(1,), settle=None):
    tof = int(np.ceil((n - src) * dx / C0 * f0)) + 2
    settle = settle if settle is not None else tof + 4
    spec = RunSpec(cfl=0.3, min_settle_periods=settle, max_settle_periods=settle,
                   convergence_tol=0.0, n_record_periods=2)
    o = StaggeredOracle(
        shape=(n,), dx=dx, pml_vox=pml, alpha=np.full(n, alpha, np.float32),
        rho=np.full(n, 1000.0, np.float32), c=np.full(n, C0, np.float32),
        beta=np.full(n, beta, np.float32), src_idx=np.array([[src]]),
        src_phase=np.zeros(1, np.float32), amplitude=amp, f0=f0, spec=spec,
        harmonics=harmonics, nonlinear=nonlinear, alpha_power=y)
    return o.run()


def _fit(ph, x0=150, x1=400, dx=0.1e-3):
    """(alpha, k) of a travelling wave P = A exp((i k - alpha) x) by least squares."""
    x = np.arange(x0, x1) * dx
    a = np.vstack([np.ones_like(x), x]).T
    lam = np.linalg.lstsq(a, np.log(np.abs(ph[x0:x1])), rcond=None)[0][1]
    k = np.linalg.lstsq(a, np.unwrap(np.angle(ph[x0:x1])), rcond=None)[0][1]
    return -lam, k


@pytest.mark.parametrize("n", [32, 27])
def test_shifted_derivatives_of_a_sinusoid(n):
    dx = 1e-3
    x = np.arange(n) * dx
    ks = k_vectors((n,), dx, np.float64)
    for m in (1, 3, (n - 1) // 2 - 1):
        kk = 2.0 * np.pi * m / (n * dx)
        f = np.sin(kk * x)
        for sign in (+1, -1):
            d = np.fft.irfft(shift_derivative(ks, 1.0, 0, dx, sign) * np.fft.rfft(f), n)
            want = kk * np.cos(kk * (x + sign * dx / 2))
            assert np.abs(d - want).max() < 1e-9 * kk, (m, sign)


@pytest.fixture(scope="module")
def lossless():
    return _run_1d()


def test_lossless_plane_wave_speed_is_c0(lossless):
    alpha, k = _fit(lossless.phasors[1])
    assert abs(2 * np.pi * 1e6 / k / C0 - 1.0) < 1e-4
    assert abs(alpha) < 1e-3


def test_split_pml_reflection_below_1e_4(lossless):
    a = np.abs(lossless.phasors[1][150:400])
    assert (a.max() - a.min()) / (a.max() + a.min()) < 1e-4


@pytest.mark.parametrize("y", [1.5, 1.1])
def test_power_law_absorption_and_dispersion(y):
    meas = {}
    for f0 in (1e6, 2e6):
        # the medium's alpha is given at the drive frequency: alpha0 is the same in both
        res = _run_1d(f0=f0, alpha=10.0 * (f0 / 1e6) ** y, y=y)
        meas[f0] = _fit(res.phasors[1])
        k = powerlaw_wavenumber(f0, 10.0, 1e6, y)
        assert abs(meas[f0][0] / k.imag - 1.0) < 0.02, (f0, meas[f0][0], k.imag)
        cp, cp_want = 2 * np.pi * f0 / meas[f0][1], 2 * np.pi * f0 / k.real
        assert abs(cp / cp_want - 1.0) < 6e-4, (f0, cp, cp_want)
    slope = np.log(meas[2e6][0] / meas[1e6][0]) / np.log(2.0)
    assert abs(slope - y) < 0.03, slope
    # the dispersion the tests above hold is not vanishing: c_phase moves off c0
    assert 2 * np.pi * 1e6 / meas[1e6][1] - C0 > (1.0 if y > 1.2 else 15.0)


def test_powerlaw_wavenumber_limits():
    k = powerlaw_wavenumber(1e6, 1e-6, 1e6, 1.5)
    assert abs(k.real / (2 * np.pi * 1e6 / C0) - 1.0) < 1e-8
    assert abs(k.imag / 1e-6 - 1.0) < 1e-6


@pytest.fixture(scope="module")
def fubini_run():
    dx = C0 / (1e6 * 16.0)
    return _run_1d(n=600, dx=dx, pml=40, src=60, amp=2e6, beta=3.5, nonlinear=True,
                   harmonics=(1, 2, 3), settle=45), dx


def test_fubini_harmonics(fubini_run):
    res, dx = fubini_run
    a1, a2, a3 = (np.abs(res.phasors[h]) for h in (1, 2, 3))
    sig0 = (30 * dx) / shock_distance(a1[90], 1e6, C0, beta=3.5)
    x_sh = shock_distance(a1[90] / fubini(1, sig0), 1e6, C0, beta=3.5)
    for iv in (150, 250, 350, 450, 540):
        s = ((iv - 60) * dx) / x_sh
        assert abs(a2[iv] / a1[iv] / (fubini(2, s) / fubini(1, s)) - 1.0) < 0.03, s
        assert abs(a3[iv] / a1[iv] / (fubini(3, s) / fubini(1, s)) - 1.0) < 0.05, s


def test_beta_zero_is_the_linear_solve():
    lin = _run_1d(n=128, src=30, settle=8)
    nl = _run_1d(n=128, src=30, settle=8, nonlinear=True, beta=0.0)
    assert np.array_equal(lin.phasors[1], nl.phasors[1])
