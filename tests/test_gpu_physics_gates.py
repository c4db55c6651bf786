"""The upstream physics gates, run through the hip engine with the upstream thresholds.

Each test restates one upstream gate with ``backend="hip"`` in place of numpy:
  * 1-D plane waves (upstream tests/test_linear_planewave.py:54-117): exact-period dt
    policy, phase-speed error < 0.1 %, phasor within the cos(pi/spp) ceiling of the
    time peak, absorption slope within 1 % of the configured alpha, PML ripple < 3 %,
    200-period stability (late peak drift < 1 %), honest convergence metadata;
  * the 3-D focused bowl in water against O'Neil and Rayleigh (upstream
    tests/test_linear_oneill_3d.py:62-133): focus within one voxel of O'Neil's
    maximum, axial correlation > 0.99 and -6 dB axial width within 5 %, lateral -6 dB
    width within 5 % and correlation > 0.98 against the Rayleigh integral, metadata;
  * Westervelt against Fubini (upstream tests/test_westervelt.py:39-121): A2/A1
    within 5 % at four or more sigma stations, A3/A1 within 10 %, amp <= p_max /
    cos(pi/spp) in the > 10 % band with the near-source f0 fraction in [0.85, 1];
    the harmonics API contract;
  * the divergence guard (upstream tests/test_divergence_guard.py:38-71): a 1e38 drive
    overflows float32 within the first period and the solve raises
    SolverDivergedError naming the settling period, while the same setup at 1e4 returns
    finite fields.
The analytic references (O'Neil, Rayleigh, Fubini, shock distance) are the oracle's
(oracle/analytic.py), pinned by their limits in tests/test_oracle_physics.py.
"""

import numpy as np
import pytest

from oracle.analytic import fubini, oneill_axial, rayleigh_pressure, shock_distance
from oracle.sources import spherical_cap_points

pytestmark = pytest.mark.gpu

F0 = 1.0e6
C0 = 1500.0


# ---------------------------------------------------------------- 1-D plane wave
PPW = 4.0
DX = C0 / (F0 * PPW)
K0 = 2.0 * np.pi * F0 / C0
SPAN = slice(80, 200)


def _run_1d(alpha=0.0, n=256, pml_vox=24, src_vox=40, min_settle=40, max_settle=90, tol=0.005,
            record=2):
    import caustica_amd.solvers as solvers
    from caustica_amd import Grid, Medium, PMLSpec, plane_cw_source, water
    from caustica_amd.solvers import CWRunSpec

    grid = Grid((n,), DX, PMLSpec(thickness=pml_vox * DX))
    medium = Medium.homogeneous(grid.shape, water(alpha_np_m=alpha, c=C0))
    source = plane_cw_source(grid, f0=F0, amplitude=1.0e5, position_vox=src_vox)
    spec = CWRunSpec(min_settle_periods=min_settle, max_settle_periods=max_settle,
                     convergence_tol=tol, n_record_periods=record)
    return solvers.get("linear")().run(grid, medium, source, spec, backend="hip")


@pytest.fixture(scope="module")
def lossless():
    print("[gates] 1-D lossless plane wave", flush=True)
    return _run_1d()


def test_exact_period_dt_policy(lossless):
    assert lossless.spp == 8
    assert lossless.dt == pytest.approx(1.0 / F0 / 8, rel=1e-12)
    assert lossless.dt * C0 / DX == pytest.approx(0.5, rel=1e-9)


def test_plane_wave_dispersion_below_0p1_percent(lossless):
    ph = lossless.phasor[SPAN]
    k_num = np.angle(ph[1:] * np.conj(ph[:-1])).mean() / DX
    assert abs(k_num - K0) / K0 < 1e-3


def test_phasor_consistent_with_time_peak(lossless):
    ratio = lossless.amp[SPAN] / lossless.p_max[SPAN]
    assert ratio.max() <= (1.0 / np.cos(np.pi / lossless.spp)) * (1 + 1e-3)
    assert ratio.min() > 0.90


def test_absorption_matches_configured_alpha_within_1_percent():
    print("[gates] 1-D absorbing plane wave", flush=True)
    alpha = 30.0
    res = _run_1d(alpha=alpha, pml_vox=32, n=288, min_settle=45, max_settle=95)
    x = np.arange(*SPAN.indices(288)) * DX
    slope = np.polyfit(x, np.log(res.amp[SPAN]), 1)[0]
    assert abs(-slope - alpha) / alpha < 0.01


def test_pml_reflection_below_3_percent(lossless):
    amp = lossless.amp[SPAN]
    assert (amp.max() - amp.min()) / (amp.max() + amp.min()) < 0.03


def test_long_run_stability_no_energy_growth():
    print("[gates] 1-D 200-period stability", flush=True)
    res = _run_1d(min_settle=200, max_settle=200, tol=0.0)
    hist = np.array([pk for (_p, pk, _r) in res.convergence_history])
    periods = np.array([p for (p, _pk, _r) in res.convergence_history])
    late = hist[periods >= periods.max() - 100]
    assert res.settle_capped
    assert (late.max() - late.min()) / late.mean() < 0.01


def test_convergence_metadata_is_honest(lossless):
    assert lossless.tof_periods >= 50
    assert lossless.converged_period >= lossless.tof_periods
    assert not lossless.settle_capped
    assert lossless.steps_total == (lossless.converged_period + 2) * lossless.spp
    assert lossless.meta["backend"] == "hip"


# ---------------------------------------------------------------- 3-D bowl vs O'Neil
A = 7.5e-3
ROC = 18e-3
APEX = (40, 40, 16)
SHAPE = (80, 80, 120)
FOCUS_VOX = (40, 40, 16 + 48)
BOWL_DX = 0.375e-3


def minus6db_width(coord, prof):
    """Width of the half-amplitude main lobe around the peak (linear crossings)."""
    prof = np.asarray(prof, np.float64)
    i_pk = int(np.argmax(prof))
    half = prof[i_pk] / 2.0
    lo = i_pk
    while lo > 0 and prof[lo] > half:
        lo -= 1
    hi = i_pk
    while hi < prof.size - 1 and prof[hi] > half:
        hi += 1
    assert 0 < lo and hi < prof.size - 1, "-6 dB lobe not inside the profile"

    def cross(i0, i1):
        return coord[i0] + (half - prof[i0]) / (prof[i1] - prof[i0]) * (coord[i1] - coord[i0])

    return cross(hi - 1, hi) - cross(lo + 1, lo)


@pytest.fixture(scope="module")
def bowl():
    import caustica_amd.solvers as solvers
    from caustica_amd import Grid, Medium, PMLSpec, water
    from caustica_amd.solvers import CWRunSpec
    from caustica_amd.sources import bowl_cw_source

    print("[gates] 3-D bowl 80 x 80 x 120 against O'Neil / Rayleigh", flush=True)
    grid = Grid(SHAPE, BOWL_DX, PMLSpec(thickness=5e-3))
    medium = Medium.homogeneous(grid.shape, water(c=C0))
    source = bowl_cw_source(grid, f0=F0, amplitude=1.0e5, aperture_radius=A, roc=ROC, apex_vox=APEX)
    spec = CWRunSpec(min_settle_periods=8, max_settle_periods=30, n_record_periods=2)
    return solvers.get("linear")().run(grid, medium, source, spec, backend="hip",
                                       reference_point=FOCUS_VOX)


def test_focus_lands_within_one_voxel(bowl):
    amp = bowl.amp
    peak = np.unravel_index(np.argmax(amp), amp.shape)
    z_phys = (np.arange(SHAPE[2]) - APEX[2]) * BOWL_DX
    p_on = np.abs(oneill_axial(z_phys[24:100], A, ROC, F0, C0))
    z_true_vox = APEX[2] + z_phys[24:100][np.argmax(p_on)] / BOWL_DX
    assert abs(peak[0] - FOCUS_VOX[0]) <= 1
    assert abs(peak[1] - FOCUS_VOX[1]) <= 1
    assert abs(peak[2] - z_true_vox) <= 1.0


def test_axial_profile_matches_oneill(bowl):
    axial = bowl.amp[APEX[0], APEX[1], :]
    z_phys = (np.arange(SHAPE[2]) - APEX[2]) * BOWL_DX
    z_pml_limit = (SHAPE[2] - 13 - 2 - APEX[2]) * BOWL_DX
    sel = (z_phys > 0.5 * ROC) & (z_phys < min(1.5 * ROC, z_pml_limit))
    a = axial[sel] / axial[sel].max()
    b = np.abs(oneill_axial(z_phys[sel], A, ROC, F0, C0))
    b = b / b.max()
    assert np.corrcoef(a, b)[0, 1] > 0.99
    sel_w = (z_phys > 0.5 * ROC) & (z_phys < z_pml_limit)
    aw = axial[sel_w] / axial[sel_w].max()
    bw = np.abs(oneill_axial(z_phys[sel_w], A, ROC, F0, C0))
    w_solver = minus6db_width(z_phys[sel_w], aw)
    w_oneill = minus6db_width(z_phys[sel_w], bw / bw.max())
    assert abs(w_solver - w_oneill) / w_oneill < 0.05


def test_lateral_profile_matches_rayleigh(bowl):
    amp = bowl.amp
    z_pk = int(np.argmax(amp[APEX[0], APEX[1], :]))
    lateral = amp[:, APEX[1], z_pk]
    assert abs(int(np.argmax(lateral)) - APEX[0]) <= 1
    x_phys = (np.arange(SHAPE[0]) - APEX[0]) * BOWL_DX
    z_phys = (z_pk - APEX[2]) * BOWL_DX
    points, areas = spherical_cap_points(A, ROC, C0 / F0 / 5.0)
    targets = np.column_stack([x_phys, np.zeros_like(x_phys), np.full_like(x_phys, z_phys)])
    p_ray = np.abs(rayleigh_pressure(points, areas, 1.0, targets, k=2 * np.pi * F0 / C0, c=C0))
    sel = np.abs(x_phys) < 6e-3
    a = lateral[sel] / lateral[sel].max()
    b = p_ray[sel] / p_ray[sel].max()
    w_solver = minus6db_width(x_phys[sel], a)
    w_ray = minus6db_width(x_phys[sel], b)
    assert abs(w_solver - w_ray) / w_ray < 0.05
    assert np.corrcoef(a, b)[0, 1] > 0.98


def test_bowl_run_metadata(bowl):
    assert bowl.meta["padded_shape"] == SHAPE
    assert bowl.tof_periods in (12, 13)
    assert not bowl.settle_capped
    assert bowl.meta["backend"] == "hip"


# ---------------------------------------------------------------- Westervelt vs Fubini
BETA = 3.5


@pytest.fixture(scope="module")
def fubini_run():
    import caustica_amd.solvers as solvers
    from caustica_amd import Grid, Medium, PMLSpec, plane_cw_source, water
    from caustica_amd.solvers import CWRunSpec

    print("[gates] 1-D Westervelt, 600 voxels at 16 ppw, 2 MPa", flush=True)
    dx = C0 / (F0 * 16.0)
    grid = Grid((600,), dx, PMLSpec(thickness=40 * dx))
    med = Medium.homogeneous(grid.shape, water(c=C0, beta=BETA))
    src = plane_cw_source(grid, f0=F0, amplitude=2.0e6, position_vox=60)
    spec = CWRunSpec(min_settle_periods=45, max_settle_periods=100, convergence_tol=0.003)
    res = solvers.get("westervelt")().run(grid, med, src, spec, backend="hip", harmonics=(1, 2, 3))
    return res, dx, 60


def _sigma_axis(res, dx, src_vox):
    a1 = res.harmonic_amp(1)
    i0 = 90
    sig0 = ((i0 - src_vox) * dx) / shock_distance(a1[i0], F0, C0, beta=BETA)
    return a1, shock_distance(a1[i0] / fubini(1, sig0), F0, C0, beta=BETA)


def test_fubini_second_harmonic_within_5_percent(fubini_run):
    res, dx, src_vox = fubini_run
    a1, x_sh = _sigma_axis(res, dx, src_vox)
    a2 = res.harmonic_amp(2)
    assert res.meta["nonlinear_active"] is True
    checked = 0
    for iv in (150, 250, 350, 450, 540):
        sigma = ((iv - src_vox) * dx) / x_sh
        if not 0.05 <= sigma <= 0.95:
            continue
        analytic = fubini(2, sigma) / fubini(1, sigma)
        assert abs(a2[iv] / a1[iv] - analytic) / analytic < 0.05, sigma
        checked += 1
    assert checked >= 4


def test_fubini_third_harmonic_reasonable(fubini_run):
    res, dx, src_vox = fubini_run
    a1, x_sh = _sigma_axis(res, dx, src_vox)
    a3 = res.harmonic_amp(3)
    for iv in (350, 450, 540):
        sigma = ((iv - src_vox) * dx) / x_sh
        analytic = fubini(3, sigma) / fubini(1, sigma)
        assert abs(a3[iv] / a1[iv] - analytic) / analytic < 0.10, sigma


def test_amp_pmax_ceiling_invariant_nonlinear(fubini_run):
    res, _dx, _src = fubini_run
    band = res.p_max > 0.10 * res.p_max.max()
    ratio = res.amp[band] / res.p_max[band]
    assert ratio.max() <= (1.0 / np.cos(np.pi / res.spp)) * (1 + 1e-3)
    assert 0.85 <= res.amp[90] / res.p_max[90] <= 1.0 + 1e-3


def test_harmonics_api_contract():
    import caustica_amd.solvers as solvers
    from caustica_amd import Grid, Medium, PMLSpec, plane_cw_source, water
    from caustica_amd.solvers import CWRunSpec

    dx = C0 / (F0 * 4.0)
    grid = Grid((128,), dx, PMLSpec(thickness=16 * dx))
    med = Medium.homogeneous(grid.shape, water(c=C0))
    src = plane_cw_source(grid, f0=F0, amplitude=1e5, position_vox=24)
    spec = CWRunSpec(min_settle_periods=10, max_settle_periods=30)
    res = solvers.get("westervelt")().run(grid, med, src, spec, backend="hip", harmonics=(1, 2))
    assert set(res.phasors) == {1, 2}
    np.testing.assert_array_equal(res.phasors[1], res.phasor)
    with pytest.raises(KeyError, match="harmonic 3"):
        res.harmonic_amp(3)
    with pytest.raises(ValueError, match="fundamental"):
        solvers.get("westervelt")().run(grid, med, src, spec, backend="hip", harmonics=(2,))


# ---------------------------------------------------------------- divergence guard
def _tiny(amplitude):
    from caustica_amd import Grid, Medium, water
    from caustica_amd.sources import CWSource

    grid = Grid((24, 24, 24), 0.5e-3, None)
    medium = Medium.homogeneous(grid.shape, water())
    src = CWSource(indices=np.array([[12, 12, 6]], np.int32), phases=np.zeros(1, np.float32),
                   amplitude=amplitude, f0=1e6)
    return grid, medium, src


def test_a_diverging_run_raises_instead_of_returning_nan():
    import caustica_amd.solvers as solvers
    from caustica_amd.solvers import CWRunSpec
    from caustica_amd.solvers.base import SolverDivergedError

    print("[gates] 1e38 drive must raise", flush=True)
    grid, medium, src = _tiny(1e38)
    with pytest.raises(SolverDivergedError) as exc:
        solvers.get("linear")().run(grid, medium, src,
                                    CWRunSpec(min_settle_periods=1, max_settle_periods=3),
                                    backend="hip")
    msg = str(exc.value)
    assert "diverged" in msg and "settling period" in msg and "CFL" in msg


def test_a_healthy_run_of_the_same_shape_still_returns():
    import caustica_amd.solvers as solvers
    from caustica_amd.solvers import CWRunSpec

    grid, medium, src = _tiny(1e4)
    res = solvers.get("linear")().run(grid, medium, src,
                                      CWRunSpec(min_settle_periods=1, max_settle_periods=3),
                                      backend="hip")
    assert np.isfinite(res.phasor).all() and np.isfinite(res.p_max).all()
    assert res.meta["backend"] == "hip"
