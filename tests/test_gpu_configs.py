"""Parity of the hip path at the benchmark configurations, against the float64 oracle.

* 64 x 64 x 96 water bowl (520 steps): the float64 oracle runs here, on the whole grid.
* 256^3 heterogeneous linear solve (900 steps) and 256 x 256 x 384 Westervelt solve
  with harmonics 1-3 (2720 steps): against fixtures the float64 oracle wrote once in
  the build container (tests/golden/make_fullsize.py); the test rebuilds the seeded
  medium, checks it against the fixture's sample first, solves on hip through the solver
  API and compares focal planes, a strided volume sample and a full-resolution box
  around the focus.
* the Pennes + CEM43 stage on the 256^3 grid (200 heating + 400 cooling steps) heated
  by the hip 256^3 phasor through the device join, against the float64 oracle's run
  heated by its own phasor (tests/golden/full_cfg3.npz).
* one step from a random state at 256^3 and 256 x 256 x 384 under every thread-layout
  variant of the z passes, against the float64 oracle's step.
* the Pennes two-step kernel on 256^3 (64-plane x chunks) against the float64 oracle.
* two requests of the mixed-size serving queue (steered bowl, fresh random-field seed,
  pinned schedule): Westervelt 128 x 128 x 256 with harmonics 1-2 and linear 192^3,
  against float64 fixtures of the same requests (tests/golden/make_fullsize.py).

Tolerances (the path's accuracy targets): fundamental phasor and focal peak rel <= 1e-4, harmonics 2-3
rel-L2 <= 1e-3, p_max rel-L2 <= 1e-4, temperature <= 1e-3 K, CEM43 <= 1e-3 relative,
one step <= 2e-5 of max |field|; schedules (spp, step counts, time of flight) equal and
period peaks within 1e-4.
"""

import os

import numpy as np
import pytest

from oracle.kspace import rel_l2
from tests import _full

pytestmark = pytest.mark.gpu

WORKERS = max(1, min(8, os.cpu_count() or 1))


def _solve_hip(wl):
    import caustica_amd.solvers as solvers

    return solvers.get(wl.solver)().run(wl.grid, wl.medium, wl.source, wl.spec, backend="hip",
                                        harmonics=wl.harmonics, record_region=wl.record_region,
                                        reference_point=wl.reference_point)


def _peak(field, focus):
    return _full.peak_info(field, focus)


def test_cfg0_water_bowl_vs_fp64_oracle():
    from caustica_amd import workloads

    wl = workloads.cfg0_water_bowl()
    print(f"[configs] {wl.name}: hip solve", flush=True)
    res = _solve_hip(wl)
    print(f"[configs] {wl.name}: float64 oracle ({WORKERS} workers)", flush=True)
    orc = _full.oracle_for(wl, "fp64", WORKERS).run()
    assert (res.spp, res.steps_total, res.tof_periods) == (orc.spp, orc.steps_total, orc.tof_periods)
    assert res.steps_total == 40 * res.spp
    assert res.converged_period == orc.converged_period and res.settle_capped == orc.settle_capped
    assert [p for p, _, _ in res.convergence_history] == [p for p, _, _ in orc.history]
    np.testing.assert_allclose([pk for _, pk, _ in res.convergence_history],
                               [pk for _, pk, _ in orc.history], rtol=1e-4)
    e_ph = rel_l2(res.phasor, orc.phasors[1])
    e_pm = rel_l2(res.p_max, orc.p_max)
    a, b = _peak(res.phasor, wl.reference_point), _peak(orc.phasors[1], wl.reference_point)
    print(f"[configs] cfg0 rel-L2 phasor {e_ph:.2e} p_max {e_pm:.2e}; focal peak "
          f"{a['peak']:.6e} vs {b['peak']:.6e}", flush=True)
    assert e_ph <= 1e-4 and e_pm <= 1e-4
    assert a["peak_at"] == b["peak_at"]
    assert abs(a["peak"] - b["peak"]) <= 1e-4 * b["peak"]
    assert abs(a["at_focus"] - b["at_focus"]) <= 1e-4 * b["peak"]


def _fixture(case):
    path = _full.GOLDEN / f"full_{case}.npz"
    if not path.exists():
        pytest.skip(f"{path.name} not generated (tests/golden/make_fullsize.py)")
    return _full.load(f"full_{case}")


def _check_medium(wl, fx):
    got = _full.medium_sample(wl.medium)
    for k in ("c_sample", "rho_sample"):
        np.testing.assert_allclose(got[k], fx[k], rtol=1e-6, err_msg=k)
    same = _full.medium_digest(wl.medium) == fx["meta"]["medium_sha256"]
    print(f"[configs] medium sample matches; digest {'equal' if same else 'differs (ulp-level)'}",
          flush=True)


def _check_schedule(res, fx):
    m = fx["meta"]
    assert (res.spp, res.steps_total, res.tof_periods) == (m["spp"], m["steps_total"], m["tof_periods"])
    assert res.converged_period == m["converged_period"] and res.settle_capped == m["settle_capped"]
    hist = np.array([(p, pk) for p, pk, _ in res.convergence_history], np.float64)
    np.testing.assert_array_equal(hist[:, 0], fx["history"][:, 0])
    np.testing.assert_allclose(hist[:, 1], fx["history"][:, 1], rtol=1e-4)


def _compare_samples(field, focus, tag, fx, tol):
    got = _full.field_samples(field, focus, tag)
    errs = {k: rel_l2(got[k], fx[k]) for k in got}
    print(f"[configs] {tag}: " + ", ".join(f"{k.split('_')[1]} {v:.2e}" for k, v in errs.items()),
          flush=True)
    for k, v in errs.items():
        assert v <= tol, f"{k}: rel-L2 {v:.3e} > {tol:g}"
    ref = fx["meta"]["peaks"][tag]
    pk = _peak(field, focus)
    assert pk["peak_at"] == ref["peak_at"], (pk, ref)
    assert abs(pk["peak"] - ref["peak"]) <= tol * ref["peak"], (pk, ref)
    assert abs(pk["at_focus"] - ref["at_focus"]) <= tol * ref["peak"], (pk, ref)


@pytest.fixture(scope="module")
def cfg1_solved():
    from caustica_amd import workloads

    fx = _fixture("cfg1")
    wl = workloads.cfg1_hetero_linear(seed=fx["meta"]["seed"])
    _check_medium(wl, fx)
    print(f"[configs] {wl.name}: hip solve", flush=True)
    return wl, fx, _solve_hip(wl)


def test_cfg1_256_vs_fp64_fixture(cfg1_solved):
    wl, fx, res = cfg1_solved
    _check_schedule(res, fx)
    _compare_samples(res.phasors[1], wl.reference_point, "ph1", fx, 1e-4)
    _compare_samples(res.p_max, wl.reference_point, "pmax", fx, 1e-4)


@pytest.mark.parametrize("case", ["serving_w128x128x256", "serving_l192"])
def test_serving_request_vs_fp64_fixture(case):
    from caustica_amd import workloads

    fx = _fixture(case)
    m = fx["meta"]
    wl = workloads.serving_workload(m["seed"], tuple(m["shape"]), m["solver"])
    _check_medium(wl, fx)
    print(f"[configs] {wl.name}: hip solve", flush=True)
    res = _solve_hip(wl)
    _check_schedule(res, fx)
    _compare_samples(res.phasors[1], wl.reference_point, "ph1", fx, 1e-4)
    _compare_samples(res.p_max, wl.reference_point, "pmax", fx, 1e-4)
    for h in wl.harmonics[1:]:
        _compare_samples(res.phasors[h], wl.reference_point, f"ph{h}", fx, 1e-3)


def test_cfg2_westervelt_vs_fp64_fixture():
    from caustica_amd import workloads

    fx = _fixture("cfg2")
    wl = workloads.cfg2_westervelt(seed=fx["meta"]["seed"])
    _check_medium(wl, fx)
    print(f"[configs] {wl.name}: hip solve", flush=True)
    res = _solve_hip(wl)
    assert res.meta["nonlinear_active"] is True
    _check_schedule(res, fx)
    _compare_samples(res.phasors[1], wl.reference_point, "ph1", fx, 1e-4)
    _compare_samples(res.p_max, wl.reference_point, "pmax", fx, 1e-4)
    for h in (2, 3):
        _compare_samples(res.phasors[h], wl.reference_point, f"ph{h}", fx, 1e-3)


def test_cfg3_pennes_from_the_256_phasor_vs_fp64_fixture():
    """The thermal benchmark end to end on hip: the 256^3 solve, Q formed on the device
    from the resident phasor and scaled to 150 W/cm^2 at the focus, 200 + 400 Pennes
    steps with CEM43, against the float64 chain."""
    from caustica_amd import workloads
    from caustica_amd.sensors import DeviceHeatingSource
    from caustica_amd.serving.batcher import schedule_of
    from caustica_amd.solvers.kspace.native import KSpaceSession
    from caustica_amd.solvers.kspace.operators import plan_member
    from caustica_amd.thermal import PennesSolver

    fx = _fixture("cfg3")
    m = fx["meta"]
    wl = workloads.cfg1_hetero_linear()
    plan = plan_member(wl.grid, wl.medium, wl.source, wl.spec, harmonics=wl.harmonics,
                       reference_point=wl.reference_point)
    total, rec = schedule_of(plan, wl.spec)
    focus = tuple(m["focus"])
    with KSpaceSession(plan.padded3, 1) as sess:
        slot = sess.add(plan, max_periods=plan.eff_max + 4)
        sess.schedule(slot, total - rec)
        print("[configs] cfg3: 256^3 acoustic solve on hip", flush=True)
        sess.advance([slot], total)
        ph = sess.result(slot, 0, 2.0 / rec, with_pmax=False)[0].reshape(plan.rec_shape)
        scale = workloads.focal_intensity_scale(ph[focus], float(wl.medium.rho[focus]),
                                                float(wl.medium.c[focus]))
        assert abs(scale - m["q_scale"]) <= 2e-4 * m["q_scale"], (scale, m["q_scale"])
        src = DeviceHeatingSource.from_member(sess, slot, plan, wl.medium, wl.grid.dx,
                                              rec_steps=rec).scaled(scale)
        med = workloads.pennes_tissue(plan.rec_shape, wl.grid.dx)
        ps = PennesSolver("hip")
        t0 = np.full(plan.rec_shape, 37.0, np.float32)
        print("[configs] cfg3: Pennes 200 + 400 steps on hip", flush=True)
        hot = ps.solve(t0, src, med, dt=m["dt"], n_steps=m["n_heat"], dose=True)
        cool = ps.solve(hot.temperature, None, med, dt=m["dt"], n_steps=m["n_cool"], dose=True,
                        dose0=hot.dose_cem43)
    tmax = np.maximum(hot.temperature_max, cool.temperature_max)
    for tag, f, kind in (("temp", cool.temperature, "abs"), ("tmax", tmax, "abs"),
                         ("dose", cool.dose_cem43, "rel")):
        got = _full.field_samples(f, focus, tag)
        for k, v in got.items():
            err = float(np.abs(v.astype(np.float64) - fx[k]).max())
            bar = 1e-3 if kind == "abs" else 1e-3 * m["peaks"]["dose"]["peak"]
            print(f"[configs] cfg3 {k}: max |err| {err:.3e} (bar {bar:.3e})", flush=True)
            assert err <= bar, k
    assert m["peaks"]["tmax"]["peak"] > 43.0  # the benchmark reaches ablative temperatures


def _one_step_fullsize(wl, variants):
    from caustica_amd.solvers.kspace.native import KSpaceSession
    from caustica_amd.solvers.kspace.operators import plan_member

    plan = plan_member(wl.grid, wl.medium, wl.source, wl.spec, harmonics=wl.harmonics,
                       reference_point=wl.reference_point, nonlinear=wl.solver == "westervelt")
    shp = plan.padded3
    rng = np.random.default_rng(11)
    # a smooth-ish state of acoustic magnitude: the scheme's step at every bin
    p0 = (rng.standard_normal(shp) * 1e5).astype(np.float32)
    u0 = [(rng.standard_normal(shp) * 0.05).astype(np.float32) for _ in range(3)]
    n0 = 7 * plan.spp + 5
    orc = _full.oracle_for(wl, "fp64", WORKERS)
    orc.p[...] = p0
    for i in range(3):
        orc.u[i][...] = u0[i]
    print(f"[configs] one step {shp}: float64 oracle", flush=True)
    orc.step(n0)
    ref_p = orc.p
    ref_u = [u.copy() for u in orc.u]
    del orc
    with KSpaceSession(shp, 1) as s:
        slot = s.add(plan)
        for zu, zp in variants:
            s.set_variant(2, zu)
            s.set_variant(3, zp)
            s.set_state(slot, n0, p0, *u0)
            s.advance([slot], 1)
            (p1, *u1), n1 = s.get_state(slot)
            assert n1 == n0 + 1
            err = [float(np.abs(p1 - ref_p).max() / np.abs(ref_p).max())]
            err += [float(np.abs(u1[i] - ref_u[i]).max() / np.abs(ref_u[i]).max()) for i in range(3)]
            print(f"[configs] one step {shp} zu {zu} zp {zp}: p {err[0]:.2e}, u "
                  + ", ".join(f"{e:.2e}" for e in err[1:]), flush=True)
            assert max(err) <= 2e-5, (zu, zp, err)


def test_one_step_256_all_variants():
    from caustica_amd import workloads

    wl = workloads.cfg1_hetero_linear()
    _one_step_fullsize(wl, [(zu, zp) for zu in (0, 1, 2) for zp in (0, 1)])


def test_one_step_256x256x384_westervelt_all_variants():
    from caustica_amd import workloads

    wl = workloads.cfg2_westervelt()
    _one_step_fullsize(wl, [(zu, zp) for zu in (0, 1, 2) for zp in (0, 1)])


@pytest.mark.parametrize("boundary", ["insulated", "dirichlet"])
def test_pennes_256_cubed_two_step_vs_fp64(boundary):
    """256^3 runs the two-step kernel with 64-plane x chunks (512 blocks, one round on
    the chip); heterogeneous conductivity, perfusion and source so that every halo and
    chunk-border plane enters the result; 6 steps = 3 two-step launches."""
    from caustica_amd.thermal import PennesSolver, ThermalMedium
    from oracle.pennes import PennesOracle

    shape = (256, 256, 256)
    print(f"[configs] Pennes 256^3 {boundary}", flush=True)
    rng = np.random.default_rng(5)
    k = (0.3 + 0.4 * rng.random(shape, dtype=np.float32)).astype(np.float32)
    rho = np.full(shape, 1000.0, np.float32)
    cp = (3400.0 + 400.0 * rng.random(shape, dtype=np.float32)).astype(np.float32)
    perf = (5e-4 * rng.random(shape, dtype=np.float32)).astype(np.float32)
    q = (2e7 * rng.random(shape, dtype=np.float32) ** 6).astype(np.float32)
    t0 = (37.0 + 2.0 * rng.random(shape, dtype=np.float32)).astype(np.float32)
    dx, dt, n = 3.75e-4, 0.05, 6
    med = ThermalMedium(k=k, rho=rho, specific_heat=cp, perfusion=perf, dx=dx)
    dev = PennesSolver("hip", boundary=boundary, finite_check_every=n, profile=True).solve(
        t0, q, med, dt=dt, n_steps=n, dose=True)
    assert dev.meta["profile"]["launches"] == n // 2
    print("[configs] Pennes 256^3: float64 oracle", flush=True)
    ref = PennesOracle(k, rho, cp, perf, dx, boundary=boundary, precision="fp64").solve(
        t0, q, dt, n, dose=True)
    e_t = float(np.abs(dev.temperature - ref.temperature).max())
    e_m = float(np.abs(dev.temperature_max - ref.temperature_max).max())
    e_d = float(np.abs(dev.dose_cem43 - ref.dose).max() / np.abs(ref.dose).max())
    print(f"[configs] Pennes 256^3 {boundary}: T {e_t:.2e} K, T_max {e_m:.2e} K, dose {e_d:.2e}",
          flush=True)
    assert float((ref.temperature - t0).max()) > 0.5
    assert e_t < 1e-3 and e_m < 1e-3 and e_d <= 1e-3
