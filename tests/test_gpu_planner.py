"""Planner against the device (upstream on-device planner gates: VRAM within 10 %,
calibrated time within 25 %; here the hip inventory is exact).

* ``hip_memory(plan)`` equals the bytes the library allocates for that member (read
  back as the pooled block after the member is removed), for linear and Westervelt,
  one and three harmonics, a record region and several sources;
* a solve's step count lies between the estimate's expected and worst counts;
* after ``calibrate`` of a padded shape, the estimated time of N steps matches the
  measured time of N steps within 25 %.
"""

import time

import numpy as np
import pytest

import caustica_amd.solvers as solvers
from caustica_amd import planner, workloads
from caustica_amd.solvers.kspace.native import KSpaceSession
from caustica_amd.solvers.kspace.operators import plan_member

pytestmark = pytest.mark.gpu


def _req(seed, solver, shape=(48, 48, 64)):
    return workloads.serving_request(0, np.random.default_rng(seed), 0.0, shape=shape,
                                     solver=solver)


@pytest.mark.parametrize("solver,harmonics,region", [
    ("linear", (1,), None),
    ("westervelt", (1, 2, 3), None),
    ("linear", (1,), (slice(4, 40), slice(8, 20), slice(0, 64))),
])
def test_hip_inventory_is_the_allocation(solver, harmonics, region):
    req = _req(11, solver)
    plan = plan_member(req.grid, req.medium, req.source, req.spec, harmonics=harmonics,
                       record_region=region, nonlinear=solver == "westervelt")
    want = planner.hip_memory(plan)
    print(f"[planner] {solver} h={harmonics} region={region}: {want.total_bytes} B", flush=True)
    with KSpaceSession(plan.padded3, 1) as s:
        slot = s.add(plan)
        s.remove(slot)
        assert s.pooled_bytes() == want.total_bytes


@pytest.mark.parametrize("solver,alpha_zero", [("linear", True), ("westervelt", False)])
def test_hip_inventory_of_a_staggered_member_is_the_allocation(solver, alpha_zero):
    """The same for the opt-in staggered scheme (split density, its own maps and tables)."""
    from caustica_amd.solvers.kspace.staggered import plan_staggered

    req = _req(14, solver)
    medium = req.medium
    if alpha_zero:
        medium = type(medium)(alpha=np.zeros_like(medium.alpha), rho=medium.rho, c=medium.c,
                              beta=medium.beta)
    plan = plan_staggered(req.grid, medium, req.source, req.spec, harmonics=req.harmonics,
                          nonlinear=solver == "westervelt", alpha_power=1.2)
    assert plan.stg["absorbing"] == (not alpha_zero)
    want = planner.hip_memory(plan)
    print(f"[planner] staggered {solver}: {want.total_bytes} B", flush=True)
    with KSpaceSession(plan.padded3, 1, scheme="staggered") as s:
        slot = s.add(plan)
        s.remove(slot)
        assert s.pooled_bytes() == want.total_bytes
    assert abs(plan.device_bytes() - want.total_bytes) <= (1 << 20) + 256 * 32


def test_steps_bracketed_by_estimate():
    req = _req(12, "linear")
    est = planner.estimate(req.grid, req.medium, req.source, req.spec, solver="linear",
                           reference_point=req.reference_point)
    res = solvers.get("linear")().run(req.grid, req.medium, req.source, req.spec,
                                      backend="hip", reference_point=req.reference_point)
    assert est.spp == res.spp and est.dt == pytest.approx(res.dt, rel=1e-12)
    assert est.steps_expected <= res.steps_total <= est.steps_worst


def test_calibrated_time_within_25_percent(tmp_path):
    shape = (128, 128, 128)
    path = tmp_path / "cal.json"
    entry = planner.calibrate((shape,), steps=40, path=path)
    key = "128x128x128:lin"
    print(f"[planner] calibrated {entry[key]}", flush=True)
    assert entry[key]["bytes_per_s"] > 0
    req = _req(13, "linear", shape)
    est = planner.estimate(req.grid, req.medium, req.source, req.spec, solver="linear",
                           calibration_path=path, reference_point=req.reference_point)
    assert est.source == "calibrated" and est.padded_shape == shape
    plan = plan_member(req.grid, req.medium, req.source, req.spec)
    n = 200
    with KSpaceSession(plan.padded3, 1) as s:
        slot = s.add(plan)
        s.advance([slot], 5)
        s.sync()
        t0 = time.perf_counter()
        s.advance([slot], n)
        s.sync()
        t = time.perf_counter() - t0
    assert abs(est.t_step_s * n - t) <= 0.25 * t, (est.t_step_s * n, t)
