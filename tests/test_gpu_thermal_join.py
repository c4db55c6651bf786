"""Device-side acoustic -> thermal join (cst_thermal_heat_from_member).

The heat source formed on the device from the resident phasors must drive the Pennes
solve to the same temperatures and dose as the host path (phasors read back,
HeatingSource.from_phasors, Q uploaded), for a linear fundamental-only case, a
Westervelt case with harmonics 1-2 under a power-law absorption model, a scaled source
and a reference temperature away from arterial (the perfusion offset folded into Q).
The two paths differ only in the rounding of Q (float64 |P|^2 from the float32
components on the device, the float32 magnitude squared on the host: about 1e-7
relative), so temperatures agree to 1e-5 K and the dose to 1e-5 relative. The device Q
itself, read back, equals the oracle's float64 heat source (oracle/heating.py, pinned to
the upstream fixture heat_power_law.npz) on the read-back phasors to 1e-6 relative.
"""

import numpy as np
import pytest

from caustica_amd import workloads
from oracle.heating import heat_source
from caustica_amd.sensors import DeviceHeatingSource, HeatingSource
from caustica_amd.serving.batcher import schedule_of
from caustica_amd.solvers.kspace.native import KSpaceSession
from caustica_amd.solvers.kspace.operators import plan_member

pytestmark = pytest.mark.gpu


def _acoustic(solver, seed):
    rng = np.random.default_rng(seed)
    req = workloads.serving_request(0, rng, 0.0, shape=(48, 48, 64), solver=solver)
    plan = plan_member(req.grid, req.medium, req.source, req.spec, harmonics=req.harmonics,
                       reference_point=req.reference_point, nonlinear=solver == "westervelt")
    total, rec = schedule_of(plan, req.spec)
    sess = KSpaceSession(plan.padded3, 1)
    slot = sess.add(plan, max_periods=plan.eff_max + 4)
    sess.schedule(slot, total - rec)
    sess.advance([slot], total)
    return req, plan, sess, slot, rec


@pytest.mark.parametrize("solver,y,scale,t_ref", [("linear", None, 1.0, None),
                                                  ("westervelt", 1.1, 3.0, None),
                                                  ("linear", None, 1.0, 20.0)])
def test_device_join_matches_host_join(solver, y, scale, t_ref):
    from caustica_amd.thermal import PennesSolver

    print(f"[join] {solver} y={y} scale={scale} t_ref={t_ref}", flush=True)
    req, plan, sess, slot, rec = _acoustic(solver, seed=21)
    try:
        phasors = {h: sess.result(slot, hi, 2.0 / rec, with_pmax=False)[0].reshape(plan.rec_shape)
                   for hi, h in enumerate(plan.harmonics)}
        host_q = HeatingSource.from_phasors(phasors, alpha=req.medium.alpha, rho=req.medium.rho,
                                            c=req.medium.c, dx=req.grid.dx, alpha_power_law_y=y)
        dev_q = DeviceHeatingSource.from_member(sess, slot, plan, req.medium, req.grid.dx,
                                                rec_steps=rec, alpha_power_law_y=y)
        assert dev_q.harmonics == host_q.harmonics
        # x100: the small grid's focus sits in weakly absorbing water; the case must heat
        host_q, dev_q = host_q.scaled(100.0 * scale), dev_q.scaled(100.0 * scale)
        med = workloads.pennes_tissue(plan.rec_shape, req.grid.dx)
        kw = {} if t_ref is None else {"reference_temperature_c": t_ref}
        ps = PennesSolver("hip", **kw)
        t0 = np.full(plan.rec_shape, 37.0, np.float32)
        a = ps.solve(t0, host_q, med, dt=0.05, n_steps=40, dose=True)
        b = ps.solve(t0, dev_q, med, dt=0.05, n_steps=40, dose=True)
    finally:
        sess.close()
    rise = float(a.temperature.max() - 37.0)
    assert rise > 0.5, rise  # the case heats
    assert np.abs(a.temperature - b.temperature).max() <= 1e-5
    assert np.abs(a.temperature_max - b.temperature_max).max() <= 1e-5
    assert np.abs(a.dose_cem43 - b.dose_cem43).max() <= 1e-5 * max(float(a.dose_cem43.max()), 1e-12)
    assert b.meta["q"].startswith("DeviceHeatingSource")


def test_device_join_refuses_harmonics_without_model():
    req, plan, sess, slot, rec = _acoustic("westervelt", seed=22)
    try:
        with pytest.raises(ValueError, match="absorption|alpha"):
            DeviceHeatingSource.from_member(sess, slot, plan, req.medium, req.grid.dx, rec_steps=rec)
    finally:
        sess.close()


@pytest.mark.parametrize("solver,y,scale", [("linear", None, 1.0), ("westervelt", 1.1, 2.5)])
def test_device_join_q_equals_oracle_heat_source(solver, y, scale):
    print(f"[join] Q vs oracle: {solver} y={y} scale={scale}", flush=True)
    req, plan, sess, slot, rec = _acoustic(solver, seed=23)
    try:
        phasors = {h: sess.result(slot, hi, 2.0 / rec, with_pmax=False)[0].reshape(plan.rec_shape)
                   for hi, h in enumerate(plan.harmonics)}
        dev_q = DeviceHeatingSource.from_member(sess, slot, plan, req.medium, req.grid.dx,
                                                rec_steps=rec, alpha_power_law_y=y).scaled(scale)
        q = dev_q.to_host()
    finally:
        sess.close()
    ref = scale * heat_source(phasors, req.medium.alpha, req.medium.rho, req.medium.c,
                              alpha_power_law_y=y)
    assert q.dtype == np.float32 and q.shape == ref.shape
    assert float(ref.max()) > 0
    np.testing.assert_allclose(q, ref, rtol=1e-6, atol=1e-6 * float(ref.max()))
