"""Device-resident exchange through the DLPack entry points.

Coefficient maps uploaded once to device tensors feed a member device to device; the
finalized phasor and p_max come back as device tensors; a checkpoint taken into device
tensors restores another member without touching the host. Each path must be bitwise
the host path (plain numpy arrays through the same entry points).
"""

import ctypes as C

import numpy as np
import pytest

from caustica_amd import _native as N
from caustica_amd import dlpack as D
from caustica_amd.solvers.kspace.native import KSpaceSession
from caustica_amd.solvers.kspace.operators import plan_member
from tests._cases import product_case

pytestmark = pytest.mark.gpu


def _plan(name="cw_3d_westervelt"):
    g, grid, medium, source, spec, kw, solver, args = product_case(name)
    return plan_member(grid, medium, source, spec, nonlinear=solver == "westervelt", **kw)


def test_device_maps_results_and_state_equal_the_host_path():
    plan = _plan()
    maps = [plan.dt_over_rho, plan.rhoc2_dt, plan.absorb, plan.beta2_dt]
    dev_maps = [D.to_device(np.ascontiguousarray(m, np.float32)) for m in maps]
    assert all(t.on_device for t in dev_maps)
    with KSpaceSession(plan.padded3, 4) as s:
        host = s.add(plan, max_periods=40)
        dev = s.add(plan, max_periods=40, maps=dev_maps)
        for sl in (host, dev):
            s.schedule(sl, 90)
        s.advance([host, dev], 60)
        # a checkpoint of `dev` into device tensors, restored into a third member, and
        # the same through host arrays from `host` into a fourth (the real-space state
        # round trip rounds the velocity alike on both paths)
        st = [D.empty(plan.padded3, np.float32, device=0) for _ in range(4)]
        n = s.get_state_into(dev, *st)
        third = s.add(plan, max_periods=40, maps=dev_maps)
        s.set_state(third, n, *st)
        (p, *u), n4 = s.get_state(host)
        fourth = s.add(plan, max_periods=40)
        s.set_state(fourth, n4, p, *u)
        for sl in (third, fourth):
            s.schedule(sl, 90)
        s.advance([host, dev, third, fourth], 60)
        ph_h, pm_h = s.result(host, 0, 0.05)
        e_ph, e_pm = s.result_export(dev, 0, 0.05, device=True)
        assert e_ph.on_device and e_ph.shape == plan.rec_shape
        t_ph, _ = s.result_export(third, 0, 0.05, device=False, with_pmax=False)
        host_view = np.from_dlpack(t_ph)  # a host export numpy can consume
        ph_4 = s.result(fourth, 0, 0.05)[0]
        for h in range(1, len(plan.harmonics)):
            a = s.result(host, h, 0.05)[0]
            into = D.empty((a.size,), np.complex64, device=0)
            s.result_into(dev, h, 0.05, into, None)
            np.testing.assert_array_equal(into.to_numpy(), a)
    np.testing.assert_array_equal(e_ph.to_numpy().ravel(), ph_h)
    np.testing.assert_array_equal(e_pm.to_numpy().ravel(), pm_h)
    np.testing.assert_array_equal(host_view.ravel(), ph_4)
    assert np.abs(ph_h).max() > 0
    assert np.abs(ph_4 - ph_h).max() <= 1e-4 * np.abs(ph_h).max()


def test_tensors_on_another_gpu_or_wrong_dtype_are_refused():
    plan = _plan("cw_3d_bowl")
    with KSpaceSession(plan.padded3, 1) as s:
        sl = s.add(plan, max_periods=40)
        s.advance([sl], 3)
        R = int(np.prod(plan.rec_shape))
        host = np.zeros(R, np.complex64)
        with D.borrow(host) as t:
            t.contents.device.device_type = D.kDLROCM
            t.contents.device.device_id = 7  # not the session's GPU: refused before any copy
            assert N.lib().cst_member_result_into(s.handle, sl, 0, 1.0, t, None) == N.CST_ERR_ARG
        assert "GPU 7" in N.last_error()
        with pytest.raises(N.NativeError, match="complex64"):
            s.result_into(sl, 0, 1.0, np.zeros(R, np.float32), None)
        with pytest.raises(N.NativeError, match="elements"):
            s.result_into(sl, 0, 1.0, np.zeros(R + 1, np.complex64), None)
