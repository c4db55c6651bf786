"""Size-independent properties of the hip path at the benchmark sizes.

The float64 oracle cannot run a 256^3 solve in seconds, so the full-size solves are
checked through properties the scheme fixes exactly:
  * linearity: every operation of the linear step is homogeneous, and doubling is exact
    in binary floating point, so the 256^3 heterogeneous solve driven at 2 A gives
    bit-for-bit twice the phasor, p_max and period peaks of the solve at A;
  * batching: two differently steered 256^3 solves advanced together give bitwise the
    same results as each alone;
  * the 256 x 256 x 384 Westervelt solve (harmonics 1-3) stays finite, its harmonics
    fall off at the focus (|P1| > |P2| > |P3| > 0), and with beta = 0 it is bitwise the
    linear solve of the same medium.
"""

import dataclasses

import numpy as np
import pytest

from caustica_amd import workloads
from caustica_amd.solvers.kspace.native import KSpaceSession
from caustica_amd.solvers.kspace.operators import plan_member

pytestmark = pytest.mark.gpu


def _plan(wl, **kw):
    return plan_member(wl.grid, wl.medium, wl.source, wl.spec, harmonics=wl.harmonics,
                       reference_point=wl.reference_point, **kw)


def _solve(sess, slots, plans, periods_total, n_record):
    """Advance members together through a pinned schedule; return (phasors, p_max, peaks)."""
    spp = plans[0].spp
    total, rec = periods_total * spp, n_record * spp
    for s in slots:
        sess.reset(s)
        sess.schedule(s, total - rec)
    sess.advance(slots, total)
    out = []
    for s, p in zip(slots, plans):
        ph = [sess.result(s, hi, 2.0 / rec, with_pmax=False)[0] for hi in range(len(p.harmonics))]
        pm = sess.result(s, 0, 2.0 / rec)[1]
        out.append((ph, pm, sess.peaks(s, 0, periods_total - n_record)))
    return out


@pytest.fixture(scope="module")
def cfg1():
    print("[fullsize] building the 256^3 workload", flush=True)
    return workloads.cfg1_hetero_linear()


def test_linearity_is_exact_at_256(cfg1):
    wl = cfg1
    plan = _plan(wl)
    plan2 = dataclasses.replace(plan, amplitude=2.0 * plan.amplitude)
    with KSpaceSession(plan.padded3, 2) as sess:
        a = sess.add(plan, max_periods=wl.periods_total + 2)
        b = sess.add(plan2, max_periods=wl.periods_total + 2)
        print("[fullsize] 256^3 at A and 2A", flush=True)
        (pa, ma, ka), (pb, mb, kb) = _solve(sess, [a, b], [plan, plan2], wl.periods_total,
                                            wl.spec.n_record_periods)
    assert np.isfinite(ma).all() and ma.max() > 0
    np.testing.assert_array_equal(pb[0], 2 * pa[0])
    np.testing.assert_array_equal(mb, 2 * ma)
    np.testing.assert_array_equal(kb, 2 * ka)


def test_batched_equals_alone_at_256(cfg1):
    wl_a = cfg1
    wl_b = workloads.cfg1_hetero_linear(steer_vox=(8, -6, 10))
    pa, pb = _plan(wl_a), _plan(wl_b)
    with KSpaceSession(pa.padded3, 2) as sess:
        sa = sess.add(pa, max_periods=wl_a.periods_total + 2)
        sb = sess.add(pb, max_periods=wl_b.periods_total + 2)
        print("[fullsize] two steered 256^3 solves together", flush=True)
        both = _solve(sess, [sa, sb], [pa, pb], wl_a.periods_total, wl_a.spec.n_record_periods)
        print("[fullsize] each alone", flush=True)
        alone_a = _solve(sess, [sa], [pa], wl_a.periods_total, wl_a.spec.n_record_periods)[0]
        alone_b = _solve(sess, [sb], [pb], wl_b.periods_total, wl_b.spec.n_record_periods)[0]
    for got, ref in ((both[0], alone_a), (both[1], alone_b)):
        np.testing.assert_array_equal(got[0][0], ref[0][0])
        np.testing.assert_array_equal(got[1], ref[1])
        np.testing.assert_array_equal(got[2], ref[2])
    assert not np.array_equal(both[0][0][0], both[1][0][0])


def test_westervelt_256x256x384():
    print("[fullsize] building the 256x256x384 Westervelt workload", flush=True)
    wl = workloads.cfg2_westervelt()
    plan_nl = _plan(wl, nonlinear=True)
    assert plan_nl.beta2_dt is not None
    zero_beta = dataclasses.replace(plan_nl, beta2_dt=np.zeros_like(plan_nl.beta2_dt))
    plan_lin = _plan(wl, nonlinear=False)
    with KSpaceSession(plan_nl.padded3, 3) as sess:
        s_nl = sess.add(plan_nl, max_periods=wl.periods_total + 2)
        s_z = sess.add(zero_beta, max_periods=wl.periods_total + 2)
        s_l = sess.add(plan_lin, max_periods=wl.periods_total + 2)
        print("[fullsize] Westervelt, beta = 0 and linear together", flush=True)
        nl, z, lin = _solve(sess, [s_nl, s_z, s_l], [plan_nl, zero_beta, plan_lin],
                            wl.periods_total, wl.spec.n_record_periods)
    for h in range(3):
        assert np.isfinite(nl[0][h]).all()
        np.testing.assert_array_equal(z[0][h], lin[0][h])
    np.testing.assert_array_equal(z[1], lin[1])
    shape = plan_nl.rec_shape
    f = tuple(int(v) for v in wl.reference_point)
    amp = [abs(nl[0][h].reshape(shape)[f]) for h in range(3)]
    assert amp[0] > amp[1] > amp[2] > 0, amp
