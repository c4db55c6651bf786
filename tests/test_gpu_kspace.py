"""Parity of the HIP CW k-space path against the CPU oracle, through the C ABI.

Tolerances (float32 device arithmetic against the float64 oracle):
  one step from a random state          max |err| <= 2e-5 of max |field|
  fundamental phasor, p_max (rel-L2)    <= 1e-4
  2nd / 3rd harmonic phasor (rel-L2)    <= 1e-3
and bit-identity where the scheme demands it (beta = 0 Westervelt == linear,
a solve alone == the same solve batched with others).
"""

import numpy as np
import pytest

from oracle.kspace import KSpaceOracle, RunSpec, rel_l2
from tests._cases import product_case

pytestmark = pytest.mark.gpu

CASES = ["cw_1d_linear", "cw_1d_westervelt", "cw_2d_hetero", "cw_3d_bowl", "cw_3d_westervelt",
         "cw_3d_padded"]


def _oracle(args, precision="fp64"):
    a = dict(args)
    a["spec"] = RunSpec(**a["spec"])
    return KSpaceOracle(precision=precision, **a)


@pytest.mark.parametrize("name", CASES)
def test_one_step_from_random_state(name):
    from caustica_amd.solvers.kspace.native import KSpaceSession
    from caustica_amd.solvers.kspace.operators import plan_member

    print(f"[one-step] {name}", flush=True)
    g, grid, medium, source, spec, kw, solver, args = product_case(name)
    plan = plan_member(grid, medium, source, spec, nonlinear=solver == "westervelt", **kw)
    rng = np.random.default_rng(7)
    nd = grid.ndim
    orc = _oracle(args)
    shp = plan.padded
    p0 = rng.standard_normal(shp).astype(np.float32) * 1e4
    u0 = [rng.standard_normal(shp).astype(np.float32) * 1e-2 for _ in range(nd)]
    n0 = 5 * plan.spp + 3
    orc.p[...] = p0
    for i in range(nd):
        orc.u[i][...] = u0[i]
    orc.step(n0)
    dev_u = [np.zeros(plan.padded3, np.float32) for _ in range(3)]
    for i in range(nd):
        dev_u[3 - nd + i] = u0[i].reshape(plan.padded3)
    with KSpaceSession(plan.padded3, 1) as s:
        slot = s.add(plan)
        s.set_state(slot, n0, p0, *dev_u)
        s.advance([slot], 1)
        (p1, *u1), n1 = s.get_state(slot)
    assert n1 == n0 + 1
    ref_p = orc.p.reshape(plan.padded3)
    err = np.abs(p1 - ref_p).max() / np.abs(ref_p).max()
    assert err < 2e-5, f"p: {err:.3e}"
    for i in range(nd):
        ref = orc.u[i].reshape(plan.padded3)
        e = np.abs(u1[3 - nd + i] - ref).max() / np.abs(ref).max()
        assert e < 2e-5, f"u{i}: {e:.3e}"


@pytest.mark.parametrize("name", CASES)
def test_solve_matches_oracle_and_upstream(name):
    import caustica_amd.solvers as solvers

    print(f"[solve] {name}", flush=True)
    g, grid, medium, source, spec, kw, solver, args = product_case(name)
    res = solvers.get(solver)().run(grid, medium, source, spec, backend="hip", **kw)
    orc = _oracle(args).run()
    m = g["meta"]
    assert res.spp == m["spp"] and res.steps_total == m["steps_total"]
    assert res.converged_period == m["converged_period"]
    assert res.settle_capped == m["settle_capped"]
    assert res.meta["backend"] == "hip"
    assert rel_l2(res.phasor, orc.phasors[1]) < 1e-4
    assert rel_l2(res.p_max, orc.p_max) < 1e-4
    for h in m["harmonics"][1:]:
        assert rel_l2(res.phasors[h], orc.phasors[h]) < 1e-3, h
    # the upstream float32 run itself sits ~3e-6 from the fp64 oracle
    assert rel_l2(res.phasor, g["phasor1"]) < 1e-4
    # convergence history: same periods, peaks at float32 agreement
    hist_o = np.array([pk for _, pk, _ in orc.history])
    hist_d = np.array([pk for _, pk, _ in res.convergence_history])
    assert [p for p, _, _ in res.convergence_history] == [p for p, _, _ in orc.history]
    np.testing.assert_allclose(hist_d, hist_o, rtol=1e-4)


def test_beta_zero_westervelt_is_bitwise_linear():
    import caustica_amd.solvers as solvers
    from caustica_amd import Grid, Medium, PMLSpec, plane_cw_source, water
    from caustica_amd.solvers import CWRunSpec

    dx = 1500.0 / (1e6 * 4.0)
    grid = Grid((128,), dx, PMLSpec(thickness=16 * dx))
    med = Medium.homogeneous(grid.shape, water(c=1500.0))
    src = plane_cw_source(grid, f0=1e6, amplitude=1e5, position_vox=24)
    spec = CWRunSpec(min_settle_periods=20, max_settle_periods=40)
    r_lin = solvers.get("linear")().run(grid, med, src, spec, backend="hip")
    r_wes = solvers.get("westervelt")().run(grid, med, src, spec, backend="hip")
    np.testing.assert_array_equal(r_lin.phasor, r_wes.phasor)
    np.testing.assert_array_equal(r_lin.p_max, r_wes.p_max)
    assert r_wes.meta["nonlinear_active"] is False


def test_batched_members_equal_alone():
    """Three solves in one batched launch sequence == each solve alone, bit for bit."""
    from caustica_amd.solvers.kspace.native import KSpaceSession
    from caustica_amd.solvers.kspace.operators import plan_member

    names = ["cw_3d_bowl", "cw_3d_bowl", "cw_3d_bowl"]
    plans = []
    for i, name in enumerate(names):
        g, grid, medium, source, spec, kw, solver, args = product_case(name)
        med = type(medium)(alpha=medium.alpha * (1 + i), rho=medium.rho, c=medium.c * (1 + 0.01 * i),
                           beta=medium.beta)
        plans.append(plan_member(grid, med, source, spec, **kw))
    steps = [100, 90, 90]
    alone = []
    for plan, st in zip(plans, steps):
        with KSpaceSession(plans[0].padded3, 1) as s:
            slot = s.add(plan)
            s.schedule(slot, st - 20)
            s.advance([slot], st)
            alone.append(s.result(slot, 0, 1.0))
    with KSpaceSession(plans[0].padded3, 3) as s:
        slots = [s.add(p) for p in plans]
        for sl, st in zip(slots, steps):
            s.schedule(sl, st - 20)
        # staggered: member 1 joins late, member 2 leaves early
        s.advance([slots[0], slots[2]], 40)
        s.advance(slots, 50)
        s.advance([slots[0], slots[1]], 10)
        s.advance([slots[1]], 30)
        mixed = [s.result(sl, 0, 1.0) for sl in slots]
    for (pa, ma), (pb, mb) in zip(alone, mixed):
        np.testing.assert_array_equal(pa, pb)
        np.testing.assert_array_equal(ma, mb)
