"""Replayed HIP graphs of the step sequence, and the record-window rules of cst_member_result.

The step scalars come from each member's device clock, so every step of a launch
sequence takes the same arguments and a captured graph replays it. A solve advanced
through graphs (16-step graphs plus 1-step graphs for the remainder, several replays
with host synchronizations between them) must be bitwise the solve advanced by direct
launches: the same kernels read the same values. Batched member lists that change
between calls (members joining and leaving) each get their own graph and must match too.
"""

import numpy as np
import pytest

from caustica_amd import _native as N
from caustica_amd.solvers.kspace.native import KSpaceSession
from caustica_amd.solvers.kspace.operators import plan_member
from tests._cases import product_case

pytestmark = pytest.mark.gpu


def _plans(name, n):
    plans = []
    for i in range(n):
        g, grid, medium, source, spec, kw, solver, args = product_case(name)
        med = type(medium)(alpha=medium.alpha * (1 + i), rho=medium.rho,
                           c=medium.c * (1 + 0.01 * i), beta=medium.beta)
        plans.append(plan_member(grid, med, source, spec, nonlinear=solver == "westervelt", **kw))
    return plans


def _run(plans, graphs: bool, chunks):
    """Advance the members through ``chunks`` [(member indices, steps), ...]; return each
    member's (phasors, p_max, period peaks) and the graph statistics."""
    with KSpaceSession(plans[0].padded3, len(plans)) as s:
        s.set_graphs(graphs)
        slots = [s.add(p, max_periods=40) for p in plans]
        total = sum(k for who, k in chunks if 0 in who)
        for sl in slots:
            s.schedule(sl, total - 2 * plans[0].spp)
        for who, k in chunks:
            s.advance([slots[i] for i in who], k)
            s.sync()
        out = []
        for sl, p in zip(slots, plans):
            ph = [s.result(sl, hi, 1.0, with_pmax=False)[0] for hi in range(len(p.harmonics))]
            out.append((ph, s.result(sl, 0, 1.0)[1], s.peaks(sl, 0, 8)))
        return out, s.graph_stats()


@pytest.mark.parametrize("name", ["cw_3d_bowl", "cw_3d_westervelt"])
def test_graph_replay_is_bitwise_the_direct_launches(name):
    print(f"[graphs] {name}", flush=True)
    plans = _plans(name, 1)
    chunks = [([0], 37), ([0], 16), ([0], 1), ([0], 50)]
    direct, st0 = _run(plans, False, chunks)
    graph, st1 = _run(plans, True, chunks)
    assert st0["launches"] == 0 and st1["enabled"] and st1["launches"] > 0, (st0, st1)
    for (pa, ma, ka), (pb, mb, kb) in zip(direct, graph):
        for a, b in zip(pa, pb):
            np.testing.assert_array_equal(a, b)
        np.testing.assert_array_equal(ma, mb)
        np.testing.assert_array_equal(ka, kb)
    assert np.abs(direct[0][0][0]).max() > 0


def test_graph_replay_with_changing_member_lists():
    print("[graphs] staggered batch", flush=True)
    plans = _plans("cw_3d_bowl", 3)
    chunks = [([0, 2], 40), ([0, 1, 2], 50), ([0, 1], 10), ([1], 30), ([2], 21)]
    direct, _ = _run(plans, False, chunks)
    graph, st = _run(plans, True, chunks)
    assert st["captures"] >= 4, st
    for (pa, ma, ka), (pb, mb, kb) in zip(direct, graph):
        np.testing.assert_array_equal(pa[0], pb[0])
        np.testing.assert_array_equal(ma, mb)
        np.testing.assert_array_equal(ka, kb)


def test_result_is_final_until_reset():
    """The float64 reduction runs once: a second call returns the same phasor, another
    scale is refused, and the member cannot advance until it is reset."""
    plan = _plans("cw_3d_bowl", 1)[0]
    with KSpaceSession(plan.padded3, 1) as s:
        sl = s.add(plan, max_periods=40)
        s.schedule(sl, 60)
        s.advance([sl], 80)
        a = s.result(sl, 0, 0.1)[0]
        b = s.result(sl, 0, 0.1)[0]
        np.testing.assert_array_equal(a, b)
        with pytest.raises(N.NativeError) as e1:
            s.result(sl, 0, 0.2)
        assert e1.value.code == N.CST_ERR_ARG
        with pytest.raises(N.NativeError) as e2:
            s.advance([sl], 1)
        assert e2.value.code == N.CST_ERR_STATE
        s.reset(sl)
        s.schedule(sl, 60)
        s.advance([sl], 80)
        np.testing.assert_array_equal(s.result(sl, 0, 0.1)[0], a)


def test_set_state_restarts_the_record_window():
    """Restoring a state clears the record sums: recording after set_state gives the same
    phasor as a fresh run from that state."""
    plan = _plans("cw_3d_bowl", 1)[0]
    with KSpaceSession(plan.padded3, 2) as s:
        a, b = s.add(plan, max_periods=40), s.add(plan, max_periods=40)
        s.advance([a], 50)
        raw, n = s.get_raw(a)
        s.schedule(a, 50)
        s.advance([a], 20)          # records 20 steps ...
        s.set_raw(a, n, raw)        # ... then goes back to step 50
        s.set_raw(b, n, raw)
        s.schedule(b, 50)
        s.advance([a, b], 30)
        np.testing.assert_array_equal(s.result(a, 0, 1.0)[0], s.result(b, 0, 1.0)[0])
