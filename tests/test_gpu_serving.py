"""Continuous batching: every served result equals the same solve run alone.

A mixed queue of small grids (two padded shapes per group, linear and Westervelt with
harmonics 1-2, steered bowls, fresh random media) is served with staggered arrivals,
so members join and leave batches at arbitrary step boundaries and groups of different
shapes overlap on their streams. Each result must be bit-identical to the plain form
(one solve at a time through the solver API): a member's trajectory reads only its own
arrays and scalars. A tight HBM budget must hold admissions under the budget and still
serve every request. One served result is also checked against the float64 oracle at
the path's tolerance (fundamental phasor rel-L2 <= 1e-4).
"""

import numpy as np
import pytest

from caustica_amd import workloads
from caustica_amd.serving import ContinuousBatcher, serve_serial, summarize
from caustica_amd.solvers.kspace.operators import plan_member

pytestmark = pytest.mark.gpu

MIX = [((64, 64, 64), "linear"), ((48, 48, 64), "westervelt"), ((64, 64, 64), "westervelt"),
       ((48, 48, 64), "linear"), ((64, 64, 96), "linear"), ((64, 64, 64), "linear"),
       ((48, 48, 64), "linear")]


def _queue(spacing_s):
    rng = np.random.default_rng(11)
    return [workloads.serving_request(i, rng, spacing_s * i, shape=shape, solver=solver)
            for i, (shape, solver) in enumerate(MIX)]


def _by_rid(responses):
    out = {}
    for r in responses:
        assert r.error is None, f"request {r.rid}: {r.error}"
        out[r.rid] = r.result
    return out


def _assert_same(a, b, rid):
    assert a.steps_total == b.steps_total, rid
    assert sorted(a.phasors) == sorted(b.phasors), rid
    for h in a.phasors:
        np.testing.assert_array_equal(a.phasors[h], b.phasors[h], err_msg=f"request {rid} h={h}")
    np.testing.assert_array_equal(a.p_max, b.p_max, err_msg=f"request {rid} p_max")
    assert [x[1] for x in a.convergence_history] == [x[1] for x in b.convergence_history], rid


@pytest.fixture(scope="module")
def plain():
    reqs = _queue(0.0)
    print(f"[serving] plain form: {len(reqs)} solves", flush=True)
    return _by_rid(serve_serial(reqs))


def test_served_mix_equals_alone(plain):
    reqs = _queue(0.004)
    b = ContinuousBatcher(max_batch=4, chunk_steps=2, planners=2)
    try:
        print("[serving] continuous batcher, staggered arrivals", flush=True)
        resp = b.serve(reqs)
    finally:
        b.close()
    served = _by_rid(resp)
    assert sorted(served) == sorted(plain)
    for rid in plain:
        _assert_same(served[rid], plain[rid], rid)
    s = summarize(resp)
    assert s["errors"] == 0
    assert s["max_batch_seen"] >= 2, "the mix never shared a launch"


def test_admission_respects_hbm_budget(plain):
    reqs = _queue(0.0)
    plans = [plan_member(r.grid, r.medium, r.source, r.spec, harmonics=r.harmonics,
                         reference_point=r.reference_point, nonlinear=r.solver == "westervelt")
             for r in reqs]
    budget = 2 * max(p.device_bytes() for p in plans)
    b = ContinuousBatcher(hbm_budget=budget, max_batch=4, chunk_steps=4, planners=2)
    try:
        print(f"[serving] budget {budget} B", flush=True)
        resp = b.serve(reqs)
        assert 0 < b.peak_in_use <= budget
    finally:
        b.close()
    served = _by_rid(resp)
    for rid in plain:
        _assert_same(served[rid], plain[rid], rid)


def test_served_result_matches_fp64_oracle(plain):
    from oracle.kspace import KSpaceOracle, RunSpec, rel_l2

    req = _queue(0.0)[3]  # 48 x 48 x 64 linear
    res = plain[req.rid]
    orc = KSpaceOracle(
        shape=req.grid.shape, dx=req.grid.dx, pml_vox=req.grid.pml_vox, pml_edge=req.grid.pml.edge,
        alpha=req.medium.alpha, rho=req.medium.rho, c=req.medium.c, beta=req.medium.beta,
        src_idx=req.source.indices, src_phase=req.source.phases, amplitude=req.source.amplitude,
        f0=req.source.f0, ramp_periods=req.source.ramp_periods,
        spec=RunSpec(**req.spec.model_dump()), harmonics=req.harmonics,
        reference_point=req.reference_point, precision="fp64",
    )
    ref = orc.run()
    assert ref.steps_total == res.steps_total
    assert rel_l2(res.phasor, ref.phasor) <= 1e-4
    assert rel_l2(res.p_max, ref.p_max) <= 1e-4


def test_served_benchmark_shapes_equal_alone():
    """The serving benchmark's four grid sizes (128^3, 192^3, 256^3, 128 x 128 x 256 at
    dx 0.375 mm, steered bowls, fresh media, linear and Westervelt with harmonics 1-2),
    arriving close together so that groups of different shapes overlap on the device and
    a same-shape pair shares launches: every served result bitwise the plain form's."""
    rng = np.random.default_rng(29)
    mix = [((128, 128, 128), "linear"), ((192, 192, 192), "linear"), ((256, 256, 256), "linear"),
           ((128, 128, 256), "westervelt"), ((128, 128, 128), "westervelt"),
           ((128, 128, 128), "linear")]
    reqs = [workloads.serving_request(i, rng, 0.01 * i, shape=shape, solver=solver)
            for i, (shape, solver) in enumerate(mix)]
    print(f"[serving] benchmark shapes: plain form, {len(reqs)} solves", flush=True)
    plain = _by_rid(serve_serial(reqs))
    b = ContinuousBatcher(planners=3)
    try:
        print("[serving] benchmark shapes: continuous batcher", flush=True)
        resp = b.serve(reqs)
    finally:
        b.close()
    served = _by_rid(resp)
    for rid in plain:
        _assert_same(served[rid], plain[rid], rid)
    assert summarize(resp)["max_batch_seen"] >= 2
