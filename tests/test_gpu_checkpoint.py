"""Checkpoint / resume on the hip backend (upstream tests/test_checkpoint.py:61-83).

An interrupted solve resumed from its checkpoint must reproduce the uninterrupted
solve bit for bit (the device state p, u0..u2 is read back exactly and the spectrum of
p is rebuilt by the same transform the step uses), and a checkpoint written by a
different run must be refused.
"""

import dataclasses

import numpy as np
import pytest

import caustica_amd.solvers as solvers
from caustica_amd import workloads
from caustica_amd.io.checkpoint import CheckpointMismatch, CheckpointSpec, RunInterrupted

pytestmark = pytest.mark.gpu


def _request(seed=5, solver="linear"):
    rng = np.random.default_rng(seed)
    return workloads.serving_request(0, rng, 0.0, shape=(48, 48, 64), solver=solver)


def _run(req, **kw):
    return solvers.get(req.solver)().run(req.grid, req.medium, req.source, req.spec, backend="hip",
                                         harmonics=req.harmonics,
                                         reference_point=req.reference_point, **kw)


@pytest.mark.parametrize("solver", ["linear", "westervelt"])
def test_resume_is_bitwise(tmp_path, solver):
    req = _request(solver=solver)
    print(f"[checkpoint] {solver}: uninterrupted", flush=True)
    ref = _run(req)
    calls = {"n": 0}

    def stop_after_three():
        calls["n"] += 1
        return calls["n"] >= 3

    ck = tmp_path / "run.ckpt.npz"
    print(f"[checkpoint] {solver}: interrupted", flush=True)
    with pytest.raises(RunInterrupted) as ei:
        _run(req, checkpoint=CheckpointSpec(ck, every_periods=2, stop_when=stop_after_three))
    assert ck.exists() and ei.value.periods_done == 3
    print(f"[checkpoint] {solver}: resumed", flush=True)
    res = _run(req, checkpoint=CheckpointSpec(ck, every_periods=2))
    assert res.steps_total == ref.steps_total
    for h in ref.phasors:
        np.testing.assert_array_equal(res.phasors[h], ref.phasors[h])
    np.testing.assert_array_equal(res.p_max, ref.p_max)
    assert [x[1] for x in res.convergence_history] == [x[1] for x in ref.convergence_history]
    assert not ck.exists()  # removed on success (keep_on_success=False)


def test_foreign_checkpoint_is_refused(tmp_path):
    req = _request()
    ck = tmp_path / "run.ckpt.npz"
    with pytest.raises(RunInterrupted):
        _run(req, checkpoint=CheckpointSpec(ck, every_periods=1, stop_when=lambda: True))
    other = _request()
    other.source = dataclasses.replace(other.source, amplitude=other.source.amplitude * 2.0)
    with pytest.raises(CheckpointMismatch):
        _run(other, checkpoint=CheckpointSpec(ck, every_periods=1))


def test_staggered_resume_is_bitwise_and_the_scheme_is_fingerprinted(tmp_path):
    """The opt-in staggered scheme resumes from the library's raw state (p, the velocity
    spectra and the split density) bit for bit, and a checkpoint of the reference scheme
    does not resume a staggered solve (nor the other way round)."""
    req = _request(seed=9, solver="westervelt")
    kw = dict(scheme="staggered", alpha_power=1.1)
    print("[checkpoint] staggered: uninterrupted", flush=True)
    ref = _run(req, **kw)
    calls = {"n": 0}

    def stop_after_three():
        calls["n"] += 1
        return calls["n"] >= 3

    ck = tmp_path / "stg.ckpt.npz"
    print("[checkpoint] staggered: interrupted", flush=True)
    with pytest.raises(RunInterrupted):
        _run(req, checkpoint=CheckpointSpec(ck, every_periods=2, stop_when=stop_after_three), **kw)
    print("[checkpoint] staggered: resumed", flush=True)
    res = _run(req, checkpoint=CheckpointSpec(ck, every_periods=2, keep_on_success=True), **kw)
    assert res.steps_total == ref.steps_total
    for h in ref.phasors:
        np.testing.assert_array_equal(res.phasors[h], ref.phasors[h])
    np.testing.assert_array_equal(res.p_max, ref.p_max)
    assert res.meta["scheme"] == "cw-kspace-pstd-staggered/1"
    with pytest.raises(CheckpointMismatch):
        _run(req, checkpoint=CheckpointSpec(ck, every_periods=2))  # the reference scheme
