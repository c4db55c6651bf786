"""The full-size float64 fixtures' inputs are rebuilt exactly on the host.

Every full-size fixture (tests/golden/full_*.npz, written by tests/golden/make_fullsize.py)
records the seeded workload it was run on: a digest and strided samples of the medium,
the time step and samples per period, the time of flight and the step count. The GPU
parity tests rebuild the workload from the seed and compare hip results with the
fixture, so a change in the workload generators (random field, layering, bowl, steering,
pinned schedule) would show there as a parity failure. This checks the inputs alone, on
the CPU: the rebuilt medium matches the recorded sample and digest, and the host
discretization gives the recorded schedule.
"""

import numpy as np
import pytest

from tests import _full

CASES = ["cfg1", "cfg2", "serving_w128x128x256", "serving_l192"]


def _workload(case, meta):
    from caustica_amd import workloads

    if case == "cfg1":
        return workloads.cfg1_hetero_linear(seed=meta["seed"])
    if case == "cfg2":
        return workloads.cfg2_westervelt(seed=meta["seed"])
    return workloads.serving_workload(meta["seed"], tuple(meta["shape"]), meta["solver"])


@pytest.mark.parametrize("case", CASES)
def test_fixture_inputs_rebuild(case):
    from caustica_amd.solvers.kspace.operators import cw_discretization, cw_tof_periods

    path = _full.GOLDEN / f"full_{case}.npz"
    if not path.exists():
        pytest.skip(f"{path.name} not generated (tests/golden/make_fullsize.py)")
    fx = _full.load(f"full_{case}")
    meta = fx["meta"]
    wl = _workload(case, meta)
    assert wl.name == meta["workload"] and wl.solver == meta["solver"]
    assert list(wl.grid.shape) == meta["shape"] and list(wl.harmonics) == meta["harmonics"]
    assert [int(v) for v in wl.reference_point] == meta["focus"]
    got = _full.medium_sample(wl.medium)
    for k in ("c_sample", "rho_sample"):
        np.testing.assert_array_equal(got[k], fx[k], err_msg=k)
    assert _full.medium_digest(wl.medium) == meta["medium_sha256"]
    spp, dt = cw_discretization(wl.grid.dx, wl.medium.c_max, wl.spec, wl.source.f0, wl.harmonics)
    tof = cw_tof_periods(wl.grid.shape, wl.grid.dx, wl.medium.c_min, wl.source.indices,
                         wl.source.f0, wl.reference_point)
    assert (spp, tof) == (meta["spp"], meta["tof_periods"])
    assert dt == pytest.approx(meta["dt"], rel=1e-15)
    assert wl.periods_total * spp == meta["steps_total"]


def test_serving_workload_is_deterministic():
    from caustica_amd import workloads

    a = workloads.serving_workload(7, (64, 64, 96), "westervelt")
    b = workloads.serving_workload(7, (64, 64, 96), "westervelt")
    c = workloads.serving_workload(8, (64, 64, 96), "westervelt")
    assert a.name == b.name and a.periods_total == b.periods_total and a.harmonics == (1, 2)
    for k in ("c", "rho", "alpha", "beta"):
        assert np.array_equal(getattr(a.medium, k), getattr(b.medium, k)), k
    np.testing.assert_array_equal(a.source.indices, b.source.indices)
    np.testing.assert_array_equal(a.source.phases, b.source.phases)
    assert a.reference_point == b.reference_point
    # another seed draws another random field and steering offset
    assert not np.array_equal(a.medium.c, c.medium.c)
