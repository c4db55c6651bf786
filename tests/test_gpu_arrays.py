"""Phased spiral array through the hip solver (upstream tests/test_arrays.py:60-103 and
the solver integration case :134-160, which upstream runs on its CPU backend).

A 32-element spiral (30 mm aperture and ROC) voxelized on a 96^3 water grid at 0.5 mm:

* uniform phases focus on the axis (within one voxel laterally) between the O'Neil
  axial peak of the full bowl (minus one voxel) and one voxel past the geometric focus;
* delay-and-sum phases towards a target 3 mm / -2 mm lateral and 4 mm proximal move
  the peak by the commanded displacement within lambda/2 laterally and lambda axially
  (the focal-shift bias cancels in the difference), and the amplitude at the target
  rises above three times its uniform-phase value.
"""

import numpy as np
import pytest

import caustica_amd.solvers as solvers
from caustica_amd import Grid, Medium, PMLSpec
from caustica_amd.arrays import archimedean_spiral
from caustica_amd.materials import water
from caustica_amd.solvers import CWRunSpec
from oracle.analytic import oneill_axial

pytestmark = pytest.mark.gpu

F0, C0 = 1.0e6, 1500.0
DX = 0.5e-3
APEX = (48, 48, 12)
FOCUS = (48, 48, 72)  # apex + ROC / dx


@pytest.fixture(scope="module")
def setup():
    arr = archimedean_spiral(n_elements=32, d_outer=0.030, d_inner=0.010, roc=0.030,
                             active_fraction=0.60)
    grid = Grid((96, 96, 96), DX, PMLSpec(thickness=5e-3))
    med = Medium.homogeneous(grid.shape, water(c=C0))
    return arr, grid, med


def _solve(arr, grid, med, phases):
    asrc = arr.voxelize(grid, APEX, f0=F0, amplitude=1e5, phases=phases)
    spec = CWRunSpec(min_settle_periods=8, max_settle_periods=30)
    res = solvers.get("linear")().run(grid, med, asrc.source, spec, backend="hip",
                                      reference_point=FOCUS)
    amp = res.amp
    assert np.isfinite(amp).all()
    return amp, np.array(np.unravel_index(np.argmax(amp), amp.shape))


def test_uniform_phases_focus_at_geometric_focus(setup):
    arr, grid, med = setup
    print("[arrays] uniform-phase spiral solve", flush=True)
    _, peak = _solve(arr, grid, med, None)
    assert abs(peak[0] - FOCUS[0]) <= 1 and abs(peak[1] - FOCUS[1]) <= 1
    z = np.linspace(0.010, 0.045, 2001)
    z_on = z[np.argmax(np.abs(oneill_axial(z, 0.015, 0.030, F0, C0)))]
    assert APEX[2] + z_on / DX - 1.0 <= peak[2] <= FOCUS[2] + 1, peak


def test_das_phases_steer_by_commanded_displacement(setup):
    arr, grid, med = setup
    lam = C0 / F0
    target = np.array([3e-3, -2e-3, arr.focal_length - 4e-3])
    commanded = target - arr.focus
    print("[arrays] uniform and steered spiral solves", flush=True)
    amp0, p0 = _solve(arr, grid, med, None)
    amp1, p1 = _solve(arr, grid, med, arr.das_phases(target, F0, C0))
    realized = (p1 - p0) * DX
    assert np.linalg.norm((realized - commanded)[:2]) <= lam / 2, realized
    assert abs((realized - commanded)[2]) <= lam, realized
    tv = tuple(int(round(v)) for v in np.asarray(APEX) + target / DX)
    assert amp1[tv] > 3.0 * amp0[tv]
