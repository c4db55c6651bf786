"""Every instantiated axis length of the pass kernels against the float64 oracle.

For each length the library supports (x and y: complex passes of length L; z: real
length 2M), one Westervelt time step from a random state on a small heterogeneous,
absorbing grid whose other axes are tiny, at a step inside the source ramp, compared
with the float64 oracle (max |err| <= 2e-5 of max |field|, the one-step bar of
tests/test_gpu_kspace.py). This exercises every FFT split, the k-space operator on
odd and even lengths (Nyquist handling), the derivative factors carried by the y
passes and the real-split z transform.
"""

import numpy as np
import pytest

from oracle.kspace import KSpaceOracle, RunSpec

pytestmark = pytest.mark.gpu


def _lengths(axis):
    from caustica_amd import _native as N

    lib = N.lib()
    # Grid needs >= 4 voxels per axis; lengths 1 and 2 only occur as the unit axes of
    # embedded 1-D / 2-D grids, which the fixture cases of test_gpu_kspace.py cover
    if axis < 2:
        return [n for n in range(4, 513) if lib.cst_supported_length(n, axis)]
    return [n for n in range(4, 1025, 2) if lib.cst_supported_length(n, 2)]


def _case(shape, seed):
    from caustica_amd import CWSource, Grid, Medium, PMLSpec
    from caustica_amd.solvers import CWRunSpec

    rng = np.random.default_rng(seed)
    dx = 2e-4
    c = (1500.0 + 80.0 * rng.standard_normal(shape)).astype(np.float32)
    rho = (1000.0 + 60.0 * rng.standard_normal(shape)).astype(np.float32)
    alpha = (5.0 + 3.0 * rng.random(shape)).astype(np.float32)
    beta = (3.5 + rng.random(shape)).astype(np.float32)
    grid = Grid(shape, dx, PMLSpec(thickness=2 * dx))
    medium = Medium(alpha=alpha, rho=rho, c=c, beta=beta)
    n_src = min(6, int(np.prod(shape)))
    flat = rng.choice(int(np.prod(shape)), size=n_src, replace=False)
    idx = np.stack(np.unravel_index(flat, shape), axis=1).astype(np.int64)
    source = CWSource(indices=idx, phases=rng.uniform(0, 6.28, n_src).astype(np.float32),
                      amplitude=1e5, f0=1e6, ramp_periods=3.0)
    spec = CWRunSpec(cfl=0.3, convergence_tol=0.0, min_settle_periods=2, max_settle_periods=2,
                     n_record_periods=1)
    return grid, medium, source, spec


def _one_step(shape, seed=3):
    from caustica_amd.solvers.kspace.native import KSpaceSession
    from caustica_amd.solvers.kspace.operators import plan_member

    grid, medium, source, spec = _case(shape, seed)
    plan = plan_member(grid, medium, source, spec, nonlinear=True)
    assert plan.padded == tuple(shape)
    orc = KSpaceOracle(shape=shape, dx=grid.dx, pml_vox=grid.pml_vox, pml_edge=grid.pml.edge,
                       alpha=medium.alpha, rho=medium.rho, c=medium.c, beta=medium.beta,
                       src_idx=source.indices, src_phase=source.phases,
                       amplitude=source.amplitude, f0=source.f0,
                       ramp_periods=source.ramp_periods, spec=RunSpec(**spec.model_dump()),
                       harmonics=(1,), nonlinear=True, precision="fp64")
    rng = np.random.default_rng(seed + 1)
    p0 = rng.standard_normal(shape).astype(np.float32) * 1e4
    u0 = [rng.standard_normal(shape).astype(np.float32) * 1e-2 for _ in range(3)]
    n0 = plan.spp + 2  # inside the 3-period ramp
    orc.p[...] = p0
    for i in range(3):
        orc.u[i][...] = u0[i]
    orc.step(n0)
    with KSpaceSession(plan.padded3, 1) as s:
        slot = s.add(plan)
        s.set_state(slot, n0, p0, *u0)
        s.advance([slot], 1)
        (p1, *u1), _ = s.get_state(slot)
    errs = [np.abs(p1 - orc.p).max() / np.abs(orc.p).max()]
    for i in range(3):
        errs.append(np.abs(u1[i] - orc.u[i]).max() / np.abs(orc.u[i]).max())
    return max(errs)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_every_supported_length(axis):
    bad = []
    for n in _lengths(axis):
        shape = [4, 4, 8]
        shape[axis] = n
        shape = tuple(shape)
        print(f"[lengths] axis {axis} n {n} shape {shape}", flush=True)
        err = _one_step(shape)
        if not err < 2e-5:
            bad.append((n, err))
    assert not bad, f"axis {axis}: lengths over the bar: {bad}"
