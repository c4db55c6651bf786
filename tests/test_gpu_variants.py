"""Every thread-layout variant of the tunable z passes computes the same step.

The variants (zu: 0, 1, 2; zp: 0, 1) change how lines are spread over threads and
blocks, and so which products the compiler fuses into FMAs: their results differ at the
float32 rounding level (2-3e-7 of the field per step). Checked on nz = 256 and 384 (the
z lengths of the benchmark workloads) and 320 (a 16 x 10 split; like 384, one-wave blocks
whose line buffers share LDS in layout 2) with small x/y, heterogeneous, absorbing and
nonlinear: one step from a random state under every combination against the float64
oracle (the one-step bar of tests/test_gpu_kspace.py); twenty more steps against the
(0, 0) combination (within 1e-5 of the field); and the same 21 steps repeated from the
same state bitwise equal (a combination whose blocks raced on LDS would not repeat).
"""

import itertools

import numpy as np
import pytest

from oracle.kspace import KSpaceOracle, RunSpec
from tests.test_gpu_fft_lengths import _case

pytestmark = pytest.mark.gpu


def _rel(a, b) -> float:
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("shape", [(16, 24, 256), (8, 16, 384), (8, 8, 320)])
def test_variants_agree_and_match_oracle(shape):
    from caustica_amd.solvers.kspace.native import KSpaceSession
    from caustica_amd.solvers.kspace.operators import plan_member

    grid, medium, source, spec = _case(shape, seed=11)
    plan = plan_member(grid, medium, source, spec, nonlinear=True)
    rng = np.random.default_rng(12)
    p0 = rng.standard_normal(shape).astype(np.float32) * 1e4
    u0 = [rng.standard_normal(shape).astype(np.float32) * 1e-2 for _ in range(3)]
    n0 = plan.spp + 2
    orc = KSpaceOracle(shape=shape, dx=grid.dx, pml_vox=grid.pml_vox, pml_edge=grid.pml.edge,
                       alpha=medium.alpha, rho=medium.rho, c=medium.c, beta=medium.beta,
                       src_idx=source.indices, src_phase=source.phases,
                       amplitude=source.amplitude, f0=source.f0,
                       ramp_periods=source.ramp_periods, spec=RunSpec(**spec.model_dump()),
                       harmonics=(1,), nonlinear=True, precision="fp64")
    orc.p[...] = p0
    for i in range(3):
        orc.u[i][...] = u0[i]
    orc.step(n0)
    after = {}
    with KSpaceSession(plan.padded3, 1) as s:
        slot = s.add(plan)
        for vu, vp in itertools.product((0, 1, 2), (0, 1)):
            print(f"[variants] {shape} zu {vu} zp {vp}", flush=True)
            s.set_variant(2, vu)
            s.set_variant(3, vp)
            s.set_state(slot, n0, p0, *u0)
            s.advance([slot], 1)
            (p1, *u1), _ = s.get_state(slot)
            err = max(_rel(p1, orc.p), *(_rel(u1[i], orc.u[i]) for i in range(3)))
            assert err < 2e-5, (vu, vp, err)
            s.advance([slot], 20)
            (p21, *u21), _ = s.get_state(slot)
            raw = s.get_raw(slot)[0]
            s.set_state(slot, n0, p0, *u0)
            s.advance([slot], 21)
            assert np.array_equal(s.get_raw(slot)[0], raw), ("repeat", vu, vp)
            after[(vu, vp)] = (p21, u21)
    p_ref, u_ref = after[(0, 0)]
    for (vu, vp), (p21, u21) in after.items():
        d = max(_rel(p21, p_ref), *(_rel(u21[i], u_ref[i]) for i in range(3)))
        print(f"[variants] {shape} zu {vu} zp {vp}: 21 steps vs (0, 0) {d:.2e}", flush=True)
        assert d < 1e-5, (vu, vp, d)
