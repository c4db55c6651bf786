"""Every variant of the tunable passes computes the same step.

The variants (tune.py: tile widths 0..2 of the y and x passes; thread layouts zu 0..2 /
zp 0..1, each with the second radix plan + 4) change how lines are spread over threads
and blocks, how each line's FFT is split into stages, and so which products the compiler
fuses into FMAs: their results differ at the float32 rounding level (2-3e-7 of the field
per step). Checked on nz = 256 and 384 (the z lengths of the benchmark workloads), 128
(the serving shapes' 8 x 8 split, whose two-threads-per-row layout shares each stage-2
DFT between two threads), 320 (a 16 x 10 split; like 384, one-wave blocks whose line
buffers share LDS in layout 2), 1024 (a 32 x 16 split whose layout-2 one-wave blocks both
share LDS and pair their stage-2 DFTs) and 96 (an 8 x 6 split), with x/y lengths below 64
(one tile width) and at 64 and 128 (three),
heterogeneous, absorbing and nonlinear: one step from a random state under each pass's
variants (the other passes at 0) against the float64 oracle (the one-step bar of
tests/test_gpu_kspace.py); twenty more steps against the all-default combination (within
1e-5 of the field); and the same 21 steps repeated from the same state bitwise equal (a
variant whose blocks raced on LDS would not repeat).
"""

import numpy as np
import pytest

from caustica_amd.tune import variants_for
from oracle.kspace import KSpaceOracle, RunSpec
from tests.test_gpu_fft_lengths import _case

pytestmark = pytest.mark.gpu


def _rel(a, b) -> float:
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("shape", [(16, 24, 256), (8, 16, 384), (8, 8, 320), (64, 128, 96),
                                   (128, 64, 256), (16, 24, 128), (8, 8, 1024)])
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
    combos = [(0, 0, 0, 0)]
    for p_ in range(4):
        for v in variants_for(p_, plan.padded3):
            c = [0, 0, 0, 0]
            c[p_] = v
            if tuple(c) not in combos:
                combos.append(tuple(c))
    after = {}
    with KSpaceSession(plan.padded3, 1) as s:
        slot = s.add(plan)
        for combo in combos:
            print(f"[variants] {shape} y/x/zu/zp {combo}", flush=True)
            for p_, v in enumerate(combo):
                s.set_variant(p_, v)
            s.set_state(slot, n0, p0, *u0)
            s.advance([slot], 1)
            (p1, *u1), _ = s.get_state(slot)
            err = max(_rel(p1, orc.p), *(_rel(u1[i], orc.u[i]) for i in range(3)))
            assert err < 2e-5, (combo, err)
            s.advance([slot], 20)
            (p21, *u21), _ = s.get_state(slot)
            raw = s.get_raw(slot)[0]
            s.set_state(slot, n0, p0, *u0)
            s.advance([slot], 21)
            assert np.array_equal(s.get_raw(slot)[0], raw), ("repeat", combo)
            after[combo] = (p21, u21)
    p_ref, u_ref = after[(0, 0, 0, 0)]
    for combo, (p21, u21) in after.items():
        d = max(_rel(p21, p_ref), *(_rel(u21[i], u_ref[i]) for i in range(3)))
        print(f"[variants] {shape} {combo}: 21 steps vs defaults {d:.2e}", flush=True)
        assert d < 1e-5, (combo, d)
