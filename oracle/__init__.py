"""CPU oracle for the CW k-space solver and the Pennes/CEM43 stage.

TEST INFRASTRUCTURE ONLY. This package restates, in NumPy, the upstream
caustica algorithms that the HIP path replaces, so that tests can check the
GPU results against them. It is imported only by tests/, by
``__graft_entry__.smoke()`` (as the checker) and by ``bench.py``'s CPU-baseline
leg (as the timed CPU restatement). The product package ``caustica_amd`` never
imports it and never falls back to it.

Modules and what each restates (upstream file:line):
  kspace.py   solvers/kspace/engine.py:86-586, solvers/kspace/operators.py:23-157,
              core/pml.py:53-73, sources.py:98-102, spectral.py:14-54
  sources.py  sources.py:105-175, analytic/geometry.py:15-64
  pennes.py   thermal/pennes.py:376-689, thermal/dose.py:74-89
  heating.py  sensors.py:193-295, sensors.py:356-435
  analytic.py analytic/oneill.py:33-67, analytic/planewave.py:18-57,
              analytic/rayleigh.py:30-84; the power-law dispersion relation
  staggered.py the opt-in staggered / power-law / split-PML scheme, which has no
              upstream counterpart (upstream MILESTONES.md:1201 plans it); float64
              only, pinned by closed forms (tests/test_oracle_staggered.py)

Every function runs in one of two precisions:
  "fp32"  the upstream dtype and operation order (float32 state, complex64
          spectra, scipy.fft); pinned bit-for-bit or to float32 round-off
          against tests/golden/*.npz, which were written by the upstream code;
  "fp64"  the same mathematics in float64/complex128, with the source and the
          phasor kernels evaluated at exact phases 2*pi*(n mod spp)/spp. This
          is the parity oracle the GPU path is graded against.
"""
