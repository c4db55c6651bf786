"""Westervelt k-space PSTD solver, registry name ``"westervelt"`` (upstream ``solvers/kspace/westervelt.py``).

Identical engine to ``linear`` plus the term dp = -2 beta dt p div(u); a beta = 0
medium takes the linear code path bit for bit.
"""

from __future__ import annotations

from typing import Any

from caustica_amd.solvers.base import CWRunSpec, SolverBase, SolverCaps, SolverResult
from caustica_amd.solvers.kspace.engine import run_cw_kspace_pstd
from caustica_amd.solvers.registry import register


@register
class WesterveltKSpacePSTD(SolverBase):
    """Nonlinear (Westervelt) full-wave k-space PSTD (1/2/3-D, CW steady state) on hip."""

    name = "westervelt"
    caps = SolverCaps(
        ndim=frozenset({1, 2, 3}),
        nonlinear=True,
        drive=frozenset({"cw"}),
        backends=frozenset({"hip"}),
        absorption=frozenset({"exponential", "power_law"}),
    )

    def run(self, grid, medium, source, spec: CWRunSpec | None = None, backend: str = "auto",
            record_region=None, reference_point=None, harmonics=(1,), checkpoint=None,
            progress=None, *, scheme: str = "reference", alpha_power: float = 1.5,
            pml_alpha: float = 2.0, **kwargs: Any) -> SolverResult:
        """``scheme="staggered"`` selects the opt-in staggered / power-law / split-PML
        scheme (no upstream counterpart; ``alpha_power`` y of alpha0 omega^y, ``pml_alpha``
        the PML strength); the default runs the upstream scheme."""
        if kwargs:
            raise TypeError(f"unknown run() options: {sorted(kwargs)}")
        spec = spec or CWRunSpec()
        self.validate(grid, medium, source)
        return run_cw_kspace_pstd(self.name, grid, medium, source, spec, backend=backend,
                                  record_region=record_region, reference_point=reference_point,
                                  nonlinear=True, harmonics=harmonics, checkpoint=checkpoint,
                                  progress=progress, scheme=scheme, alpha_power=alpha_power,
                                  pml_alpha=pml_alpha)
