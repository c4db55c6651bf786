"""Solver registry and the built-in CW solvers (upstream ``src/caustica/solvers/__init__.py``).

    import caustica_amd.solvers as solvers
    result = solvers.get("linear")().run(grid, medium, source, backend="hip")
"""

from caustica_amd.solvers.base import (
    CWRunSpec,
    SolverBase,
    SolverCapabilityError,
    SolverCaps,
    SolverDivergedError,
    SolverResult,
    interior_slices,
)
from caustica_amd.solvers.kspace.linear import LinearKSpacePSTD
from caustica_amd.solvers.kspace.westervelt import WesterveltKSpacePSTD
from caustica_amd.solvers.registry import available, get, register

__all__ = [
    "CWRunSpec",
    "LinearKSpacePSTD",
    "SolverBase",
    "SolverCapabilityError",
    "SolverCaps",
    "SolverDivergedError",
    "SolverResult",
    "WesterveltKSpacePSTD",
    "available",
    "get",
    "interior_slices",
    "register",
]
