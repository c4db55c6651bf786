"""caustica_amd: the CW k-space solver (linear / Westervelt) and the Pennes/CEM43 stage of
caustica, as a ``hip`` backend for AMD Instinct MI355X (gfx950).

Same object API as upstream caustica (Grid, PMLSpec, Medium, CWSource, solver registry,
CWRunSpec, SolverResult, get_backend); the step loop runs in libcaustica_hip.so.
"""

from caustica_amd.core.backend import Backend, CausticaWarning, backends, get_backend
from caustica_amd.core.grid import Grid
from caustica_amd.core.pml import PMLSpec
from caustica_amd.materials import Material, water
from caustica_amd.medium import Medium
from caustica_amd.sources import CWSource, bowl_cw_source, plane_cw_source

__version__ = "0.1.0"

__all__ = [
    "Backend",
    "CWSource",
    "CausticaWarning",
    "Grid",
    "Material",
    "Medium",
    "PMLSpec",
    "backends",
    "bowl_cw_source",
    "get_backend",
    "plane_cw_source",
    "water",
]
