"""Solver contract: capabilities, run policy, result (upstream ``src/caustica/solvers/base.py``).

``CWRunSpec`` keeps the upstream defaults and validators (:55-83); ``SolverResult``
keeps the upstream field list (:85-120); ``check_source_clears_pml`` (:132-169) and
``SolverBase.validate`` (:178-201) refuse the same setups with the same errors.
"""

from __future__ import annotations

from abc import ABC, abstractmethod
from dataclasses import dataclass, field
from typing import Any, ClassVar

import numpy as np
from pydantic import BaseModel, ConfigDict, Field, model_validator

from caustica_amd.core.grid import Grid
from caustica_amd.medium import Medium
from caustica_amd.sources import CWSource


class SolverCapabilityError(ValueError):
    """A setup asked a solver for something it declares it cannot do."""


class SolverDivergedError(RuntimeError):
    """The field stopped being a number (inf/NaN)."""


@dataclass(frozen=True)
class SolverCaps:
    ndim: frozenset
    nonlinear: bool
    drive: frozenset
    backends: frozenset
    absorption: frozenset = frozenset({"exponential"})


class CWRunSpec(BaseModel):
    """Steady-state CW run policy (exact-period dt implied)."""

    model_config = ConfigDict(extra="forbid", frozen=True)

    cfl: float = Field(0.48, gt=0.0, le=0.6)
    cfl_hard_max: float = Field(0.50, gt=0.0, le=0.6)
    min_settle_periods: int = Field(8, ge=0)
    max_settle_periods: int = Field(96, ge=1)
    convergence_tol: float = Field(0.01, ge=0.0)
    n_record_periods: int = Field(2, ge=1)
    t_end_min_us: float | None = Field(None, gt=0.0)

    @model_validator(mode="after")
    def _ranges(self) -> CWRunSpec:
        if self.max_settle_periods < self.min_settle_periods:
            raise ValueError(
                f"max_settle_periods ({self.max_settle_periods}) < min_settle_periods "
                f"({self.min_settle_periods})"
            )
        if self.cfl > self.cfl_hard_max:
            raise ValueError(f"cfl ({self.cfl}) exceeds cfl_hard_max ({self.cfl_hard_max})")
        return self


@dataclass
class SolverResult:
    """Steady-state CW result over ``region`` of the active grid."""

    phasor: np.ndarray
    p_max: np.ndarray
    region: tuple
    dt: float
    spp: int
    steps_total: int
    t_end_s: float
    tof_periods: int
    converged_period: int
    settle_capped: bool
    convergence_history: list
    phasors: dict = field(default_factory=dict)
    meta: dict[str, Any] = field(default_factory=dict)

    @property
    def amp(self) -> np.ndarray:
        return np.abs(self.phasor).astype(np.float32)

    @property
    def phase(self) -> np.ndarray:
        return np.angle(self.phasor).astype(np.float32)

    def harmonic_amp(self, n: int) -> np.ndarray:
        if n not in self.phasors:
            raise KeyError(
                f"harmonic {n} was not recorded (available: {sorted(self.phasors)}); "
                f"pass harmonics=(1, {n}) to run()."
            )
        return np.abs(self.phasors[n]).astype(np.float32)


def interior_slices(shape, margin: int) -> tuple:
    if margin < 0:
        raise ValueError(f"margin must be >= 0, got {margin}")
    if any(n <= 2 * margin for n in shape):
        raise ValueError(f"margin {margin} leaves no interior for shape {shape}")
    return tuple(slice(margin, n - margin) for n in shape)


def check_source_clears_pml(grid: Grid, source: CWSource) -> None:
    """Refuse a source whose every voxel lies inside the sponge band."""
    band = grid.pml_vox
    if band <= 0 or any(n <= 2 * band for n in grid.shape):
        return
    idx = source.indices
    upper = np.asarray(grid.shape) - band
    if ((idx >= band) & (idx < upper)).all(axis=1).any():
        return
    lo = tuple(int(v) for v in idx.min(axis=0))
    hi = tuple(int(v) for v in idx.max(axis=0))
    raise SolverCapabilityError(
        f"every source voxel (span {lo}..{hi}) lies inside the PML band ({band} voxels on "
        f"each face of {grid.shape}), where it is damped as fast as it is driven. Move the "
        f"source inward, thin the PML, or enlarge the grid."
    )


class SolverBase(ABC):
    name: ClassVar[str]
    caps: ClassVar[SolverCaps]

    def validate(self, grid: Grid, medium: Medium, source: CWSource) -> None:
        if grid.ndim not in self.caps.ndim:
            raise SolverCapabilityError(
                f"solver '{self.name}' supports {sorted(self.caps.ndim)}-D grids, got {grid.ndim}-D."
            )
        if tuple(medium.shape) != tuple(grid.shape):
            raise ValueError(f"medium shape {medium.shape} != grid shape {grid.shape}")
        if not medium.is_linear and not self.caps.nonlinear:
            raise SolverCapabilityError(
                f"solver '{self.name}' is linear-only but the medium has beta != 0. Zero the "
                f"medium's beta or pick a nonlinear solver (e.g. 'westervelt')."
            )
        if "cw" not in self.caps.drive:
            raise SolverCapabilityError(f"solver '{self.name}' does not support CW drive.")
        source.check_inside(grid)
        check_source_clears_pml(grid, source)

    @abstractmethod
    def run(self, grid, medium, source, spec=None, **kwargs) -> SolverResult:
        """Execute one steady-state CW solve."""
