"""Isotropic Cartesian grid (1/2/3-D) -- upstream ``src/caustica/core/grid.py:28-150``."""

from __future__ import annotations

import numpy as np

from caustica_amd.core.pml import PMLSpec


class Grid:
    """Shape, spacing and an optional sponge PML; geometry only (no time step)."""

    def __init__(self, shape, dx: float, pml: PMLSpec | None = None):
        shape = tuple(int(s) for s in shape)
        if not 1 <= len(shape) <= 3:
            raise ValueError(f"Grid supports 1..3 dimensions, got shape {shape}")
        if any(s < 4 for s in shape):
            raise ValueError(f"every axis needs >= 4 voxels, got shape {shape}")
        if not (np.isfinite(dx) and dx > 0):
            raise ValueError(f"dx must be a positive finite number in meters, got {dx}")
        self.shape = shape
        self.dx = float(dx)
        self.pml = pml

    @property
    def ndim(self) -> int:
        return len(self.shape)

    @property
    def n_voxels(self) -> int:
        return int(np.prod(self.shape))

    @property
    def extent(self) -> tuple[float, ...]:
        return tuple(s * self.dx for s in self.shape)

    @property
    def pml_vox(self) -> int:
        return 0 if self.pml is None else self.pml.thickness_vox(self.dx)

    def axis_coords(self, axis: int, centered: bool = True) -> np.ndarray:
        n = self.shape[axis]
        x = np.arange(n, dtype=np.float64) * self.dx
        if centered:
            x -= x[n // 2]
        return x

    def voxels_for(self, length_m: float) -> int:
        return int(round(length_m / self.dx))

    def ppw(self, f0: float, c: float) -> float:
        if f0 <= 0 or c <= 0:
            raise ValueError("f0 and c must be > 0")
        return c / (f0 * self.dx)

    def __repr__(self) -> str:  # pragma: no cover - cosmetic
        pml = f", pml={self.pml_vox}vox" if self.pml is not None else ""
        return f"Grid(shape={self.shape}, dx={self.dx * 1e3:.3g} mm{pml})"
