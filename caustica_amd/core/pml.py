"""Gaussian sponge PML: specification and 1-D profile (upstream ``core/pml.py:30-73``).

``s(r) = exp(-edge ((width - r) / width)^2)`` for r = 0..width-1 at both ends of an
axis, 1 inside; applied multiplicatively to p and u each step as a separable product.
The profile is computed in float32 exactly as upstream, and handed to the device as
three 1-D vectors (the device forms ((sx sy) sz) per voxel instead of reading a
dense sponge volume).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass(frozen=True)
class PMLSpec:
    thickness: float
    edge: float = 2.0

    def __post_init__(self) -> None:
        if self.thickness < 0:
            raise ValueError(f"PML thickness must be >= 0, got {self.thickness}")
        if self.edge <= 0:
            raise ValueError(f"PML edge factor must be > 0, got {self.edge}")

    def thickness_vox(self, dx: float) -> int:
        if dx <= 0:
            raise ValueError(f"dx must be > 0, got {dx}")
        return int(round(self.thickness / dx))


def sponge_profile_1d(n: int, width: int, edge: float = 2.0) -> np.ndarray:
    """float32 profile of length n: ones with Gaussian ramps of ``width`` cells at both ends."""
    if n <= 0:
        raise ValueError(f"profile length must be > 0, got {n}")
    if width < 0:
        raise ValueError(f"sponge width must be >= 0, got {width}")
    if 2 * width > n:
        raise ValueError(f"sponge width {width} does not fit twice into length {n}")
    s = np.ones(n, dtype=np.float32)
    if width > 0:
        r = np.arange(width, dtype=np.float32)
        ramp = np.exp(-edge * ((width - r) / width) ** 2)
        s[:width] = ramp
        s[-width:] = ramp[::-1]
    return s
