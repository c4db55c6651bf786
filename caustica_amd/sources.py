"""CW sources: voxel indices + per-voxel phase + amplitude + f0 + ramp.

Upstream ``src/caustica/sources.py:25-175`` and ``analytic/geometry.py:15-64``
(Fibonacci equal-area spherical-cap sampling). The bowl voxel set is the rounded
cap sampling at dx/2 spacing, deduplicated with ``np.unique`` (same set as upstream:
pinned by tests/golden/bowl_*.npz).
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from caustica_amd.core.grid import Grid

_GOLDEN_ANGLE = np.pi * (3.0 - np.sqrt(5.0))


@dataclass(frozen=True)
class CWSource:
    indices: np.ndarray
    phases: np.ndarray
    amplitude: float
    f0: float
    ramp_periods: float = 3.0
    label: str = field(default="", compare=False)

    def __post_init__(self) -> None:
        idx = np.asarray(self.indices)
        ph = np.asarray(self.phases, dtype=np.float32)
        if idx.ndim != 2:
            raise ValueError(f"indices must be (n, ndim), got shape {idx.shape}")
        if not np.issubdtype(idx.dtype, np.integer):
            raise TypeError(f"indices must be integer voxel coords, got {idx.dtype}")
        if ph.shape != (idx.shape[0],):
            raise ValueError(f"phases must be ({idx.shape[0]},), got {ph.shape}")
        if self.amplitude <= 0 or self.f0 <= 0:
            raise ValueError("amplitude and f0 must be > 0")
        if self.ramp_periods <= 0:
            raise ValueError("ramp_periods must be > 0")
        if np.unique(idx, axis=0).shape[0] != idx.shape[0]:
            raise ValueError(
                "source contains duplicate voxel indices; deduplicate (and decide the phase) "
                "before constructing the CWSource."
            )
        object.__setattr__(self, "indices", np.ascontiguousarray(idx.astype(np.int64)))
        object.__setattr__(self, "phases", np.ascontiguousarray(ph))

    @property
    def n_points(self) -> int:
        return int(self.indices.shape[0])

    @property
    def ndim(self) -> int:
        return int(self.indices.shape[1])

    def check_inside(self, grid: Grid) -> None:
        if self.ndim != grid.ndim:
            raise ValueError(f"source is {self.ndim}-D but grid is {grid.ndim}-D")
        upper = np.asarray(grid.shape)
        bad = np.where((self.indices < 0).any(1) | (self.indices >= upper).any(1))[0]
        if bad.size:
            raise ValueError(
                f"{bad.size} source voxel(s) fall outside grid {grid.shape} "
                f"(first offender: {self.indices[bad[0]].tolist()})"
            )


def ramp_envelope(t: float, period: float, ramp_periods: float) -> float:
    """Cosine taper 0 -> 1 over ``ramp_periods`` periods."""
    x = min(t / (ramp_periods * period), 1.0)
    return 0.5 * (1.0 - np.cos(np.pi * x))


def spherical_cap_points(aperture_radius: float, roc: float, spacing: float):
    """Equal-area Fibonacci points on a concave cap (apex at 0, focus at (0, 0, roc))."""
    a, f = float(aperture_radius), float(roc)
    if not 0 < a < f:
        raise ValueError(f"need 0 < aperture_radius < roc, got a={a}, roc={f}")
    if spacing <= 0:
        raise ValueError(f"spacing must be > 0, got {spacing}")
    cos_tmax = np.sqrt(1.0 - (a / f) ** 2)
    cap_area = 2.0 * np.pi * f**2 * (1.0 - cos_tmax)
    n = max(int(np.ceil(cap_area / spacing**2)), 16)
    j = np.arange(n, dtype=np.float64)
    cos_t = 1.0 - (1.0 - cos_tmax) * (j + 0.5) / n
    sin_t = np.sqrt(np.clip(1.0 - cos_t**2, 0.0, None))
    phi = j * _GOLDEN_ANGLE
    points = np.column_stack((f * sin_t * np.cos(phi), f * sin_t * np.sin(phi), f * (1.0 - cos_t)))
    normals = (np.array([0.0, 0.0, f]) - points) / f
    areas = np.full(n, cap_area / n)
    return points, normals, areas


def plane_cw_source(grid: Grid, f0: float, amplitude: float, axis: int = 0,
                    position_vox: int | None = None, phase: float = 0.0) -> CWSource:
    if not 0 <= axis < grid.ndim:
        raise ValueError(f"axis {axis} invalid for a {grid.ndim}-D grid")
    pos = position_vox if position_vox is not None else grid.pml_vox + 8
    if not 0 <= pos < grid.shape[axis]:
        raise ValueError(f"plane position {pos} outside axis of length {grid.shape[axis]}")
    other = [np.arange(n) for i, n in enumerate(grid.shape) if i != axis]
    if other:
        mesh = np.meshgrid(*other, indexing="ij")
        cols = iter([m.reshape(-1) for m in mesh])
        n_pts = mesh[0].size
    else:
        cols, n_pts = iter([]), 1
    idx = np.empty((n_pts, grid.ndim), dtype=np.int64)
    for d in range(grid.ndim):
        idx[:, d] = pos if d == axis else next(cols)
    return CWSource(indices=idx, phases=np.full(n_pts, phase, np.float32), amplitude=amplitude,
                    f0=f0, label=f"plane(axis={axis}, pos={pos})")


def bowl_cw_source(grid: Grid, f0: float, amplitude: float, aperture_radius: float, roc: float,
                   apex_vox, spacing: float | None = None, phases=None) -> CWSource:
    """Focused spherical cap on a 3-D grid, axis +z, apex at ``apex_vox``.

    ``phases`` (optional, per voxel of the returned set, or a callable of the
    (n, 3) index array) steers the bowl electronically; the default is uniform
    phase (geometric focus).
    """
    if grid.ndim != 3:
        raise ValueError("bowl_cw_source requires a 3-D grid")
    if len(apex_vox) != 3:
        raise ValueError(f"apex_vox must have 3 entries, got {apex_vox}")
    ds = grid.dx / 2.0 if spacing is None else float(spacing)
    points, _n, _a = spherical_cap_points(aperture_radius, roc, ds)
    idx = np.round(points / grid.dx).astype(np.int64) + np.asarray(apex_vox, np.int64)
    idx = np.unique(idx, axis=0)
    if phases is None:
        ph = np.zeros(idx.shape[0], np.float32)
    elif callable(phases):
        ph = np.asarray(phases(idx), np.float32)
    else:
        ph = np.asarray(phases, np.float32)
    src = CWSource(indices=idx, phases=ph, amplitude=amplitude, f0=f0,
                   label=f"bowl(a={aperture_radius * 1e3:.1f}mm, roc={roc * 1e3:.1f}mm)")
    src.check_inside(grid)
    return src


def focus_phases(indices: np.ndarray, dx: float, target_vox, c: float, f0: float) -> np.ndarray:
    """Delay-and-sum phases that steer a voxelized aperture to ``target_vox``.

    phi_j = k |r_j - r_target| (mod 2 pi) with k = 2 pi f0 / c: with the drive
    sin(omega t - phi_j), every voxel's wave arrives at the target in phase
    (upstream arrays/transducer.py:59-69 gives element phases by the same rule).
    """
    r = (np.asarray(indices, np.float64) - np.asarray(target_vox, np.float64)) * dx
    dist = np.sqrt((r**2).sum(axis=1))
    k = 2.0 * np.pi * f0 / c
    return np.mod(-k * dist, 2.0 * np.pi).astype(np.float32)
