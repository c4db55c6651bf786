"""Per-voxel thermal property volumes (upstream ``src/caustica/thermal/properties.py:83-253``).

k [W/m/K], rho [kg/m^3], specific heat [J/kg/K], perfusion w_b [1/s] as float32 host
arrays, plus the grid spacing dx the stencil needs.
"""

from __future__ import annotations

import numpy as np

from caustica_amd.materials import Material


class ThermalPropertyError(ValueError):
    """A material lacks the thermal fields a Pennes solve needs."""


class ThermalMedium:
    def __init__(self, k, rho, specific_heat, perfusion, dx: float, id_map=None):
        vols = {"k": k, "rho": rho, "specific_heat": specific_heat, "perfusion": perfusion}
        shapes = {n: np.shape(v) for n, v in vols.items()}
        if len(set(shapes.values())) != 1:
            raise ValueError(f"thermal property volumes disagree in shape: {shapes}")
        for name, v in vols.items():
            arr = np.ascontiguousarray(v, dtype=np.float32)
            if not np.isfinite(arr).all():
                raise ValueError(f"thermal volume '{name}' contains non-finite values")
            setattr(self, name, arr)
        for name in ("k", "rho", "specific_heat"):
            if (getattr(self, name) <= 0).any():
                raise ValueError(f"thermal volume '{name}' must be > 0 everywhere")
        if (self.perfusion < 0).any():
            raise ValueError("perfusion must be >= 0")
        if not (np.isfinite(dx) and dx > 0):
            raise ValueError(f"dx must be a positive finite spacing in meters, got {dx}")
        self.dx = float(dx)
        self.id_map = id_map

    @classmethod
    def homogeneous(cls, shape, material: Material, dx: float) -> ThermalMedium:
        missing = [f for f in ("thermal_conductivity", "specific_heat", "perfusion_rate")
                   if getattr(material, f) is None]
        if missing:
            raise ThermalPropertyError(f"material {material.name!r} has no {', '.join(missing)}")
        full = np.ones(shape, np.float32)
        return cls(k=full * material.thermal_conductivity, rho=full * material.rho,
                   specific_heat=full * material.specific_heat,
                   perfusion=full * material.perfusion_rate, dx=dx)

    @property
    def shape(self):
        return self.k.shape

    @property
    def ndim(self) -> int:
        return self.k.ndim

    @property
    def volumetric_heat_capacity(self) -> np.ndarray:
        return (self.rho.astype(np.float64) * self.specific_heat).astype(np.float32)

    @property
    def diffusivity(self) -> np.ndarray:
        return (self.k.astype(np.float64) / self.volumetric_heat_capacity).astype(np.float32)

    @property
    def is_perfused(self) -> bool:
        return bool((self.perfusion > 0).any())
