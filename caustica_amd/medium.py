"""Dense per-voxel acoustic property volumes (upstream ``src/caustica/medium.py:23-106``).

alpha [Np/m], rho [kg/m^3], c [m/s], beta [1] as float32 C-contiguous host arrays;
a solve uploads them to HBM once.
"""

from __future__ import annotations

import numpy as np

from caustica_amd.materials import Material


class Medium:
    def __init__(self, alpha, rho, c, beta, id_map=None):
        vols = {"alpha": alpha, "rho": rho, "c": c, "beta": beta}
        shapes = {name: np.shape(v) for name, v in vols.items()}
        if len(set(shapes.values())) != 1:
            raise ValueError(f"property volumes disagree in shape: {shapes}")
        for name, v in vols.items():
            setattr(self, name, np.ascontiguousarray(v, dtype=np.float32))
        if id_map is not None and np.shape(id_map) != self.alpha.shape:
            raise ValueError(f"id_map shape {np.shape(id_map)} != property shape {self.alpha.shape}")
        self.id_map = id_map

    @classmethod
    def homogeneous(cls, shape, material: Material) -> Medium:
        full = np.full(shape, 1.0, dtype=np.float32)
        return cls(
            alpha=full * material.alpha_np_m,
            rho=full * material.rho,
            c=full * material.c,
            beta=full * material.beta,
        )

    @classmethod
    def from_id_map(cls, id_map, materials: dict[int, Material]) -> Medium:
        id_map = np.asarray(id_map)
        if not np.issubdtype(id_map.dtype, np.integer):
            raise TypeError(f"id_map must be integer-typed, got {id_map.dtype}")
        present = [int(i) for i in np.unique(id_map)]
        unknown = [i for i in present if i not in materials]
        if unknown:
            raise ValueError(f"id_map contains ids {unknown} missing from the material table")
        out = {k: np.zeros(id_map.shape, np.float32) for k in ("alpha", "rho", "c", "beta")}
        for i in present:
            m = materials[i]
            mask = id_map == i
            out["alpha"][mask] = m.alpha_np_m
            out["rho"][mask] = m.rho
            out["c"][mask] = m.c
            out["beta"][mask] = m.beta
        return cls(**out, id_map=id_map)

    @property
    def shape(self):
        return self.alpha.shape

    @property
    def c_min(self) -> float:
        return float(self.c.min())

    @property
    def c_max(self) -> float:
        return float(self.c.max())

    @property
    def is_linear(self) -> bool:
        return bool((self.beta == 0).all())

    def __repr__(self) -> str:  # pragma: no cover - cosmetic
        return f"Medium(shape={self.shape}, c=[{self.c_min:.0f},{self.c_max:.0f}] m/s, linear={self.is_linear})"
