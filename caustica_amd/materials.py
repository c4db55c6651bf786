"""Material constants and unit conversions (upstream ``src/caustica/materials.py:23-140``)."""

from __future__ import annotations

from dataclasses import dataclass

#: 1 dB/cm = 100 / (20 log10 e) Np/m
DB_CM_TO_NP_M = 100.0 / 8.685889638065035


def db_cm_to_np_m(alpha_db_cm: float) -> float:
    return alpha_db_cm * DB_CM_TO_NP_M


def np_m_to_db_cm(alpha_np_m: float) -> float:
    return alpha_np_m / DB_CM_TO_NP_M


def power_law_np_m(alpha0_db_cm_mhz: float, y: float, f0_hz: float) -> float:
    """alpha0 f^y [dB/(cm MHz^y)] evaluated at f0, in Np/m."""
    return db_cm_to_np_m(alpha0_db_cm_mhz * (f0_hz / 1e6) ** y)


def beta_from_bona(bona: float) -> float:
    """Nonlinearity coefficient beta = 1 + B/(2A)."""
    return 1.0 + 0.5 * bona


@dataclass(frozen=True)
class Material:
    """One tissue: acoustic properties at f0 and optional thermal properties."""

    name: str
    alpha_np_m: float
    rho: float
    c: float
    beta: float = 0.0
    thermal_conductivity: float | None = None
    specific_heat: float | None = None
    perfusion_rate: float | None = None

    def __post_init__(self) -> None:
        if self.rho <= 0 or self.c <= 0 or self.alpha_np_m < 0:
            raise ValueError(f"material {self.name!r}: rho, c must be > 0 and alpha >= 0")


def water(alpha_np_m: float = 0.0, c: float = 1500.0, rho: float = 1000.0, beta: float = 0.0,
          name: str = "water") -> Material:
    """Water; defaults are the linear lossless validation medium (upstream materials.py:53-66)."""
    return Material(name=name, alpha_np_m=alpha_np_m, rho=rho, c=c, beta=beta)
