"""Pennes bioheat + CEM43 on the hip backend (upstream ``src/caustica/thermal``)."""

from caustica_amd.thermal.dose import cem43_minutes, cem43_rate
from caustica_amd.thermal.pennes import (
    PennesSolver,
    ThermalDivergedError,
    ThermalResult,
    ThermalStabilityError,
)
from caustica_amd.thermal.properties import ThermalMedium, ThermalPropertyError

__all__ = [
    "PennesSolver",
    "ThermalDivergedError",
    "ThermalMedium",
    "ThermalPropertyError",
    "ThermalResult",
    "ThermalStabilityError",
    "cem43_minutes",
    "cem43_rate",
]
