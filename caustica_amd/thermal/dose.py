"""CEM43 thermal dose (Sapareto-Dean) -- upstream ``src/caustica/thermal/dose.py``.

t_CEM43 = integral R^(43 - T(t)) dt in equivalent minutes, R = 0.5 above 43 C and 0.25
at or below (at T = 43 both give 1). The Pennes solve accumulates it on the device per
sub-step with the trapezoid rule; these helpers are the closed forms.
"""

from __future__ import annotations

import numpy as np

CEM43_REFERENCE_C = 43.0
CEM43_R_ABOVE = 0.5
CEM43_R_BELOW = 0.25
ITRUSST_CEM43_LIMITS = {"brain": 2.0, "bone": 16.0, "skin": 21.0}
ITRUSST_DELTA_T_LIMIT_C = 2.0


def cem43_rate(temperature_c):
    """R^(43 - T): equivalent minutes accrued per minute at T (dtype of T kept)."""
    t = np.asarray(temperature_c)
    if t.dtype.kind != "f":
        t = t.astype(np.float64)
    s = t.dtype.type
    r = np.where(t > CEM43_REFERENCE_C, s(CEM43_R_ABOVE), s(CEM43_R_BELOW))
    return r ** (s(CEM43_REFERENCE_C) - t)


def cem43_minutes(temperature_c, seconds: float):
    """Dose of a constant temperature held for ``seconds`` (43 C for 60 s = 1 min)."""
    if seconds < 0:
        raise ValueError(f"seconds must be >= 0, got {seconds}")
    return cem43_rate(temperature_c) * (seconds / 60.0)
