"""Acoustic -> thermal heat source restated on the CPU (test oracle; see oracle/__init__.py).

Restates upstream ``src/caustica/sensors.py:193-295`` (HeatingSource.from_phasors,
Q = sum_n alpha_n |P_n|^2 / (rho c), accumulated in float64 and stored float32)
and ``sensors.py:356-435`` (alpha_n = alpha_f0 * n**y for the power-law model).

Pinning: against the upstream fixture heat_power_law.npz (tests/test_oracle_golden.py)
and against the closed form Q = 2 alpha I for a plane wave (tests/test_oracle_physics.py).
"""

from __future__ import annotations

import numpy as np


def heat_source(phasors: dict, alpha, rho, c, alpha_power_law_y=None) -> np.ndarray:
    harmonics = tuple(sorted(int(n) for n in phasors))
    if any(n >= 2 for n in harmonics) and alpha_power_law_y is None:
        raise ValueError("harmonics above f0 need an absorption model")
    shape = np.shape(phasors[harmonics[0]])
    alpha_f0 = np.broadcast_to(np.asarray(alpha, np.float64), shape)
    inv_rho_c = 1.0 / (
        np.broadcast_to(np.asarray(rho, np.float64), shape)
        * np.broadcast_to(np.asarray(c, np.float64), shape)
    )
    q = np.zeros(shape, np.float64)
    for n in harmonics:
        a_n = alpha_f0 if alpha_power_law_y is None else alpha_f0 * float(n) ** float(alpha_power_law_y)
        p2 = np.abs(np.asarray(phasors[n])).astype(np.float64) ** 2
        q += a_n * p2 * inv_rho_c
    return q
