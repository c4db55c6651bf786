"""Single-bin DFT phasor contract (upstream ``src/caustica/spectral.py:14-54``).

p(t) = Re{P exp(-i omega t)}: P = (2/N) sum_n p(t_n) exp(+i omega t_n) over an integer
number of periods. The device accumulates exactly this sum during the record window.
"""

from __future__ import annotations

import numpy as np


def single_bin_phasor(series, dt: float, f0: float, t0: float = 0.0, axis: int = -1) -> np.ndarray:
    x = np.asarray(series)
    n = x.shape[axis]
    if n < 2:
        raise ValueError(f"need at least 2 samples, got {n}")
    if dt <= 0 or f0 <= 0:
        raise ValueError("dt and f0 must be > 0")
    n_periods = n * dt * f0
    if abs(n_periods - round(n_periods)) > 1e-6 * max(1.0, n_periods):
        raise ValueError(f"window covers {n_periods:.6f} periods -- not an integer; the single-bin DFT would leak.")
    t = t0 + dt * np.arange(n)
    kernel = np.exp(+1j * 2.0 * np.pi * f0 * t)
    shape = [1] * x.ndim
    shape[axis] = n
    return (x * kernel.reshape(shape)).sum(axis=axis) * (2.0 / n)
