"""Float16 storage of result fields with a measured error bound.

Contract of upstream ``src/caustica/io/quantize.py:19-59``: a field is divided by its
own peak |value| and cast to float16; the round trip ``stored.astype(float32) *
float32(scale)`` is compared with the original and the largest error, relative to
the peak, is recorded. When it exceeds ``max_norm_err`` (default 1e-3) the field is
kept in float32 with scale 1. An all-zero field is stored as float16 zeros, exactly.
"""

from __future__ import annotations

from typing import NamedTuple

import numpy as np

DEFAULT_MAX_NORM_ERR = 1e-3


class Quantized(NamedTuple):
    stored: np.ndarray   # float16 (scaled by 1/scale) or float32 (as is)
    scale: float
    dtype_name: str      # "float16" | "float32"
    norm_err: float      # max |restored - original| / peak

    def restore(self) -> np.ndarray:
        return restore(self.stored, self.scale)


def restore(stored, scale) -> np.ndarray:
    """The reload rule shared by every reader: float32 cast times float32(scale)."""
    return np.asarray(stored).astype(np.float32) * np.float32(scale)


def try_float16(arr, max_norm_err: float = DEFAULT_MAX_NORM_ERR) -> Quantized:
    a = np.asarray(arr, dtype=np.float32)
    peak = float(np.max(np.abs(a))) if a.size else 0.0
    if not peak > 0.0:
        return Quantized(a.astype(np.float16), 1.0, "float16", 0.0)
    q = (a / peak).astype(np.float16)
    err = float(np.max(np.abs(q.astype(np.float32) * peak - a)) / peak)
    if err <= max_norm_err:
        return Quantized(q, peak, "float16", err)
    return Quantized(a, 1.0, "float32", 0.0)
