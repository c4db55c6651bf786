"""In-run checkpoint / resume for CW solves (upstream ``src/caustica/io/checkpoint.py``).

A checkpoint is an uncompressed ``.npz`` holding the engine counters, the
convergence history, the device state (p, u0, u1, u2 read back from HBM) and a
fingerprint of everything that shapes the trajectory; loading refuses a file whose
fingerprint differs. Writes go to a temporary name and are renamed into place.
"""

from __future__ import annotations

import hashlib
import json
import os
from collections.abc import Callable
from dataclasses import dataclass, field
from pathlib import Path
from typing import Any

import numpy as np

CHECKPOINT_FORMAT = "caustica-checkpoint/1"
STAGES = ("settle", "record")


class RunInterrupted(RuntimeError):
    """Graceful stop requested through ``CheckpointSpec.stop_when``; state is on disk."""

    def __init__(self, checkpoint_path, stage: str, periods_done: int, steps_done: int):
        super().__init__(
            f"run interrupted on request at period {periods_done} (stage '{stage}', {steps_done} "
            f"steps); state saved to {checkpoint_path} -- rerun the same call to resume."
        )
        self.checkpoint_path = Path(checkpoint_path)
        self.stage = stage
        self.periods_done = periods_done
        self.steps_done = steps_done


class CheckpointMismatch(RuntimeError):
    """The checkpoint on disk belongs to a different run."""


@dataclass
class CheckpointSpec:
    path: Path
    every_periods: int = 8
    stop_when: Callable[[], bool] | None = field(default=None, compare=False)
    keep_on_success: bool = False

    def __post_init__(self) -> None:
        self.path = Path(self.path)
        if self.every_periods < 1:
            raise ValueError(f"every_periods must be >= 1, got {self.every_periods}")


def arrays_sha1(*arrays) -> str:
    h = hashlib.sha1()
    for a in arrays:
        a = np.ascontiguousarray(a)
        h.update(str(a.dtype).encode())
        h.update(str(a.shape).encode())
        h.update(a.tobytes())
    return h.hexdigest()


def _fp_json(fp: dict[str, Any]) -> str:
    return json.dumps(fp, sort_keys=True, separators=(",", ":"))


def write_checkpoint(path, *, stage, n, periods_done, prev_peak, converged, converged_period,
                     history, fields, fingerprint) -> Path:
    if stage not in STAGES:
        raise ValueError(f"stage must be one of {STAGES}, got {stage!r}")
    path = Path(path)
    meta = {
        "format": CHECKPOINT_FORMAT,
        "stage": stage,
        "n": int(n),
        "periods_done": int(periods_done),
        "prev_peak": None if prev_peak is None else float(prev_peak),
        "converged": bool(converged),
        "converged_period": int(converged_period),
        "fingerprint": fingerprint,
    }
    hist = np.asarray(history, dtype=np.float64).reshape(len(history), 3)
    tmp = path.with_name(f".{path.name}.{os.getpid()}.tmp")
    with open(tmp, "wb") as fh:
        np.savez(fh, meta_json=np.asarray(json.dumps(meta)), history=hist,
                 **{f"field_{k}": np.ascontiguousarray(v) for k, v in fields.items()})
    os.replace(tmp, path)
    return path


def load_checkpoint(path, fingerprint: dict[str, Any]) -> dict[str, Any] | None:
    path = Path(path)
    if not path.exists():
        return None
    with np.load(path, allow_pickle=False) as npz:
        meta = json.loads(str(npz["meta_json"]))
        if meta.get("format") != CHECKPOINT_FORMAT:
            raise CheckpointMismatch(f"{path.name}: format {meta.get('format')!r}")
        if _fp_json(meta["fingerprint"]) != _fp_json(fingerprint):
            keys = sorted(set(meta["fingerprint"]) | set(fingerprint))
            diff = [f"  {k}: checkpoint={meta['fingerprint'].get(k)!r} vs run={fingerprint.get(k)!r}"
                    for k in keys if meta["fingerprint"].get(k) != fingerprint.get(k)]
            raise CheckpointMismatch(
                f"{path.name} was written by a different run; refusing to splice. Differing keys:\n"
                + "\n".join(diff)
            )
        hist = npz["history"]
        fields = {k[len("field_"):]: npz[k] for k in npz.files if k.startswith("field_")}
    return {
        "stage": meta["stage"],
        "n": int(meta["n"]),
        "periods_done": int(meta["periods_done"]),
        "prev_peak": meta["prev_peak"],
        "converged": bool(meta["converged"]),
        "converged_period": int(meta["converged_period"]),
        "history": [(int(p), float(pk), float(r)) for p, pk, r in hist],
        "fields": fields,
    }
