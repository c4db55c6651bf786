"""Load the upstream-written fixtures under tests/golden/ as plain arrays."""

from __future__ import annotations

import json
from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).resolve().parent / "golden"


def load(name: str) -> dict:
    with np.load(GOLDEN / f"{name}.npz", allow_pickle=False) as d:
        out = {k: d[k] for k in d.files}
    if "meta_json" in out:
        out["meta"] = json.loads(bytes(out.pop("meta_json")).decode())
    return out


def cw_case(name: str) -> tuple[dict, dict]:
    """(fixture arrays + meta, solve arguments) for a cw_*.npz fixture."""
    g = load(name)
    m = g["meta"]
    dx = m["dx"]
    pml_vox = int(round(m["pml_thickness"] / dx)) if m["pml_thickness"] is not None else 0
    spec = dict(m["spec"])
    rec = tuple(slice(a, b) for a, b in m["record_region"])
    args = dict(
        shape=tuple(m["shape"]),
        dx=dx,
        pml_vox=pml_vox,
        pml_edge=m["pml_edge"] if m["pml_edge"] is not None else 2.0,
        alpha=g["alpha"],
        rho=g["rho"],
        c=g["c"],
        beta=g["beta"],
        src_idx=g["src_indices"].astype(np.int64),
        src_phase=g["src_phases"],
        amplitude=m["amplitude"],
        f0=m["f0"],
        ramp_periods=m["ramp_periods"],
        spec=spec,
        harmonics=tuple(m["harmonics"]),
        record_region=rec,
        reference_point=tuple(m["reference_point"]) if m["reference_point"] else None,
        nonlinear=m["solver"] == "westervelt",
    )
    return g, args
