"""Result files in the ``caustica-result/1`` layout (upstream ``src/caustica/io/store.py``).

Layout of one file (upstream store.py:90-170):

* root attributes: format, version, solver, backend, f0_hz, dt_s, spp, steps_total,
  t_end_s, tof_periods, converged_period, settle_capped, harmonics, dx_m, grid_shape,
  pml_vox, region_start, region_stop, phase_convention, absorption_model, plus the
  caller's extra attributes;
* ``input/``: source_indices (int32), source_phases_rad (float32); attributes f0_hz,
  amplitude_pa, ramp_periods, label, n_points;
* ``output/``: p_real_h{n}, p_imag_h{n} for every recorded harmonic and p_max, each
  float16-quantized when the measured error meets the contract (attributes scale,
  stored_dtype, quant_norm_err, reload), gzip level 4, chunked; amplitude and phase
  are derived, never stored;
* ``convergence/history``: (period, peak_pa, rel_change) float64 rows.

``result_tree`` builds that tree as plain Python objects (what is tested without an
HDF5 library); ``save_result`` / ``load_result`` write and read it with h5py, which
must be installed for them (they raise ImportError naming it otherwise). Writes go to
a temporary sibling and are renamed into place, so a file that exists is complete.
"""

from __future__ import annotations

import os
import secrets
from pathlib import Path
from typing import Any

import numpy as np

from caustica_amd.io.quantize import DEFAULT_MAX_NORM_ERR, Quantized, restore, try_float16

RESULT_FORMAT = "caustica-result/1"
FORMAT_VERSION = "caustica_amd-1"

PHASE_CONVENTION = (
    "p(t) = Re{P exp(-i omega t)}; an outgoing wave goes as exp(+ikx), as in the analytic "
    "O'Neil and Rayleigh fields. The absolute phase is not referenced to the source; "
    "relative phase between voxels is exact."
)
ABSORPTION_MODEL = (
    "frequency-independent exponential decay exp(-alpha*c*dt) per step on both p and u, "
    "so the spatial amplitude decay rate is alpha; harmonics above f0 are absorbed at the "
    "f0 rate (less than a power-law tissue would)."
)
_RELOAD = 'arr = ds[:].astype("float32") * ds.attrs["scale"]'


def _h5py():
    try:
        import h5py  # noqa: PLC0415
    except ImportError as exc:  # pragma: no cover - depends on the environment
        raise ImportError("writing or reading caustica-result/1 files needs h5py; "
                          "install h5py or use result_tree() for the data") from exc
    return h5py


def _field(arr, quantize: bool, max_norm_err: float) -> tuple[np.ndarray, dict]:
    q = try_float16(arr, max_norm_err) if quantize else Quantized(
        np.asarray(arr, np.float32), 1.0, "float32", 0.0)
    return q.stored, {"scale": float(q.scale), "stored_dtype": q.dtype_name,
                      "quant_norm_err": float(q.norm_err), "reload": _RELOAD}


def result_tree(result, source, *, dx: float, grid_shape, pml_vox: int = 0,
                quantize: bool = True, max_norm_err: float = DEFAULT_MAX_NORM_ERR,
                extra_attrs: dict[str, Any] | None = None) -> dict:
    """The file's content: {"attrs": {...}, group: {"attrs": {...}, "datasets":
    {name: (array, attrs)}}} for groups input, output, convergence."""
    hs = sorted(int(h) for h in result.phasors)
    attrs = {
        "format": RESULT_FORMAT, "caustica_version": FORMAT_VERSION,
        "solver": str(result.meta.get("solver", "")), "backend": str(result.meta.get("backend", "")),
        "f0_hz": float(source.f0), "dt_s": float(result.dt), "spp": int(result.spp),
        "steps_total": int(result.steps_total), "t_end_s": float(result.t_end_s),
        "tof_periods": int(result.tof_periods), "converged_period": int(result.converged_period),
        "settle_capped": bool(result.settle_capped), "harmonics": hs, "dx_m": float(dx),
        "grid_shape": [int(v) for v in grid_shape], "pml_vox": int(pml_vox),
        "region_start": [int(s.start) for s in result.region],
        "region_stop": [int(s.stop) for s in result.region],
        "phase_convention": PHASE_CONVENTION, "absorption_model": ABSORPTION_MODEL,
    }
    attrs.update(extra_attrs or {})
    out = {}
    for h in hs:
        ph = np.asarray(result.phasors[h])
        out[f"p_real_h{h}"] = _field(ph.real, quantize, max_norm_err)
        out[f"p_imag_h{h}"] = _field(ph.imag, quantize, max_norm_err)
    out["p_max"] = _field(result.p_max, quantize, max_norm_err)
    return {
        "attrs": attrs,
        "input": {
            "attrs": {"f0_hz": float(source.f0), "amplitude_pa": float(source.amplitude),
                      "ramp_periods": float(source.ramp_periods), "label": str(source.label),
                      "n_points": int(source.n_points)},
            "datasets": {
                "source_indices": (np.asarray(source.indices, np.int32), {}),
                "source_phases_rad": (np.asarray(source.phases, np.float32), {}),
            },
        },
        "output": {
            "attrs": {
                "stored": "p_real_h{n}, p_imag_h{n} (phasor of harmonic n), p_max (peak |p| "
                          "over the record window)",
                "derived": "p_amp = hypot(p_real, p_imag), p_phase = atan2(p_imag, p_real); "
                           "recomputed by readers, never stored",
                "phase_convention": PHASE_CONVENTION, "absorption_model": ABSORPTION_MODEL,
                "pmax_window": "peak over the record window only (steady state); settling "
                               "transients are excluded",
            },
            "datasets": out,
        },
        "convergence": {
            "attrs": {"columns": "period, peak_pa, rel_change"},
            "datasets": {"history": (np.asarray(result.convergence_history,
                                                np.float64).reshape(-1, 3), {})},
        },
    }


def tree_to_result(tree: dict):
    """SolverResult from a tree (the inverse of ``result_tree`` up to the recorded
    quantization error)."""
    from caustica_amd.solvers.base import SolverResult

    a = tree["attrs"]
    if a.get("format") != RESULT_FORMAT:
        raise ValueError(f"format {a.get('format')!r} != {RESULT_FORMAT!r}")
    ds = tree["output"]["datasets"]
    phasors = {}
    for h in (int(x) for x in a["harmonics"]):
        re_, rea = ds[f"p_real_h{h}"]
        im_, ima = ds[f"p_imag_h{h}"]
        phasors[h] = (restore(re_, rea["scale"]) + 1j * restore(im_, ima["scale"])).astype(np.complex64)
    pm, pma = ds["p_max"]
    hist = tree["convergence"]["datasets"]["history"][0]
    return SolverResult(
        phasor=phasors[1] if 1 in phasors else phasors[min(phasors)],
        p_max=restore(pm, pma["scale"]),
        region=tuple(slice(int(s), int(e)) for s, e in zip(a["region_start"], a["region_stop"])),
        dt=float(a["dt_s"]), spp=int(a["spp"]), steps_total=int(a["steps_total"]),
        t_end_s=float(a["t_end_s"]), tof_periods=int(a["tof_periods"]),
        converged_period=int(a["converged_period"]), settle_capped=bool(a["settle_capped"]),
        convergence_history=[(int(p), float(k), float(r)) for p, k, r in hist],
        phasors=phasors,
        meta={"solver": str(a["solver"]), "backend": str(a["backend"]),
              "caustica_version": str(a["caustica_version"])},
    )


def save_result(path, result, source, *, dx: float, grid_shape, pml_vox: int = 0,
                quantize: bool = True, max_norm_err: float = DEFAULT_MAX_NORM_ERR,
                extra_attrs: dict[str, Any] | None = None) -> Path:
    """Write one solve as a ``caustica-result/1`` HDF5 file (atomic rename)."""
    h5py = _h5py()
    tree = result_tree(result, source, dx=dx, grid_shape=grid_shape, pml_vox=pml_vox,
                       quantize=quantize, max_norm_err=max_norm_err, extra_attrs=extra_attrs)
    path = Path(path)
    tmp = path.with_name(f"{path.name}.{os.getpid():x}-{secrets.token_hex(3)}.tmp")
    try:
        with h5py.File(tmp, "w") as hf:
            for k, v in tree["attrs"].items():
                hf.attrs[k] = v
            for gname in ("input", "output", "convergence"):
                g = hf.create_group(gname)
                for k, v in tree[gname]["attrs"].items():
                    g.attrs[k] = v
                for name, (arr, dattrs) in tree[gname]["datasets"].items():
                    kw = (dict(chunks=True, compression="gzip", compression_opts=4)
                          if gname == "output" else {})
                    d = g.create_dataset(name, data=arr, **kw)
                    for k, v in dattrs.items():
                        d.attrs[k] = v
        os.replace(tmp, path)
    finally:
        if tmp.exists():
            tmp.unlink()
    return path


def load_result(path):
    """SolverResult from a ``caustica-result/1`` file."""
    h5py = _h5py()
    with h5py.File(path, "r") as hf:
        tree = {"attrs": dict(hf.attrs)}
        for gname in ("input", "output", "convergence"):
            g = hf[gname]
            tree[gname] = {"attrs": dict(g.attrs),
                           "datasets": {n: (g[n][:], dict(g[n].attrs)) for n in g}}
    tree["attrs"]["harmonics"] = [int(h) for h in tree["attrs"]["harmonics"]]
    return tree_to_result(tree)


def load_field(path, name: str) -> np.ndarray:
    """One ``output/`` field, float32, with its scale applied."""
    h5py = _h5py()
    with h5py.File(path, "r") as hf:
        d = hf[f"output/{name}"]
        return restore(d[:], d.attrs["scale"])


__all__ = ["ABSORPTION_MODEL", "PHASE_CONVENTION", "RESULT_FORMAT", "load_field", "load_result",
           "result_tree", "save_result", "tree_to_result"]
