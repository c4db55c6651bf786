"""Planner: VRAM inventory, step-time model and device table, with a hip entry.

Same interface as upstream ``src/caustica/planner`` (``estimate`` __init__.py:281-477,
``Estimate``/``summary`` :236-270, ``compare`` :504-516, ``GPUSpec`` :104-128,
``kspace_memory`` model.py:86-115, ``db_time_coeffs``/``step_time`` model.py:138-154,
device table gpu_db.json) so a plan-first runner keeps its gates. What differs on the
hip target (the MI355X row, or any row with a measured ``bw_eff``):

* memory is the library's own carve of one member (``hip_memory``: one allocation in
  256-byte aligned pieces -- p, seven padded half spectra (four work spectra and the
  velocity state), coefficient maps, 1-D k and sponge tables, source CSR, record sums,
  convergence peak slots), exact to the byte; the
  upstream inventory of the same solve is reported beside it
  (``vram_bytes_reference_model``);
* time is a bandwidth model: each of the eight passes of a step is HBM-bound, so a step
  costs the bytes it moves (``step_bytes``) over the sustained rate, datasheet bandwidth
  times ``bw_eff`` until ``calibrate`` measures the rate of that padded shape on the card.

Rows where hip does not run keep the upstream (a, b) model and inventory, for comparison.
"""

from __future__ import annotations

import json
import time
from dataclasses import dataclass
from math import ceil, log2, prod
from pathlib import Path

import numpy as np

from caustica_amd.solvers.kspace.operators import (
    MemberPlan,
    cw_discretization,
    cw_tof_periods,
    normalize_record_region,
    pad_shape,
)

_GIB = 2**30
_F32, _C64 = 4, 8

#: Flat margin of the upstream inventory (model.py:26).
ALLOCATOR_MARGIN = 1.15
#: Device memory held before any solve allocates (upstream model.py:30: CUDA context and
#: runtime); the same figure is kept for the hip row (HIP runtime, torch, code objects).
CONTEXT_OVERHEAD_BYTES = int(1.2e9)
#: Upstream db time model efficiencies (model.py:35-39) and one-time warm-up (:59).
FFT_FLOP_EFF = 0.10
BW_EFF = 0.75
GPU_WARMUP_S = 3.0
#: hip one-time cost per solve before the step clock: session tables, one member upload
#: and the first launches (no plan creation or JIT: kernels are prebuilt).
HIP_WARMUP_S = 0.05

_DB_PATH = Path(__file__).with_name("gpu_db.json")


# ----------------------------------------------------------------- devices
@dataclass(frozen=True)
class GPUSpec:
    key: str
    vram_gib: float
    mem_bw_gbs: float
    fp32_tflops: float
    bw_eff: float | None = None  # sustained fraction of mem_bw_gbs on hip; None: no hip
    arch: str = ""
    notes: str = ""

    @property
    def vram_bytes(self) -> int:
        return int(self.vram_gib * _GIB)

    @property
    def usable_bytes(self) -> int:
        return max(0, self.vram_bytes - CONTEXT_OVERHEAD_BYTES)

    @property
    def runs_hip(self) -> bool:
        return self.bw_eff is not None


def load_gpu_db() -> tuple[dict, dict]:
    raw = json.loads(_DB_PATH.read_text())
    devs = {k: GPUSpec(key=k, **v) for k, v in raw["devices"].items()}
    return devs, dict(raw.get("aliases", {}))


def _resolve_gpu(gpu) -> GPUSpec:
    if isinstance(gpu, GPUSpec):
        return gpu
    devs, aliases = load_gpu_db()
    key = aliases.get(gpu, gpu)
    if key not in devs:
        raise ValueError(f"unknown gpu {gpu!r}; known: {', '.join(sorted(devs))}")
    return devs[key]


# ----------------------------------------------------------------- memory
@dataclass(frozen=True)
class MemoryModel:
    breakdown: dict
    total_bytes: int
    padded_shape: tuple


def fft_sizes(active_shape) -> tuple:
    padded = pad_shape(tuple(active_shape))
    return padded, prod(padded), prod(padded[:-1]) * (padded[-1] // 2 + 1)


def kspace_memory(active_shape, nonlinear: bool, n_harmonics: int,
                  rec_elems: int | None = None) -> MemoryModel:
    """Upstream inventory of one CW solve (model.py:86-115): dense state, maps, sponge
    volume, spectral factor volumes, record buffers, step temporaries, FFT workspace."""
    padded, p, r = fft_sizes(active_shape)
    nd = len(padded)
    rec = int(rec_elems) if rec_elems is not None else prod(active_shape)
    b = {
        "state (p + u)": (1 + nd) * _F32 * p,
        "property maps": (3 + (1 if nonlinear else 0)) * _F32 * p,
        "sponge": _F32 * p,
        "spectral factors (i*k*kappa)": nd * _C64 * r,
        "record buffers": rec * (_C64 * n_harmonics + _F32),
        "step temporaries": 3 * _C64 * r + (2 + (2 if nonlinear else 0)) * _F32 * p,
        "fft workspace": 2 * _C64 * r,
    }
    return MemoryModel(b, ceil(sum(b.values()) * ALLOCATOR_MARGIN), padded)


def _al(n: int) -> int:
    return (max(int(n), 0) + 255) // 256 * 256


def ramp_steps(dt: float, f0: float, ramp_periods: float) -> int:
    """Steps before the cosine ramp completes (the length of the library's drive-amplitude
    table): the n with n dt / (ramp_periods / f0) < 1, evaluated as the library does."""
    period = 1.0 / f0
    n = 0
    while min(n * dt / (ramp_periods * period), 1.0) < 1.0:
        n += 1
    return n


def hip_inventory(padded3, *, nonlinear: bool, n_harmonics: int, rec_elems: int,
                  n_src: int, max_periods: int, spp: int = 1, ramp_n: int = 0,
                  staggered_maps: int = 0) -> MemoryModel:
    """Bytes of one member on the hip engine, piece by piece as the library carves them.
    ``staggered_maps`` > 0: a member of the opt-in staggered scheme with that many
    coefficient maps (5, + 2 absorbing, + 2 nonlinear), its split density and its PML and
    shift-factor tables in place of the reference maps and sponge."""
    nx, ny, nz = (int(v) for v in padded3)
    nzh = nz // 2 + 1
    ms = (nz // 2 + 15) // 16 * 16
    P = nx * ny * nz
    H = nx * (ny * ms + 16) + (nx * ny + 15) // 16 * 16  # rows + Nyquist side plane
    ns = max(int(n_src), 1)
    if staggered_maps:
        pst = (nz + nx + ny + 63) // 64 * 64
        scheme = {
            "split density (rho_0-2)": _al(3 * P * 4),
            "property maps": _al(int(staggered_maps) * P * 4),
            "PML tables": _al(4 * pst * 4),
            "shift factors": _al((nx + ny + nzh) * 8),
        }
    else:
        scheme = {
            "property maps": _al((4 if nonlinear else 3) * P * 4),
            "sponge profiles": _al((nz + nx + ny) * 4),
        }
    b = {
        "pressure (p)": _al(P * 4),
        "half spectra (sp, sg0-2, velocity state su0-2)": _al(7 * H * 8),
        **scheme,
        "k tables": _al(2 * (nx + ny + nzh) * 4),
        "source table": _al((nx * ny + 1) * 4) + _al(ns * 4) + _al(ns * 8),
        "record buffers": _al(2 * n_harmonics * rec_elems * 8) + _al(rec_elems * 4),
        "convergence peaks": _al(max(int(max_periods), 1) * 64 * 16 * 4),
        # clock (16 B) and the step scalars 64 B after it (48 B) share one slot
        "step clock and drive tables": _al(64 + 48) + _al(max(int(ramp_n), 1) * 4) + _al(int(spp) * 8),
    }
    return MemoryModel(b, int(sum(b.values())), tuple(padded3))


def hip_memory(plan: MemberPlan, max_periods: int | None = None) -> MemoryModel:
    """``hip_inventory`` of a planned member (the same ``max_periods`` default as
    ``KSpaceSession.add``)."""
    nmaps = sum(m is not None for m in plan.stg["maps"]) if plan.scheme == "staggered" else 0
    return hip_inventory(plan.padded3, nonlinear=plan.beta2_dt is not None,
                         n_harmonics=len(plan.harmonics), rec_elems=int(np.prod(plan.rec_shape)),
                         n_src=int(plan.src_idx3.shape[0]),
                         max_periods=plan.eff_max + 4 if max_periods is None else max_periods,
                         spp=plan.spp, ramp_n=ramp_steps(plan.dt, plan.f0, plan.ramp_periods),
                         staggered_maps=nmaps)


# ------------------------------------------------------------------- time
def step_bytes(padded3, nonlinear: bool, n_harmonics: int = 1, recording: bool = False) -> int:
    """HBM bytes one hip step moves: 248 H + 24 P (+4 P nonlinear; +(32 n_h + 8) P on a
    record step), H = nx ny (nz/2 + 1) spectrum bins, P = nx ny nz voxels."""
    nx, ny, nz = (int(v) for v in padded3)
    P = nx * ny * nz
    H = nx * ny * (nz // 2 + 1)
    b = 248 * H + 24 * P + (4 * P if nonlinear else 0)
    if recording:
        b += (32 * n_harmonics + 8) * P
    return int(b)


def db_time_coeffs(fp32_tflops: float, mem_bw_gbs: float, nd: int, nonlinear: bool):
    """Upstream (a, b) of ``t_step = a P log2 P + b P`` (model.py:118-149)."""
    n_ffts = 2 + 2 * nd
    a = 2.5 * n_ffts / (fp32_tflops * 1e12 * FFT_FLOP_EFF)
    pointwise = _F32 * (10.0 * (nd + 1) + (6.0 if nonlinear else 0.0)) + 6.0 * nd * _C64 * 0.5
    b = pointwise / (mem_bw_gbs * 1e9 * BW_EFF)
    return a, b


def step_time(a: float, b: float, p_elems: int) -> float:
    return a * p_elems * log2(p_elems) + b * p_elems


# --------------------------------------------------------------- estimate
@dataclass(frozen=True)
class Estimate:
    solver: str
    gpu: str
    source: str                # "db" | "calibrated"
    vram_bytes: int
    vram_breakdown: dict
    vram_usable_bytes: int
    fits: bool
    dt: float
    spp: int
    t_step_s: float
    warmup_s: float
    steps_expected: int
    steps_worst: int
    t_expected_s: float
    t_worst_s: float
    warnings: tuple
    advice: tuple
    padded_shape: tuple = ()
    vram_bytes_reference_model: int = 0

    @property
    def vram_gib(self) -> float:
        return self.vram_bytes / _GIB

    def summary(self) -> str:
        lines = [
            f"planner [{self.source}] {self.solver} on {self.gpu}: "
            f"{'FITS' if self.fits else 'DOES NOT FIT'} "
            f"({self.vram_gib:.2f} GiB needed / {self.vram_usable_bytes / _GIB:.2f} GiB usable)",
            f"  t_step={self.t_step_s * 1e3:.3f} ms, spp={self.spp}, warmup={self.warmup_s:.2f} s, "
            f"expected={_fmt_t(self.t_expected_s)} ({self.steps_expected} steps), "
            f"worst-case={_fmt_t(self.t_worst_s)}",
        ]
        lines += [f"  ! {w}" for w in self.warnings]
        lines += [f"  -> {a}" for a in self.advice]
        return "\n".join(lines)


def _fmt_t(s: float) -> str:
    if s < 120.0:
        return f"{s:.2f} s"
    if s < 7200.0:
        return f"{s / 60.0:.1f} min"
    return f"{s / 3600.0:.2f} h"


def estimate(grid, medium, source, spec=None, *, solver: str = "westervelt",
             gpu="MI355X", harmonics=(1,), record_region=None, reference_point=None,
             calibration_path=None) -> Estimate:
    """VRAM and wall time of one CW solve without running it: the engine's own dt/spp,
    the settle policy (expected: time of flight + settle floor; worst: the cap; both
    raised to the t_end floor), and the inventory/time model of the target device."""
    from caustica_amd.solvers.base import CWRunSpec

    if solver not in ("linear", "westervelt"):
        raise ValueError(f"planner models the k-space engine ('linear'/'westervelt'), not {solver!r}")
    spec = spec if spec is not None else CWRunSpec()
    harmonics = tuple(dict.fromkeys(int(h) for h in harmonics))
    nonlinear = solver == "westervelt" and not medium.is_linear
    dev = _resolve_gpu(gpu)
    spp, dt = cw_discretization(grid.dx, medium.c_max, spec, source.f0, harmonics)
    tof = cw_tof_periods(grid.shape, grid.dx, medium.c_min, source.indices, source.f0,
                         reference_point)
    rec = (normalize_record_region(record_region, grid.shape) if record_region is not None
           else tuple(slice(0, n) for n in grid.shape))
    rec_elems = prod(s.stop - s.start for s in rec)

    floor_ = max(spec.min_settle_periods, ceil(source.ramp_periods) + 1)
    pre_exp = tof + floor_
    pre_worst = max(tof + spec.max_settle_periods, pre_exp)
    if spec.t_end_min_us is not None:
        need = ceil(spec.t_end_min_us * 1e-6 * source.f0) - spec.n_record_periods
        pre_exp, pre_worst = max(pre_exp, need), max(pre_worst, need)
    steps_exp = (pre_exp + spec.n_record_periods) * spp
    steps_worst = (pre_worst + spec.n_record_periods) * spp
    rec_steps = spec.n_record_periods * spp

    ref_mem = kspace_memory(grid.shape, nonlinear, len(harmonics), rec_elems)
    padded = ref_mem.padded_shape
    padded3 = (1,) * (3 - len(padded)) + tuple(padded)
    warnings: list = []
    if dev.runs_hip:
        mem = hip_inventory(padded3, nonlinear=nonlinear, n_harmonics=len(harmonics),
                            rec_elems=rec_elems, n_src=int(source.indices.shape[0]),
                            max_periods=pre_worst + spec.n_record_periods + 4, spp=spp,
                            ramp_n=ramp_steps(dt, source.f0, source.ramp_periods))
        entry = find_calibration_for(dev.key, padded3, nonlinear, calibration_path)
        if entry is not None:
            rate, warm, label = entry["bytes_per_s"], entry["warmup_s"], "calibrated"
        else:
            rate, warm, label = dev.mem_bw_gbs * 1e9 * dev.bw_eff, HIP_WARMUP_S, "db"
            warnings.append(f"db estimate: {100 * dev.bw_eff:.0f}% of datasheet bandwidth "
                            f"assumed; planner.calibrate() measures this shape")
        t_step = step_bytes(padded3, nonlinear) / rate
        t_rec = step_bytes(padded3, nonlinear, len(harmonics), recording=True) / rate

        def total(n):
            return warm + (n - rec_steps) * t_step + rec_steps * t_rec
    else:
        mem = ref_mem
        a, b = db_time_coeffs(dev.fp32_tflops, dev.mem_bw_gbs, len(padded), nonlinear)
        t_step, warm, label = step_time(a, b, prod(padded)), GPU_WARMUP_S, "db"
        warnings.append(f"{dev.key} does not run the hip engine: upstream inventory and "
                        f"time model shown for comparison")

        def total(n):
            return warm + n * t_step

    fits = mem.total_bytes <= dev.usable_bytes
    advice: list = []
    if not fits:
        m = (mem.total_bytes / max(dev.usable_bytes, 1)) ** (1.0 / grid.ndim)
        advice.append(f"increase dx by >= x{m:.2f} (grid ~"
                      f"{tuple(max(8, int(n / m)) for n in grid.shape)})")
        if mem.breakdown["record buffers"] > 0.10 * mem.total_bytes:
            advice.append("shrink the record region: pass record_region=... to run()")
        if len(harmonics) > 1:
            advice.append("drop harmonics you do not need (each costs 16 B per recorded voxel)")
        if nonlinear:
            advice.append("the 'linear' solver drops the beta map (4 B per voxel)")
        bigger = [k for k, s in load_gpu_db()[0].items()
                  if s.key != dev.key and s.usable_bytes >= mem.total_bytes]
        if bigger:
            advice.append(f"or pick a larger device: {', '.join(sorted(bigger))}")
    return Estimate(solver=solver, gpu=dev.key, source=label, vram_bytes=mem.total_bytes,
                    vram_breakdown=dict(mem.breakdown), vram_usable_bytes=dev.usable_bytes,
                    fits=fits, dt=dt, spp=spp, t_step_s=t_step, warmup_s=warm,
                    steps_expected=steps_exp, steps_worst=steps_worst,
                    t_expected_s=total(steps_exp), t_worst_s=total(steps_worst),
                    warnings=tuple(warnings), advice=tuple(advice), padded_shape=padded,
                    vram_bytes_reference_model=ref_mem.total_bytes)


@dataclass(frozen=True)
class Comparison:
    estimates: tuple

    def table(self) -> str:
        head = (f"{'GPU':<10} {'fits':<5} {'VRAM GiB':>9} {'usable':>7} {'t_step':>10} "
                f"{'expected':>10} {'worst':>10} {'src':<10}")
        rows = [head, "-" * len(head)]
        for e in self.estimates:
            rows.append(f"{e.gpu:<10} {'yes' if e.fits else 'NO':<5} {e.vram_gib:>9.2f} "
                        f"{e.vram_usable_bytes / _GIB:>7.1f} {e.t_step_s * 1e3:>8.3f}ms "
                        f"{_fmt_t(e.t_expected_s):>10} {_fmt_t(e.t_worst_s):>10} {e.source:<10}")
        return "\n".join(rows)

    def __str__(self) -> str:
        return self.table()


def compare(grid, medium, source, spec=None, *, gpus=None, **kw) -> Comparison:
    """``estimate`` across devices, fits first, then fastest first."""
    keys = tuple(gpus) if gpus is not None else tuple(sorted(load_gpu_db()[0]))
    ests = sorted((estimate(grid, medium, source, spec, gpu=k, **kw) for k in keys),
                  key=lambda e: (not e.fits, e.t_expected_s))
    return Comparison(tuple(ests))


# ------------------------------------------------------------- calibration
def _cal_key(padded3, nonlinear: bool) -> str:
    return "x".join(str(int(v)) for v in padded3) + (":nl" if nonlinear else ":lin")


def find_calibration_for(gpu_key: str, padded3, nonlinear: bool, path=None):
    """The calibration entry of this device and padded shape, or None."""
    if path is None:
        return None
    p = Path(path)
    if not p.exists():
        return None
    data = json.loads(p.read_text())
    return data.get("devices", {}).get(gpu_key, {}).get(_cal_key(padded3, nonlinear))


def calibrate(shapes=((256, 256, 256),), *, nonlinear: bool = False, steps: int = 40,
              path=None, gpu: str = "MI355X", device: int = 0) -> dict:
    """Measure the hip step rate of each padded 3-D shape on ``device`` (a homogeneous
    water member, ``steps`` steps after warm-up) and the one-time cost of opening a
    session and adding the member; write them to ``path`` (JSON) if given."""
    from caustica_amd.solvers.kspace.native import KSpaceSession
    from caustica_amd.solvers.kspace.operators import k_tables, sponge_tables

    dev = _resolve_gpu(gpu)
    out = {}
    for shape in shapes:
        shape = tuple(int(v) for v in shape)
        plan = _synthetic_plan(shape, nonlinear, k_tables, sponge_tables)
        t0 = time.perf_counter()
        with KSpaceSession(shape, 1, device) as s:
            slot = s.add(plan, max_periods=8)
            s.advance([slot], 1)
            s.sync()
            warm = time.perf_counter() - t0
            s.advance([slot], 4)
            s.sync()
            t1 = time.perf_counter()
            s.advance([slot], steps)
            s.sync()
            t_step = (time.perf_counter() - t1) / steps
        out[_cal_key(shape, nonlinear)] = {
            "bytes_per_s": step_bytes(shape, nonlinear) / t_step, "t_step_s": t_step,
            "warmup_s": warm, "steps": steps, "calibrated_at": time.strftime("%Y-%m-%dT%H:%M:%S")}
    if path is not None:
        p = Path(path)
        data = json.loads(p.read_text()) if p.exists() else {"version": 1, "devices": {}}
        data.setdefault("devices", {}).setdefault(dev.key, {}).update(out)
        p.write_text(json.dumps(data, indent=1, sort_keys=True))
    return out


def _synthetic_plan(shape, nonlinear, k_tables, sponge_tables) -> MemberPlan:
    """A homogeneous-water member of a padded 3-D shape for timing (the step cost does
    not depend on the values; one source voxel at the centre)."""
    dx, c, rho, f0 = 2e-4, 1500.0, 1000.0, 1e6
    spp = 10
    dt = 1.0 / f0 / spp
    P = prod(shape)
    full = tuple(slice(0, n) for n in shape)
    box = tuple(v for n in shape for v in (0, n))
    return MemberPlan(
        shape=shape, padded=shape, padded3=shape, dx=dx, pml_vox=4, pml_edge=2.0, spp=spp,
        dt=dt, f0=f0, c_ref=c, amplitude=1e5, ramp_periods=2.0,
        dt_over_rho=np.full(P, dt / rho, np.float32).reshape(shape),
        rhoc2_dt=np.full(P, rho * c * c * dt, np.float32).reshape(shape),
        absorb=np.ones(shape, np.float32),
        beta2_dt=np.full(shape, 7.0 * dt, np.float32) if nonlinear else None,
        ktab=k_tables(shape, dx), sponge=sponge_tables(shape, 3, 4, 2.0),
        src_idx3=np.array([[n // 2 for n in shape]], np.int32),
        src_phase=np.zeros(1, np.float32), src_scale=np.full(1, 2 * c * dt / dx, np.float32),
        conv=full, rec=full, conv3=box, rec3=box, harmonics=(1,), tof_periods=1,
        eff_min=4, eff_max=4)


__all__ = ["ALLOCATOR_MARGIN", "Comparison", "Estimate", "GPUSpec", "MemoryModel", "calibrate",
           "compare", "db_time_coeffs", "estimate", "find_calibration_for", "fft_sizes",
           "hip_inventory", "hip_memory", "kspace_memory", "load_gpu_db", "step_bytes",
           "step_time"]
