"""Pennes bioheat solver with CEM43 dose on the hip backend.

Upstream ``src/caustica/thermal/pennes.py``: same equation, discretization, API and
guards -- explicit conservative finite differences with harmonic-mean face
conductivities, insulated or Dirichlet walls, forward Euler on theta = T - T_ref,
positivity-bound stability (refuse or sub-step), theta_max over every sub-step, CEM43
accumulated per sub-step with the trapezoid rule, finite guard on a cadence,
``ThermalResult.chain`` for multi-phase exposures. The coefficient assembly (float32,
upstream dtypes) runs once on the host; the step loop is the fused kernel in
libcaustica_hip (cst_thermal_*), with the dose sum carried in float64 on the device.
"""

from __future__ import annotations

import ctypes as C
import logging
from collections.abc import Sequence
from dataclasses import dataclass, field
from math import ceil
from typing import Any

import numpy as np

from caustica_amd import _native as N
from caustica_amd.core.backend import get_backend
from caustica_amd.sensors import DeviceHeatingSource, HeatingSource
from caustica_amd.thermal.properties import ThermalMedium

log = logging.getLogger("caustica")

BLOOD_DENSITY = 1050.0
BLOOD_SPECIFIC_HEAT = 3617.0
ARTERIAL_TEMPERATURE_C = 37.0
SCHEME = "pennes-fd-explicit/1"

_ALL = slice(None)
_LOWER = slice(None, -1)
_UPPER = slice(1, None)


class ThermalDivergedError(RuntimeError):
    """The temperature (or the dose) stopped being a number."""


class ThermalStabilityError(ValueError):
    """The requested time step is above the explicit scheme's stability bound."""


def _ax(arr, axis, sl):
    idx = [_ALL] * arr.ndim
    idx[axis] = sl
    return arr[tuple(idx)]


@dataclass
class ThermalResult:
    temperature: np.ndarray
    temperature_max: np.ndarray
    dose_cem43: np.ndarray | None
    samples: list
    times: list
    dt: float
    n_steps: int
    t_end_s: float
    substeps: int
    meta: dict[str, Any] = field(default_factory=dict)

    @property
    def peak_temperature_c(self) -> float:
        return float(self.temperature_max.max())

    @property
    def peak_dose_cem43(self) -> float | None:
        return None if self.dose_cem43 is None else float(self.dose_cem43.max())

    @staticmethod
    def chain(results: Sequence[ThermalResult]) -> ThermalResult:
        """One result for a multi-phase exposure: elementwise max of temperature_max,
        the final temperature, the last phase's dose (refused if any later phase's dose
        map is below an earlier one's -- the phases were then not chained via dose0)."""
        results = list(results)
        if not results:
            raise ValueError("ThermalResult.chain() needs at least one result")
        if len({r.temperature.shape for r in results}) != 1:
            raise ValueError("the phases are not on one grid")
        dosed = [r.dose_cem43 is not None for r in results]
        if any(dosed) and not all(dosed):
            raise ValueError("some phases accumulated no CEM43; solve every phase with dose=True, "
                             "passing the previous phase's dose_cem43 as dose0.")
        for i, (prev, nxt) in enumerate(zip(results, results[1:])):
            if prev.dose_cem43 is None:
                continue
            slack = 1e-6 * np.maximum(1.0, prev.dose_cem43)
            if not bool((nxt.dose_cem43 >= prev.dose_cem43 - slack).all()):
                raise ValueError(f"phase {i + 1}'s CEM43 map is below phase {i}'s somewhere: the "
                                 f"phases were not chained (pass dose0=<previous>.dose_cem43).")
        tmax = results[0].temperature_max.copy()
        for r in results[1:]:
            np.maximum(tmax, r.temperature_max, out=tmax)
        samples, times, offset = [], [], 0.0
        for i, r in enumerate(results):
            for t, s in zip(r.times, r.samples):
                if i > 0 and t == 0.0:
                    continue
                times.append(offset + t)
                samples.append(s)
            offset += r.t_end_s
        last = results[-1]
        meta = dict(last.meta)
        meta["q"] = " -> ".join(str(r.meta.get("q")) for r in results)
        meta["chained_phases"] = [{"t_end_s": r.t_end_s, "n_steps": r.n_steps, "dt_s": r.dt}
                                  for r in results]
        return ThermalResult(temperature=last.temperature, temperature_max=tmax,
                             dose_cem43=last.dose_cem43, samples=samples, times=times, dt=last.dt,
                             n_steps=sum(r.n_steps for r in results),
                             t_end_s=float(sum(r.t_end_s for r in results)),
                             substeps=max(r.substeps for r in results), meta=meta)


class PennesSolver:
    name = "pennes"
    scheme = SCHEME

    def __init__(self, backend: str = "auto", *, blood_density: float = BLOOD_DENSITY,
                 blood_specific_heat: float = BLOOD_SPECIFIC_HEAT,
                 arterial_temperature_c: float = ARTERIAL_TEMPERATURE_C,
                 boundary: str = "insulated", boundary_temperature_c: float | None = None,
                 reference_temperature_c: float | None = None, on_unstable: str = "refuse",
                 finite_check_every: int = 25, profile: bool = False, x_chunk: int = 0,
                 uniform_maps: bool = True):
        if boundary not in ("insulated", "dirichlet"):
            raise ValueError(f"boundary must be 'insulated' or 'dirichlet', got {boundary!r}")
        if on_unstable not in ("refuse", "substep"):
            raise ValueError(f"on_unstable must be 'refuse' or 'substep', got {on_unstable!r}")
        if blood_density <= 0 or blood_specific_heat <= 0:
            raise ValueError("blood_density and blood_specific_heat must be > 0")
        if finite_check_every < 1:
            raise ValueError(f"finite_check_every must be >= 1, got {finite_check_every}")
        self.backend = backend
        self.blood_density = float(blood_density)
        self.blood_specific_heat = float(blood_specific_heat)
        self.arterial_temperature_c = float(arterial_temperature_c)
        self.boundary = boundary
        self.boundary_temperature_c = (float(arterial_temperature_c) if boundary_temperature_c is None
                                       else float(boundary_temperature_c))
        self.reference_temperature_c = (float(arterial_temperature_c) if reference_temperature_c is None
                                        else float(reference_temperature_c))
        self.on_unstable = on_unstable
        self.finite_check_every = int(finite_check_every)
        # record the stencil kernel's device time (HIP events) into result.meta
        self.profile = bool(profile)
        # x planes per block of the two-step kernel (0: chosen per grid); results do not
        # depend on it
        self.x_chunk = int(x_chunk)
        # uniform k, perfusion and dt/(rho C) maps run kernels that keep the three values
        # in registers (same bits); False reads the maps regardless
        self.uniform_maps = bool(uniform_maps)

    # ------------------------------------------------------- host assembly
    def _faces(self, k, dx):
        inv_dx2 = 1.0 / (dx * dx)
        return [((2.0 * _ax(k, ax, _LOWER) * _ax(k, ax, _UPPER) /
                  (_ax(k, ax, _LOWER) + _ax(k, ax, _UPPER))) * inv_dx2).astype(np.float32)
                for ax in range(k.ndim)]

    def _perf(self, medium):
        return (medium.perfusion * np.float32(self.blood_density * self.blood_specific_heat)).astype(np.float32)

    def _inv_rho_c(self, medium):
        return np.asarray(1.0 / medium.volumetric_heat_capacity.astype(np.float64), np.float32)

    def stable_dt(self, medium: ThermalMedium) -> float:
        """1 / max_i (sum of face coefficients + w_b rho_b C_b)_i / (rho C)_i (upstream :376-402)."""
        k = medium.k
        inv_dx2 = 1.0 / (medium.dx * medium.dx)
        diag = np.zeros(k.shape, np.float32)
        for ax, kf in enumerate(self._faces(k, medium.dx)):
            _ax(diag, ax, _LOWER)[...] += kf
            _ax(diag, ax, _UPPER)[...] += kf
        if self.boundary == "dirichlet":
            for ax in range(k.ndim):
                _ax(diag, ax, slice(0, 1))[...] += (2.0 * _ax(k, ax, slice(0, 1)) * inv_dx2).astype(np.float32)
                _ax(diag, ax, slice(-1, None))[...] += (2.0 * _ax(k, ax, slice(-1, None)) * inv_dx2).astype(np.float32)
        diag += self._perf(medium)
        diag *= self._inv_rho_c(medium)
        peak = float(diag.max())
        return float("inf") if peak <= 0.0 else 1.0 / peak

    def _resolve_substeps(self, dt, dt_stable, medium) -> int:
        if dt <= dt_stable:
            return 1
        if self.on_unstable == "substep":
            n_sub = int(ceil(dt / dt_stable))
            log.info("pennes: dt=%.6g s exceeds the stability bound %.6g s; sub-stepping %d x %.6g s",
                     dt, dt_stable, n_sub, dt / n_sub)
            return n_sub
        raise ThermalStabilityError(
            f"dt = {dt:.6g} s is above the stability bound {dt_stable:.6g} s of the explicit scheme "
            f"for this medium (dx = {medium.dx * 1e3:.6g} mm). Pass dt <= {dt_stable:.6g} s, use "
            f"PennesSolver(on_unstable='substep'), or coarsen dx (the bound scales as dx^2)."
        )

    def _prepare_q(self, q, medium):
        if q is None:
            return None, "none"
        shape = medium.shape
        if isinstance(q, DeviceHeatingSource):
            if abs(q.dx - medium.dx) > 1e-12 * max(q.dx, medium.dx):
                raise ValueError(f"the heat source was built on dx = {q.dx * 1e3:.6g} mm but this "
                                 f"thermal medium has dx = {medium.dx * 1e3:.6g} mm.")
            if q.shape != tuple(shape):
                raise ValueError(f"the heat source covers the record region {q.region} (shape "
                                 f"{q.shape}) but the thermal medium is {shape}.")
            return q, f"DeviceHeatingSource({q.alpha_model})"
        if isinstance(q, HeatingSource):
            if abs(q.dx - medium.dx) > 1e-12 * max(q.dx, medium.dx):
                raise ValueError(f"the HeatingSource was built on dx = {q.dx * 1e3:.6g} mm but this "
                                 f"thermal medium has dx = {medium.dx * 1e3:.6g} mm.")
            if q.q.shape != tuple(shape):
                raise ValueError(f"the HeatingSource covers the record region {q.region} (shape "
                                 f"{q.q.shape}) but the thermal medium is {shape}; use "
                                 f"heating_source.embed({shape}).")
            return np.ascontiguousarray(q.q, np.float32), f"HeatingSource({q.alpha_model})"
        arr = np.asarray(q, dtype=np.float32)
        if arr.ndim == 0:
            arr = np.full(shape, float(arr), np.float32)
        if arr.shape != tuple(shape):
            raise ValueError(f"Q has shape {arr.shape} but the thermal medium is {shape}.")
        if not np.isfinite(arr).all():
            raise ValueError("Q contains non-finite values.")
        return np.ascontiguousarray(arr), "array"

    # ---------------------------------------------------------------- solve
    def solve(self, temperature0, q, medium: ThermalMedium, dt: float, n_steps: int, *,
              record_every: int | None = None, dose: bool = False, dose0=None) -> ThermalResult:
        if not isinstance(medium, ThermalMedium):
            raise TypeError(f"medium must be a ThermalMedium, got {type(medium).__name__}.")
        if not (np.isfinite(dt) and dt > 0):
            raise ValueError(f"dt must be a positive finite number of seconds, got {dt}")
        n_steps = int(n_steps)
        if n_steps < 1:
            raise ValueError(f"n_steps must be >= 1, got {n_steps}")
        if record_every is not None and int(record_every) < 1:
            raise ValueError(f"record_every must be >= 1 or None, got {record_every}")
        if dose0 is not None and not dose:
            raise ValueError("dose0 was given but dose=False; pass dose=True to keep accumulating.")
        b = get_backend(self.backend)
        if not b.native:
            raise ValueError(f"backend {b.name!r} has no native Pennes solver in caustica_amd")
        shape = tuple(medium.shape)
        if tuple(np.shape(temperature0)) != shape:
            raise ValueError(f"temperature0 has shape {np.shape(temperature0)} but the medium is {shape}.")
        t_ref = np.float32(self.reference_temperature_c)
        theta0 = np.ascontiguousarray(np.asarray(temperature0, np.float32) - t_ref)
        if not np.isfinite(theta0).all():
            raise ValueError("temperature0 contains non-finite values.")
        q_arr, q_kind = self._prepare_q(q, medium)
        dt_stable = self.stable_dt(medium)
        n_sub = self._resolve_substeps(dt, dt_stable, medium)
        dt_sub = dt / n_sub
        coeff = np.ascontiguousarray((self._inv_rho_c(medium) * np.float32(dt_sub)).astype(np.float32))
        perf = np.ascontiguousarray(self._perf(medium))
        q_dev = q_arr if isinstance(q_arr, DeviceHeatingSource) else None
        q_off = 0.0
        if float(self.reference_temperature_c - self.arterial_temperature_c) != 0.0 and medium.is_perfused:
            # perfusion sink -perf (theta + (t_ref - T_a)): fold the constant into Q
            q_off = float(np.float32(self.reference_temperature_c - self.arterial_temperature_c))
            if q_dev is None:
                off = (perf * np.float32(q_off)).astype(np.float32)
                q_arr = (-off if q_arr is None else (q_arr - off)).astype(np.float32)
        if q_dev is not None:
            q_arr = None  # formed on the device after setup
        d0 = None
        if dose:
            if dose0 is not None:
                d0 = np.ascontiguousarray(np.asarray(dose0, np.float32))
                if d0.shape != shape:
                    raise ValueError(f"dose0 has shape {d0.shape} but the medium is {shape}.")
        nd = len(shape)
        shape3 = (1,) * (3 - nd) + shape
        mask = 0
        if self.boundary == "dirichlet":
            for ax in range(3 - nd, 3):
                mask |= 1 << ax
        desc = N.ThermalDesc()
        desc.dx = medium.dx
        keep = [np.ascontiguousarray(medium.k, np.float32), perf, coeff, q_arr, theta0, d0]
        desc.k = N.ptr(keep[0])
        desc.perf = N.ptr(perf)
        desc.coeff = N.ptr(coeff)
        desc.q = N.ptr(q_arr)
        desc.theta0 = N.ptr(theta0)
        desc.dose0 = N.ptr(d0)
        desc.boundary = mask
        desc.theta_wall = float(np.float32(self.boundary_temperature_c) - t_ref)
        desc.t_ref = float(t_ref)
        desc.half_dt_min = float(np.float32(0.5 * dt_sub / 60.0))
        desc.dose = 1 if dose else 0
        lib = N.lib()
        h = C.c_void_p()
        N.check(lib.cst_thermal_create(*shape3, b.device, C.byref(h)))
        samples, times = [], []
        try:
            N.check(lib.cst_thermal_setup(h, C.byref(desc)))
            if self.x_chunk:
                N.check(lib.cst_thermal_set_chunk(h, self.x_chunk))
            uni = C.c_int(0)
            N.check(lib.cst_thermal_set_uniform_maps(h, int(self.uniform_maps), C.byref(uni)))
            scalar_maps = bool(uni.value) and self.uniform_maps
            if self.profile:
                N.check(lib.cst_thermal_set_profiling(h, 1))
            if q_dev is not None:
                idx = np.array([t[0] for t in q_dev.terms], np.int32)
                fac = np.array([t[1] for t in q_dev.terms], np.float64)
                N.check(lib.cst_thermal_heat_from_member(
                    h, q_dev.session.handle, q_dev.slot, len(idx), N.ptr(idx), N.ptr(fac),
                    N.ptr(q_dev.w), float(q_dev.scale), float(q_off)))

            def read(theta=True, tmax=False, dz=False):
                out = [np.empty(shape, np.float32) if f else None for f in (theta, tmax, dz)]
                N.check(lib.cst_thermal_read(h, *[N.ptr(a) for a in out]))
                return out

            def record(step_idx):
                samples.append(read()[0] + t_ref)
                times.append(step_idx * dt)

            if record_every is not None:
                record(0)
            step = 0
            while step < n_steps:
                nxt = n_steps
                if record_every is not None:
                    nxt = min(nxt, (step // record_every + 1) * record_every)
                nxt = min(nxt, (step // self.finite_check_every + 1) * self.finite_check_every)
                N.check(lib.cst_thermal_run(h, (nxt - step) * n_sub))
                step = nxt
                if record_every is not None and step % record_every == 0:
                    record(step)
                if step % self.finite_check_every == 0 or step == n_steps:
                    self._guard(lib, h, step, dt, n_sub)
            if record_every is not None and (not times or times[-1] != n_steps * dt):
                record(n_steps)
            theta, tmax, dz = read(True, True, dose)
            prof = None
            if self.profile:
                ms, cnt = C.c_double(), C.c_int64()
                N.check(lib.cst_thermal_kernel_time(h, C.byref(ms), C.byref(cnt)))
                prof = {"kernel_ms": ms.value, "launches": cnt.value}
        finally:
            lib.cst_thermal_destroy(h)
        return ThermalResult(
            temperature=theta + t_ref,
            temperature_max=tmax + t_ref,
            dose_cem43=dz,
            samples=samples,
            times=times,
            dt=float(dt),
            n_steps=n_steps,
            t_end_s=float(dt * n_steps),
            substeps=n_sub,
            meta={
                "scheme": SCHEME,
                "backend": b.name,
                "boundary": self.boundary,
                "boundary_temperature_c": self.boundary_temperature_c,
                "dt_stable_s": dt_stable,
                "dt_sub_s": dt_sub,
                "q": q_kind,
                "perfusion_active": medium.is_perfused,
                "blood_density": self.blood_density,
                "blood_specific_heat": self.blood_specific_heat,
                "arterial_temperature_c": self.arterial_temperature_c,
                "reference_temperature_c": self.reference_temperature_c,
                "dose_accumulated": bool(dose),
                "uniform_map_kernels": scalar_maps,
                **({"profile": prof} if prof is not None else {}),
            },
        )

    @staticmethod
    def _guard(lib, h, step, dt, n_sub) -> None:
        bad = C.c_int(0)
        N.check(lib.cst_thermal_status(h, C.byref(bad)))
        if bad.value:
            raise ThermalDivergedError(
                f"the thermal state stopped being finite by outer step {step} (t = {step * dt:.6g} s, "
                f"{n_sub} sub-step(s) per step). Usual causes: a heat source far outside the range "
                f"float32 can carry, a temperature above ~170 C where R^(43-T) overflows, or a "
                f"medium the stability bound could not see."
            )
