"""Acoustic -> thermal join: volumetric heat deposition Q from a CW result.

Upstream ``src/caustica/sensors.py:100-435``. For a locally progressive wave the
absorbed power density of harmonic n is Q_n = 2 alpha_n I_n = alpha_n |P_n|^2 / (rho c)
(P a peak amplitude, p(t) = Re{P exp(-i omega t)}). Harmonics above f0 need a stated
absorption model (``alpha_power_law_y`` -> alpha_n = alpha_f0 n^y, or explicit
``alpha_at_harmonics``); a result with harmonics and no model is refused, as upstream.
The sum is formed in float64 and stored float32.
"""

from __future__ import annotations

import dataclasses
import warnings
from dataclasses import dataclass, field
from typing import Any

import numpy as np

from caustica_amd.core.backend import CausticaWarning

ALPHA_MODEL_F0 = "f0_only"
ALPHA_MODEL_POWER_LAW = "power_law"
ALPHA_MODEL_EXPLICIT = "explicit"
ALPHA_POWER_LAW_Y_RANGE = (0.0, 3.0)


def _as_map(value: Any, shape: tuple, what: str) -> np.ndarray:
    arr = np.asarray(value, dtype=np.float64)
    if arr.shape not in ((), shape):
        raise ValueError(f"{what} has shape {arr.shape}, neither a scalar nor the field shape {shape}.")
    if not np.isfinite(arr).all():
        raise ValueError(f"{what} contains non-finite values.")
    return np.broadcast_to(arr, shape)


@dataclass(frozen=True, eq=False)
class HeatingSource:
    q: np.ndarray
    region: tuple
    dx: float
    harmonics: tuple
    alpha_model: str
    meta: dict[str, Any] = field(default_factory=dict)

    @classmethod
    def from_result(cls, result, medium, dx: float, *, harmonics=None, alpha_power_law_y=None,
                    alpha_at_harmonics=None) -> HeatingSource:
        phasors = dict(result.phasors) if result.phasors else {1: result.phasor}
        region = tuple(result.region)
        if len(region) != medium.alpha.ndim:
            raise ValueError(f"result.region has rank {len(region)} but the medium is "
                             f"{medium.alpha.ndim}-D -- they are not the same grid.")
        shape = tuple(np.shape(phasors[next(iter(phasors))]))
        maps = {k: getattr(medium, k)[region] for k in ("alpha", "rho", "c")}
        for name, vol in maps.items():
            if vol.shape != shape:
                raise ValueError(f"medium.{name} sliced by result.region has shape {vol.shape} but "
                                 f"the recorded phasor has shape {shape}.")
        return cls.from_phasors(phasors, alpha=maps["alpha"], rho=maps["rho"], c=maps["c"], dx=dx,
                                region=region, harmonics=harmonics,
                                alpha_power_law_y=alpha_power_law_y,
                                alpha_at_harmonics=alpha_at_harmonics)

    @classmethod
    def from_phasors(cls, phasors: dict, *, alpha, rho, c, dx: float, region=None, harmonics=None,
                     alpha_power_law_y=None, alpha_at_harmonics=None) -> HeatingSource:
        if not phasors:
            raise ValueError("phasors is empty: there is no field to turn into heat.")
        if not (np.isfinite(dx) and dx > 0):
            raise ValueError(f"dx must be a positive finite spacing in meters, got {dx}")
        available = tuple(sorted(int(n) for n in phasors))
        if harmonics is None:
            requested = available
        else:
            requested = tuple(dict.fromkeys(int(n) for n in harmonics))
            missing = [n for n in requested if n not in phasors]
            if missing:
                raise ValueError(f"harmonics {missing} were asked for but are not in the result "
                                 f"(available: {list(available)}).")
        if any(n < 1 for n in requested):
            raise ValueError(f"harmonic numbers must be >= 1, got {requested}")
        shape = tuple(np.shape(phasors[requested[0]]))
        for n in requested:
            if tuple(np.shape(phasors[n])) != shape:
                raise ValueError(f"harmonic {n} has shape {np.shape(phasors[n])}, not {shape}.")
        alpha_f0 = _as_map(alpha, shape, "alpha")
        rho_m = _as_map(rho, shape, "rho")
        c_m = _as_map(c, shape, "c")
        if (alpha_f0 < 0).any():
            raise ValueError("alpha must be >= 0 Np/m (absorption cannot be negative).")
        if (rho_m <= 0).any() or (c_m <= 0).any():
            raise ValueError("rho and c must be > 0 everywhere to convert |P|^2 into intensity.")
        alpha_by_h = _harmonic_alpha(requested, alpha_f0, available, alpha_power_law_y, alpha_at_harmonics)
        if alpha_at_harmonics is not None:
            model = ALPHA_MODEL_EXPLICIT
        elif alpha_power_law_y is not None and max(requested) > 1:
            model = ALPHA_MODEL_POWER_LAW
        else:
            model = ALPHA_MODEL_F0
        inv_rho_c = 1.0 / (rho_m * c_m)
        q = np.zeros(shape, np.float64)
        mean_by_h = {}
        for n in requested:
            p = np.asarray(phasors[n])
            if not np.isfinite(p).all():
                raise ValueError(f"the phasor of harmonic {n} contains non-finite values.")
            contrib = alpha_by_h[n] * (np.abs(p).astype(np.float64) ** 2) * inv_rho_c
            q += contrib
            mean_by_h[n] = float(contrib.mean())
        if region is None:
            region = tuple(slice(0, n) for n in shape)
        meta = {
            "alpha_model": model,
            "alpha_power_law_y": alpha_power_law_y,
            "mean_q_by_harmonic_w_m3": mean_by_h,
            "harmonics_available": list(available),
            "intensity_model": "local plane wave: I = |P|^2 / (2 rho c)",
        }
        return cls(q=np.ascontiguousarray(q, dtype=np.float32), region=tuple(region), dx=float(dx),
                   harmonics=requested, alpha_model=model, meta=meta)

    @property
    def shape(self):
        return tuple(self.q.shape)

    @property
    def q_max(self) -> float:
        return float(self.q.max())

    @property
    def total_power_w(self) -> float:
        return float(self.q.sum()) * self.dx ** self.q.ndim

    def scaled(self, factor: float) -> HeatingSource:
        """Same distribution, Q multiplied by ``factor`` (drive scaling: Q goes as p^2)."""
        return HeatingSource(q=(self.q.astype(np.float64) * factor).astype(np.float32),
                             region=self.region, dx=self.dx, harmonics=self.harmonics,
                             alpha_model=self.alpha_model, meta={**self.meta, "scaled_by": factor})

    def embed(self, shape) -> np.ndarray:
        shape = tuple(int(s) for s in shape)
        if len(shape) != self.q.ndim:
            raise ValueError(f"shape {shape} has rank {len(shape)}, need {self.q.ndim}")
        full = np.zeros(shape, np.float32)
        try:
            full[self.region] = self.q
        except ValueError as exc:
            raise ValueError(f"the region {self.region} does not fit a grid of shape {shape}: {exc}") from None
        return full


def _harmonic_alpha(requested, alpha_f0, available, y, explicit) -> dict:
    if y is not None and explicit is not None:
        raise ValueError("pass EITHER alpha_power_law_y OR alpha_at_harmonics, not both.")
    high = [n for n in requested if n >= 2]
    if high and y is None and explicit is None:
        raise ValueError(
            f"this result carries harmonics {high} above the fundamental, but alpha is only known "
            f"at f0. State the model: alpha_power_law_y=y (alpha_n = alpha_f0 n^y), "
            f"alpha_at_harmonics={{n: alpha_np_m}}, or harmonics=(1,) for fundamental only."
        )
    if y is not None:
        lo, hi = ALPHA_POWER_LAW_Y_RANGE
        if not (np.isfinite(y) and lo <= float(y) <= hi):
            raise ValueError(f"alpha_power_law_y={y} is outside the physical range [{lo}, {hi}].")
        return {n: alpha_f0 * float(n) ** float(y) for n in requested}
    if explicit is not None:
        keys = {int(n): v for n, v in explicit.items()}
        missing = [n for n in requested if n >= 2 and n not in keys]
        if missing:
            raise ValueError(f"alpha_at_harmonics is missing harmonic(s) {missing}.")
        out = {}
        hetero = float(alpha_f0.min()) != float(alpha_f0.max())
        for n in requested:
            if n not in keys:
                out[n] = alpha_f0
                continue
            if hetero and np.asarray(keys[n]).ndim == 0:
                warnings.warn(f"alpha_at_harmonics[{n}] is a scalar but alpha at f0 is heterogeneous; "
                              f"the scalar is applied to every voxel.", CausticaWarning, stacklevel=4)
            arr = _as_map(keys[n], alpha_f0.shape, f"alpha_at_harmonics[{n}]")
            if (arr < 0).any():
                raise ValueError(f"alpha_at_harmonics[{n}] has negative values.")
            out[n] = arr
        return out
    return {n: alpha_f0 for n in requested}


@dataclass(frozen=True, eq=False)
class DeviceHeatingSource:
    """Heat source formed on the device from an acoustic member still resident in HBM.

    Same physics and validation as ``HeatingSource.from_phasors`` (Q = sum_n alpha_n
    |P_n|^2 / (rho c), float64 sum, float32 store), but the phasors never leave the
    device: ``PennesSolver.solve`` hands it to ``cst_thermal_heat_from_member``, which
    writes Q straight into the thermal session. Only the map alpha_f0 / (rho c) (float32)
    crosses the bus. |P|^2 is formed in float64 from the float32 phasor components.
    """

    session: Any
    slot: int
    w: np.ndarray
    terms: tuple          # ((harmonic index in the member, absorption factor), ...)
    region: tuple
    dx: float
    harmonics: tuple
    alpha_model: str
    scale: float = 1.0
    meta: dict[str, Any] = field(default_factory=dict)

    @classmethod
    def from_member(cls, session, slot: int, plan, medium, dx: float, *, rec_steps: int,
                    harmonics=None, alpha_power_law_y=None) -> DeviceHeatingSource:
        """Finalize the member's phasors on the device and describe the join.

        ``plan`` is the member's MemberPlan (record region and harmonics), ``medium`` the
        acoustic Medium, ``rec_steps`` the number of recorded steps (the phasor scale is
        2 / rec_steps, as for the host result)."""
        available = tuple(int(h) for h in plan.harmonics)
        requested = available if harmonics is None else tuple(dict.fromkeys(int(n) for n in harmonics))
        missing = [n for n in requested if n not in available]
        if missing:
            raise ValueError(f"harmonics {missing} were asked for but are not in the result "
                             f"(available: {list(available)}).")
        if any(n < 1 for n in requested):
            raise ValueError(f"harmonic numbers must be >= 1, got {requested}")
        if not (np.isfinite(dx) and dx > 0):
            raise ValueError(f"dx must be a positive finite spacing in meters, got {dx}")
        region = tuple(plan.rec)
        shape = tuple(s.stop - s.start for s in region)
        maps = {k: np.asarray(getattr(medium, k)[region], np.float64) for k in ("alpha", "rho", "c")}
        if (maps["alpha"] < 0).any():
            raise ValueError("alpha must be >= 0 Np/m (absorption cannot be negative).")
        if (maps["rho"] <= 0).any() or (maps["c"] <= 0).any():
            raise ValueError("rho and c must be > 0 everywhere to convert |P|^2 into intensity.")
        # the same absorption-model rules as the host path (scalar factors per harmonic)
        factors = _harmonic_alpha(requested, np.ones((), np.float64), available, alpha_power_law_y, None)
        model = (ALPHA_MODEL_POWER_LAW if alpha_power_law_y is not None and max(requested) > 1
                 else ALPHA_MODEL_F0)
        w = np.ascontiguousarray(maps["alpha"] / (maps["rho"] * maps["c"]), dtype=np.float32)
        terms = []
        for n in requested:
            hi = available.index(n)
            session.result(slot, hi, 2.0 / rec_steps, with_pmax=False, copy=False)
            terms.append((hi, float(factors[n])))
        return cls(session=session, slot=int(slot), w=w, terms=tuple(terms), region=region,
                   dx=float(dx), harmonics=requested, alpha_model=model,
                   meta={"alpha_model": model, "alpha_power_law_y": alpha_power_law_y,
                         "intensity_model": "local plane wave: I = |P|^2 / (2 rho c)"})

    @property
    def shape(self):
        return tuple(self.w.shape)

    def to_host(self) -> np.ndarray:
        """Q [W/m^3] as the device forms it (float32 over the region), read back: the join
        runs into a scratch thermal session of the region's shape (no perfusion offset)
        and cst_thermal_read_source copies it out. For inspection and tests; a Pennes
        solve never needs it."""
        import ctypes as C

        from caustica_amd import _native as N

        lib = N.lib()
        shape = self.shape
        shape3 = (1,) * (3 - len(shape)) + tuple(shape)
        zero = np.zeros(shape, np.float32)
        desc = N.ThermalDesc()
        desc.dx = self.dx
        desc.k = desc.perf = desc.coeff = desc.theta0 = N.ptr(zero)
        desc.q = desc.dose0 = None
        desc.half_dt_min = 0.0
        h = C.c_void_p()
        N.check(lib.cst_thermal_create(*shape3, self.session.device, C.byref(h)))
        try:
            N.check(lib.cst_thermal_setup(h, C.byref(desc)))
            idx = np.array([t[0] for t in self.terms], np.int32)
            fac = np.array([t[1] for t in self.terms], np.float64)
            N.check(lib.cst_thermal_heat_from_member(h, self.session.handle, self.slot, len(idx),
                                                     N.ptr(idx), N.ptr(fac), N.ptr(self.w),
                                                     float(self.scale), 0.0))
            out = np.empty(shape, np.float32)
            N.check(lib.cst_thermal_read_source(h, N.ptr(out)))
        finally:
            lib.cst_thermal_destroy(h)
        return out

    def scaled(self, factor: float) -> DeviceHeatingSource:
        """Same distribution, Q multiplied by ``factor`` (applied in float64 on the device)."""
        return dataclasses.replace(self, scale=self.scale * float(factor),
                                   meta={**self.meta, "scaled_by": self.scale * float(factor)})
