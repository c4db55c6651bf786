"""CW k-space PSTD engine on MI355X: host control loop over the native step loop.

Same contract as upstream ``src/caustica/solvers/kspace/engine.py:167-586``: the
scheme (collocated leapfrog, kappa-corrected spectral derivatives with the even-axis
Nyquist derivative zeroed, exp(-alpha c dt) absorption on p and u, multiplicative
Gaussian sponge, 2 c dt / dx mass source with a cosine ramp, optional Westervelt
term), the exact-period schedule (settle until the per-period peak converges,
capped; optional t_end floor; record n_record_periods), the single-bin phasor
convention p(t) = Re{P exp(-i omega t)}, the finite guards, the progress payload and
checkpoint/resume. The per-step work (eight fused passes, engine.py:458-482) runs in
libcaustica_hip; this module only decides how many steps to run and reads back one
peak per period -- or, when nothing can stop the run early (convergence_tol == 0, no
progress callback, no checkpoint), it queues the whole solve at once and reads the
period peaks at the end.

``scheme="staggered"`` runs the opt-in staggered scheme instead (half-cell shifted
spectral derivatives, fractional-Laplacian power-law absorption alpha0 omega^y with
``alpha_power`` y, split-field PML of strength ``pml_alpha``; see
``caustica_amd.solvers.kspace.staggered``), which has no upstream counterpart; the
schedule, record and result are the same.
"""

from __future__ import annotations

import logging
import time
import warnings
from collections.abc import Callable
from math import ceil, isfinite

import numpy as np

from caustica_amd import _native as N
from caustica_amd.core.backend import CausticaWarning, get_backend
from caustica_amd.io.checkpoint import (
    CheckpointSpec,
    RunInterrupted,
    arrays_sha1,
    load_checkpoint,
    write_checkpoint,
)
from caustica_amd.solvers.base import (
    CWRunSpec,
    SolverCapabilityError,
    SolverDivergedError,
    SolverResult,
)
from caustica_amd.solvers.kspace.native import KSpaceSession
from caustica_amd.solvers.kspace.operators import MemberPlan, plan_member
from caustica_amd.solvers.kspace.staggered import SCHEME as STAGGERED_SCHEME
from caustica_amd.solvers.kspace.staggered import plan_staggered
from caustica_amd.tune import autotune

log = logging.getLogger("caustica")

SCHEME = "cw-kspace-pstd/2"
#: grids at least this large get their pass variants autotuned (cached per shape)
TUNE_MIN_VOXELS = 1 << 21


def guard_finite(peak: float, *, where: str, solver_name: str) -> None:
    if isfinite(peak):
        return
    raise SolverDivergedError(
        f"solver '{solver_name}' diverged during {where}: the peak pressure is {peak}, so the "
        f"field is no longer finite. Usual causes: a time step the scheme cannot integrate for "
        f"this medium (check the CFL and points-per-wavelength of the fastest material), a drive "
        f"amplitude far outside the range float32 can carry, a non-physical sound speed or "
        f"density, or a source sitting in or against the PML band."
    )


def check_harmonics(harmonics) -> tuple[int, ...]:
    harmonics = tuple(dict.fromkeys(int(h) for h in harmonics))
    if not harmonics or harmonics[0] != 1 or any(h < 1 for h in harmonics):
        raise ValueError(f"harmonics must start with 1 (fundamental), got {harmonics}")
    return harmonics


def check_reference_point(grid, reference_point) -> None:
    if reference_point is None:
        return
    ref = tuple(reference_point)
    if len(ref) != grid.ndim or any(
        not float(r).is_integer() or not 0 <= int(r) < n for r, n in zip(ref, grid.shape)
    ):
        raise ValueError(
            f"reference_point must be integer VOXEL indices inside {grid.shape}, got "
            f"{reference_point} (did you pass meters?)"
        )


def _embedded_u(nd: int) -> list[int]:
    """Device velocity components that carry the grid's axes (rank-1/2 grids embed in 3-D)."""
    return list(range(3 - nd, 3))


def run_cw_kspace_pstd(
    solver_name: str,
    grid,
    medium,
    source,
    spec: CWRunSpec,
    backend: str = "auto",
    record_region=None,
    reference_point=None,
    nonlinear: bool = False,
    harmonics=(1,),
    checkpoint: CheckpointSpec | None = None,
    progress: Callable[[dict], None] | None = None,
    scheme: str = "reference",
    alpha_power: float = 1.5,
    pml_alpha: float = 2.0,
) -> SolverResult:
    harmonics = check_harmonics(harmonics)
    check_reference_point(grid, reference_point)
    if scheme not in ("reference", "staggered"):
        raise ValueError(f"scheme must be 'reference' or 'staggered', got {scheme!r}")
    b = get_backend(backend)
    if not b.native:
        raise SolverCapabilityError(
            f"backend {b.name!r} has no native CW engine in caustica_amd (only 'hip')."
        )
    if len(harmonics) > 3:
        raise SolverCapabilityError("the hip engine records at most 3 harmonics per solve")
    if scheme == "staggered":
        plan = plan_staggered(grid, medium, source, spec, harmonics=harmonics,
                              record_region=record_region, reference_point=reference_point,
                              nonlinear=nonlinear, alpha_power=alpha_power, pml_alpha=pml_alpha)
    else:
        plan = plan_member(grid, medium, source, spec, harmonics=harmonics,
                           record_region=record_region, reference_point=reference_point,
                           nonlinear=nonlinear)
    if grid.pml_vox == 0:
        log.warning(
            "grid has no PML: the FFT domain is PERIODIC and waves wrap around, producing "
            "standing-wave interference. Attach a PMLSpec to the Grid unless periodic "
            "boundaries are intended."
        )
    period = 1.0 / source.f0
    extra = 0
    if spec.t_end_min_us is not None:
        extra = max(0, ceil(spec.t_end_min_us * 1e-6 / period) - spec.n_record_periods - plan.eff_max)
    max_periods = plan.eff_max + extra + spec.n_record_periods + 4
    nvox = plan.padded3[0] * plan.padded3[1] * plan.padded3[2]
    with KSpaceSession(plan.padded3, 2, b.device, scheme=plan.scheme) as sess:
        if nvox >= TUNE_MIN_VOXELS and plan.scheme == "reference":
            autotune(sess, plan)
        slot = sess.add(plan, max_periods=max_periods)
        return _control_loop(sess, slot, plan, solver_name, grid, medium, source, spec, b,
                             reference_point, checkpoint, progress)


def _fingerprint(plan: MemberPlan, solver_name, grid, medium, source, spec, backend_name,
                 reference_point) -> dict:
    out = {
        "scheme": SCHEME if plan.scheme == "reference" else STAGGERED_SCHEME,
        "solver": solver_name,
        "backend": backend_name,
        "grid_shape": list(grid.shape),
        "dx": grid.dx,
        "pml_vox": grid.pml_vox,
        "pml_edge": plan.pml_edge,
        "spp": plan.spp,
        "f0": source.f0,
        "amplitude": float(source.amplitude),
        "ramp_periods": source.ramp_periods,
        "harmonics": list(plan.harmonics),
        "nonlinear": plan.beta2_dt is not None or bool(plan.stg and plan.stg["nonlinear"]),
        "record_region": [[s.start, s.stop] for s in plan.rec],
        "reference_point": None if reference_point is None else [int(r) for r in reference_point],
        "spec": spec.model_dump(),
        "source_sha1": arrays_sha1(source.indices, source.phases),
        "medium_sha1": arrays_sha1(medium.c, medium.rho, medium.alpha, medium.beta),
        # the raw device state's layout is the library's: a checkpoint resumes only on
        # the library version that wrote it
        "device_state_version": int(N.lib().cst_version()) if backend_name == "hip" else None,
    }
    if plan.scheme == "staggered":
        out["alpha_power"] = plan.stg["alpha_power"]
        out["pml_alpha"] = plan.stg["pml_alpha"]
    return out


def _control_loop(sess: KSpaceSession, slot: int, plan: MemberPlan, solver_name, grid, medium,
                  source, spec, b, reference_point, checkpoint, progress) -> SolverResult:
    spp = plan.spp
    nd = grid.ndim
    period = 1.0 / source.f0
    tof_periods, eff_min, eff_max = plan.tof_periods, plan.eff_min, plan.eff_max
    fingerprint = (
        _fingerprint(plan, solver_name, grid, medium, source, spec, b.name, reference_point)
        if checkpoint is not None
        else None
    )
    n = 0
    prev_peak = None
    converged = False
    converged_period = eff_max
    history: list = []
    periods_done = 0
    stage = "settle"
    ucomp = _embedded_u(nd)

    if checkpoint is not None:
        state = load_checkpoint(checkpoint.path, fingerprint)
        if state is not None:
            f = state["fields"]
            if "device_state" in f:
                sess.set_raw(slot, state["n"], f["device_state"])
            elif plan.scheme == "staggered":
                raise ValueError(f"checkpoint {checkpoint.path} holds no device state: a staggered "
                                 f"solve resumes only from the library's own state (p, u, rho_a)")
            else:
                u = [np.zeros(plan.padded3, np.float32) for _ in range(3)]
                for i, c in enumerate(ucomp):
                    u[c] = f[f"u{i}"].reshape(plan.padded3)
                sess.set_state(slot, state["n"], f["p"], *u)
            n = state["n"]
            periods_done = state["periods_done"]
            prev_peak = state["prev_peak"]
            converged = state["converged"]
            converged_period = state["converged_period"]
            history = state["history"]
            stage = state["stage"]

    def _write_ckpt(stage_now: str) -> None:
        # the real p and u for readers, and the exact device state for a bitwise resume
        (p, *u), _n = sess.get_state(slot)
        fields = {"p": p.reshape(plan.padded)}
        for i, c in enumerate(ucomp):
            fields[f"u{i}"] = u[c].reshape(plan.padded)
        fields["device_state"], _n = sess.get_raw(slot)
        write_checkpoint(checkpoint.path, stage=stage_now, n=n, periods_done=periods_done,
                         prev_peak=prev_peak, converged=converged,
                         converged_period=converged_period, history=history, fields=fields,
                         fingerprint=fingerprint)

    periods_expected = eff_max
    if spec.t_end_min_us is not None:
        periods_expected = max(periods_expected,
                               ceil(spec.t_end_min_us * 1e-6 / period) - spec.n_record_periods)
    periods_expected += spec.n_record_periods
    steps_expected = periods_expected * spp
    t0 = time.monotonic()
    steps_at_start = n
    last_peak = prev_peak
    last_delta = None
    warned = [False]

    def _snapshot() -> np.ndarray:
        (p, *_u), _n = sess.get_state(slot, with_u=False)
        p = p.reshape(plan.padded)
        if nd == 1:
            return p[: grid.shape[0]].copy()
        if nd == 2:
            return p[: grid.shape[0], : grid.shape[1]].copy()
        j = int(reference_point[1]) if reference_point is not None else grid.shape[1] // 2
        return p[: grid.shape[0], j, : grid.shape[2]].copy()

    def _emit(stage_now: str) -> None:
        if progress is None:
            return
        elapsed = time.monotonic() - t0
        done = n - steps_at_start
        eta = None
        if done > 0 and steps_expected > n:
            eta = round((steps_expected - n) * (elapsed / done), 1)
        try:
            progress({
                "period": periods_done,
                "periods_expected": periods_expected,
                "step": n,
                "steps_expected": steps_expected,
                "peak": last_peak,
                "converge_delta": last_delta,
                "elapsed_s": round(elapsed, 3),
                "eta_s": eta,
                "stage": stage_now,
                "snapshot": _snapshot,
            })
        except Exception as exc:  # noqa: BLE001 - a widget must not kill a solve
            if not warned[0]:
                warned[0] = True
                warnings.warn(
                    f"progress callback raised {type(exc).__name__}: {exc}. The solve continues; "
                    f"further progress failures are silent.",
                    CausticaWarning,
                    stacklevel=2,
                )

    def _period_boundary() -> None:
        _emit("settle")
        if checkpoint is None:
            return
        stop = checkpoint.stop_when is not None and checkpoint.stop_when()
        if stop or periods_done % checkpoint.every_periods == 0:
            _write_ckpt("settle")
        if stop:
            raise RunInterrupted(checkpoint.path, "settle", periods_done, n)

    t_end_need = None
    if spec.t_end_min_us is not None:
        t_end_need = ceil(spec.t_end_min_us * 1e-6 / period) - spec.n_record_periods
    rec_steps = spec.n_record_periods * spp

    lazy = (checkpoint is None and progress is None and spec.convergence_tol == 0.0
            and stage == "settle" and n == 0)
    if lazy:
        # nothing can end the settle loop early: queue the whole solve, read peaks once
        settle_periods = eff_max
        if t_end_need is not None:
            settle_periods = max(settle_periods, t_end_need)
        record_from = settle_periods * spp
        sess.schedule(slot, record_from)
        sess.advance([slot], record_from + rec_steps)
        peaks = sess.peaks(slot, 0, eff_max)
        for period_idx in range(1, eff_max + 1):
            peak = float(peaks[period_idx - 1])
            guard_finite(peak, where=f"settling period {period_idx} (step {period_idx * spp})",
                         solver_name=solver_name)
            rel = abs(peak - prev_peak) / max(prev_peak, 1e-9) if prev_peak is not None else None
            if period_idx >= tof_periods:
                history.append((period_idx, peak, float("nan") if rel is None else rel))
            prev_peak = peak
        n = record_from + rec_steps
        periods_done = settle_periods
    else:
        period_idx = periods_done
        while stage == "settle" and not converged and period_idx < eff_max:
            period_idx += 1
            sess.advance([slot], spp)
            n += spp
            peak = float(sess.peaks(slot, period_idx - 1, 1)[0])
            guard_finite(peak, where=f"settling period {period_idx} (step {n})",
                         solver_name=solver_name)
            periods_done = period_idx
            rel = abs(peak - prev_peak) / max(prev_peak, 1e-9) if prev_peak is not None else None
            last_peak, last_delta = peak, rel
            if period_idx >= tof_periods:
                history.append((period_idx, peak, float("nan") if rel is None else rel))
            if period_idx >= eff_min and rel is not None:
                if peak > 0.0 and rel < spec.convergence_tol:
                    converged_period = period_idx
                    converged = True
            if not converged:
                prev_peak = peak
            _period_boundary()
        if t_end_need is not None and stage == "settle":
            while periods_done < t_end_need:
                sess.advance([slot], spp)
                n += spp
                periods_done += 1
                _period_boundary()
        if stage == "settle":
            stage = "record"
            _emit("record")
            if checkpoint is not None:
                _write_ckpt("record")
                if checkpoint.stop_when is not None and checkpoint.stop_when():
                    raise RunInterrupted(checkpoint.path, "record", periods_done, n)
        sess.schedule(slot, n)
        sess.advance([slot], rec_steps)
        n += rec_steps

    rshape = plan.rec_shape
    phasors = {}
    pmax = None
    for hi, h in enumerate(plan.harmonics):
        ph, pm = sess.result(slot, hi, 2.0 / rec_steps, with_pmax=(hi == 0))
        phasors[h] = ph.reshape(rshape)
        if hi == 0:
            pmax = pm.reshape(rshape)
    guard_finite(float(pmax.max()), where=f"the record window (step {n})", solver_name=solver_name)
    if checkpoint is not None and not checkpoint.keep_on_success:
        checkpoint.path.unlink(missing_ok=True)
    return SolverResult(
        phasor=phasors[1],
        p_max=pmax,
        region=plan.rec,
        dt=plan.dt,
        spp=spp,
        steps_total=n,
        t_end_s=n * plan.dt,
        tof_periods=tof_periods,
        converged_period=converged_period,
        settle_capped=not converged,
        convergence_history=history,
        phasors=phasors,
        meta={
            "solver": solver_name,
            "backend": b.name,
            "padded_shape": plan.padded,
            "c_ref": plan.c_ref,
            "nonlinear_active": plan.beta2_dt is not None or bool(plan.stg and plan.stg["nonlinear"]),
            "source": source.label,
            "scheme": SCHEME if plan.scheme == "reference" else STAGGERED_SCHEME,
        },
    )
