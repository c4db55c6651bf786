#!/usr/bin/env python3
"""Benchmark of the hip CW k-space path on one MI355X.

Headline workload: the 256^3 heterogeneous linear CW solve -- layered water / skin /
fat / muscle with a seeded 2 % Gaussian random field on c and rho, dx 0.375 mm,
16-cell sponge, bowl ROC 64 mm / aperture 64 mm at 200 kPa, CFL 0.3, 60 periods with
the fundamental phasor from the last 2. One bench step ("--steps") is one whole solve
(900 time steps + the float64 record finalization) with the medium, operators and
source already resident in HBM; ``value`` is solves/s over the timed steps and
``ms_per_step`` milliseconds per solve (``config.ms_per_time_step`` per time step).

The same JSON line carries the per-pass HBM roofline of that solve (algorithmic bytes /
duration from HIP events on the session stream), the 256 x 256 x 384 Westervelt solve,
the Pennes/CEM43 stage heated by the resident phasor, the mixed-size serving queue
(continuous batcher against the serial plain form; the default runs a named subset of
the 512-request benchmark queue, ``--serving-requests 512 --serving-rates 2,8,32`` the
whole one) and the CPU baseline (the float32 NumPy/scipy.fft restatement of the upstream
step on this host: 3 warm-up steps, median of 20 timed steps, scipy.fft workers = 1 and
= the host's cores).
"""

from __future__ import annotations

import argparse
import json
import os
import re
import statistics
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def pass_bytes(padded3, nonlinear: bool) -> dict:
    """HBM bytes per launch each pass must move (every array read once, written once).

    H = nx ny (nz/2 + 1) half-spectrum bins, P = nx ny nz voxels. The y and z derivative
    factors ride in the 3-spectrum y passes, so the x passes carry two spectra:
    xgrad reads p-hat and writes A and Q; the y pass after it reads A, Q and writes three
    gradient spectra; zu reads each gradient spectrum and velocity-state spectrum and
    writes the velocity state back (the velocity lives on the device as its z-spectrum);
    the y pass after zu reads the three velocity spectra and writes U_x-hat and the
    combined y/z divergence terms; xdiv reads those two and writes the divergence
    spectrum."""
    nx, ny, nz = padded3
    P = nx * ny * nz
    H = nx * ny * (nz // 2 + 1)
    return {
        "xgrad": 8 * H + 16 * H,
        "y3": 16 * H + 24 * H,                     # ygrad and ydiv both move 40 H
        "zu": 24 * H + 24 * H + 8 * P + 24 * H,    # gradient spectra, u-hat, dt/rho + absorb, u-hat
        "xdiv": 16 * H + 8 * H,
        "y1": 2 * 8 * H,
        "zp": 8 * H + 4 * P + 8 * P + (4 * P if nonlinear else 0) + 4 * P + 8 * H,
    }


def reference_step_bytes(padded3, nonlinear: bool) -> int:
    """Algorithmic bytes of one step under the reference's pass schedule (every array of
    engine.py's step read and written once: 256 H + 48 P, + 4 P nonlinear); the figure
    the step-level fraction of HBM peak is quoted against."""
    nx, ny, nz = padded3
    P = nx * ny * nz
    H = nx * ny * (nz // 2 + 1)
    return 256 * H + 48 * P + (4 * P if nonlinear else 0)


def record_extra_bytes(padded3, n_h: int) -> int:
    """Extra bytes of a record step: Kahan sum + compensation r/w per harmonic, p_max r/w."""
    nx, ny, nz = padded3
    P = nx * ny * nz
    return 32 * P * n_h + 8 * P


def read_traffic(kernel: str, workload: str):
    """HBM bytes per launch measured with rocprofv3 --pmc (profiles/*traffic*.md), if any."""
    for f in sorted((ROOT / "profiles").glob("*traffic*.md")):
        for line in f.read_text().splitlines():
            m = re.match(r"\s*traffic_bytes_per_launch\s+(\S+)\s+(\S+)\s+([0-9.eE+]+)", line)
            if m and m.group(1) == kernel and m.group(2) == workload:
                return float(m.group(3))
    return None


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_cores() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        return os.cpu_count() or 1


def cpu_step_ms(wl, workers: int, warmup: int = 3, steps: int = 20) -> tuple[float, int]:
    """(median step [ms], spp) of the oracle's float32 restatement of the upstream step on
    the workload's medium and source: ``warmup`` untimed steps, then the median of ``steps``
    individually timed ones (upstream planner/calibration.py:237-255)."""
    from oracle.kspace import KSpaceOracle, RunSpec

    orc = KSpaceOracle(
        shape=wl.grid.shape, dx=wl.grid.dx, pml_vox=wl.grid.pml_vox, pml_edge=wl.grid.pml.edge,
        alpha=wl.medium.alpha, rho=wl.medium.rho, c=wl.medium.c, beta=wl.medium.beta,
        src_idx=wl.source.indices, src_phase=wl.source.phases, amplitude=wl.source.amplitude,
        f0=wl.source.f0, ramp_periods=wl.source.ramp_periods,
        spec=RunSpec(**wl.spec.model_dump()), harmonics=wl.harmonics,
        nonlinear=wl.solver == "westervelt", precision="fp32", workers=workers,
    )
    n = 5 * orc.spp  # past the ramp: every source voxel is driven
    for _ in range(warmup):
        orc.step(n)
        n += 1
    times = []
    for _ in range(steps):
        t0 = time.perf_counter()
        orc.step(n)
        times.append(time.perf_counter() - t0)
        n += 1
    spp = orc.spp
    del orc
    return statistics.median(times) * 1e3, spp


def cpu_pennes_step_ms(shape, dx: float, warmup: int = 3, steps: int = 20) -> float:
    """Per-step time of the oracle's float32 Pennes/CEM43 loop (upstream thermal/pennes.py
    :608-654) on the thermal benchmark's tissue: (t(warmup + steps) - t(warmup)) / steps of
    whole solves, which share their setup (numpy: one core)."""
    import numpy as np

    from caustica_amd import workloads
    from oracle.pennes import PennesOracle

    tm = workloads.pennes_tissue(shape, dx)
    orc = PennesOracle(tm.k, tm.rho, tm.specific_heat, tm.perfusion, dx, precision="fp32")
    t0 = np.full(shape, 37.0, np.float32)
    q = np.zeros(shape, np.float32)
    q[tuple(n // 2 for n in shape)] = 1e7

    def run(k):
        a = time.perf_counter()
        orc.solve(t0, q, 0.05, k, dose=True)
        return time.perf_counter() - a

    short, long_ = run(warmup), run(warmup + steps)
    return max(long_ - short, 1e-9) / steps * 1e3


def cpu_baseline(wl, warmup: int, steps: int, legs: str, log) -> dict:
    """The CPU baseline of the headline workload at scipy.fft workers = host cores (the
    reported value) and workers = 1 (the upstream default); with legs == "all" also the
    64 x 64 x 96 bowl, the Westervelt solve, the Pennes stage and the serving sizes."""
    cores = host_cores()
    per_solve = wl.periods_total
    out_legs = {}
    for w in sorted({cores, 1}, reverse=True):
        log(f"cpu baseline: {wl.name}, workers={w}")
        ms, spp = cpu_step_ms(wl, w, warmup, steps)
        out_legs[f"{wl.name}_workers{w}"] = {"median_step_ms": round(ms, 2), "spp": spp,
                                             "steps_per_solve": per_solve * spp,
                                             "solves_per_s": 1e3 / (ms * per_solve * spp)}
    if legs == "all":
        from caustica_amd import workloads

        extra = [("cfg0", workloads.cfg0_water_bowl()), ("cfg2", workloads.cfg2_westervelt())]
        for tag, w2 in extra:
            for w in sorted({cores, 1}, reverse=True):
                log(f"cpu baseline: {w2.name}, workers={w}")
                ms, spp = cpu_step_ms(w2, w, warmup, steps)
                out_legs[f"{w2.name}_workers{w}"] = {
                    "median_step_ms": round(ms, 2), "spp": spp,
                    "steps_per_solve": w2.periods_total * spp,
                    "solves_per_s": 1e3 / (ms * w2.periods_total * spp)}
        log("cpu baseline: Pennes 256^3")
        pms = cpu_pennes_step_ms((256, 256, 256), 0.375e-3, warmup, steps)
        out_legs["pennes_256cubed"] = {"step_ms": round(pms, 2), "steps": 600,
                                       "stage_s": round(pms * 600 / 1e3, 2), "cores": 1}
    head = out_legs[f"{wl.name}_workers{cores}"]
    return {
        "value": head["solves_per_s"],
        "unit": "solves/s",
        "cores": cores,
        "kind": "port",
        "cpu_model": cpu_model(),
        "sample": (f"{warmup} warm-up + {steps} individually timed steps of the float32 "
                   f"NumPy/scipy.fft restatement of the upstream step on the same medium and "
                   f"source (scipy.fft workers={cores}); median step {head['median_step_ms']} ms "
                   f"x {head['steps_per_solve']} steps per solve"),
        "legs": out_legs,
    }


def profile_passes(sess, one_solve, plan, total_steps) -> tuple[dict, dict]:
    """Per-pass event timing over one profiled solve: (passes, step summary)."""
    sess.set_profiling(True)
    one_solve()
    sess.sync()
    kt = sess.kernel_times()
    sess.set_profiling(False)
    nb = pass_bytes(plan.padded3, plan.beta2_dt is not None)
    passes = {}
    for name, (ms, cnt) in kt.items():
        if cnt == 0 or name not in nb:
            continue
        avg = ms / cnt
        gbs = nb[name] / (avg * 1e-3) / 1e9
        passes[name] = {"launches": cnt, "avg_us": round(avg * 1e3, 2),
                        "bytes_per_launch": nb[name], "GB_s": round(gbs, 1),
                        "frac": round(gbs / HBM_PEAK_GBS, 4)}
    step_ms = sum(passes[k]["avg_us"] * passes[k]["launches"] for k in passes) / 1e3 / total_steps
    moved = sum(nb[k] * passes[k]["launches"] for k in passes) / total_steps
    ref_bytes = reference_step_bytes(plan.padded3, plan.beta2_dt is not None)
    ref_gbs = ref_bytes / (step_ms * 1e-3) / 1e9
    step = {"alg_bytes": ref_bytes, "moved_bytes": round(moved), "kernel_ms": round(step_ms, 4),
            "GB_s": round(ref_gbs, 1), "frac": round(ref_gbs / HBM_PEAK_GBS, 4)}
    return passes, step


def westervelt_leg(log) -> dict:
    """The 256 x 256 x 384 Westervelt solve (dx 0.25 mm, 20-cell sponge, layered medium
    with B/A, 1.0 MPa bowl, CFL 0.2, 80 periods, harmonics 1-3 from the last 2): one
    untimed and one timed solve with the member resident, then a profiled one."""
    import numpy as np

    from caustica_amd import _native as N
    from caustica_amd import workloads
    from caustica_amd.solvers.kspace.native import KSpaceSession
    from caustica_amd.solvers.kspace.operators import plan_member
    from caustica_amd.tune import autotune

    log("westervelt: building the 256x256x384 workload")
    wl = workloads.cfg2_westervelt()
    plan = plan_member(wl.grid, wl.medium, wl.source, wl.spec, harmonics=wl.harmonics,
                       reference_point=wl.reference_point, nonlinear=True)
    spp = plan.spp
    rec = wl.spec.n_record_periods * spp
    total = wl.periods_total * spp
    with KSpaceSession(plan.padded3, 2) as sess:
        choice = autotune(sess, plan, use_cache=False)
        slot = sess.add(plan, max_periods=wl.periods_total + 4)

        def one_solve():
            sess.reset(slot)
            sess.schedule(slot, total - rec)
            sess.advance([slot], total)
            for hi in range(len(plan.harmonics)):
                N.check(N.lib().cst_member_result(sess.handle, slot, hi, 2.0 / rec, None, None))

        one_solve()
        sess.sync()
        t0 = time.perf_counter()
        one_solve()
        sess.sync()
        dt = time.perf_counter() - t0
        peaks = sess.peaks(slot, 0, wl.periods_total - wl.spec.n_record_periods)
        if not np.isfinite(peaks).all():
            raise RuntimeError("Westervelt benchmark produced non-finite values")
        passes, step = profile_passes(sess, one_solve, plan, total)
    return {"workload": wl.name, "grid": list(wl.grid.shape), "spp": spp,
            "steps_per_solve": total, "harmonics": list(wl.harmonics), "variants": choice,
            "solves_per_s": round(1.0 / dt, 4), "ms_per_solve": round(dt * 1e3, 1),
            "step": step, "passes": passes}


def staggered_kind_bytes(padded3, absorbing: bool, nonlinear: bool) -> dict:
    """Algorithmic HBM bytes per time step of the opt-in staggered scheme, per timing kind
    (every array a pass must read or write counted once; H half-spectrum bins, P voxels):
    sxgrad reads p-hat and writes two spectra (24 H); the derivative y passes move 40 H
    (gradient) and 48 H (divergence: three spectra in, three out); szu reads each
    gradient and velocity spectrum and one dt/rho0' map per axis and writes the velocity
    spectra (72 H + 12 P); sxdiv moves 48 H, sxabs 32 H; the plain y passes 48 H (3
    spectra), and with absorption 32 H twice more, plus 16 H for p; szrho reads the three
    divergence spectra, rho_a and dt rho0 (+ 2 m dt) and writes rho_a -- absorption-free
    it also reads c^2 (+ the Westervelt map) and writes p and its spectrum (32 H + 36 P),
    absorbing it writes rho and two spectra (40 H + 32 P) and szabs reads two spectra, rho,
    c^2, c^2 tau / dt, c^2 eta (+ the Westervelt map) and writes p and its spectrum
    (24 H + 20 P)."""
    nx, ny, nz = padded3
    P = nx * ny * nz
    H = nx * ny * (nz // 2 + 1)
    nl = 4 * P if nonlinear else 0
    if absorbing:
        zp = (40 * H + 32 * P + nl) + (24 * H + 20 * P + nl)
        x = 48 * H + 32 * H
        y = 48 * H + 32 * H + 32 * H + 16 * H
    else:
        zp = 32 * H + 36 * P + 2 * nl
        x = 48 * H
        y = 48 * H + 16 * H
    return {"xgrad": 24 * H, "y3": 88 * H, "zu": 72 * H + 12 * P, "xdiv": x, "y1": y, "zp": zp}


def staggered_leg(log) -> dict:
    """The opt-in staggered scheme (no upstream counterpart) on the 256^3 heterogeneous
    medium of the headline workload: power-law absorption with the configuration's
    exponent y = 1.1 (alpha0 from each tissue's absorption at 1 MHz, the drive
    frequency), split-field PML on the 16-cell layer, the same bowl,
    CFL and 900-step schedule. One untimed and one timed solve with the member resident,
    then a profiled one: per timing kind and per step, the algorithmic bytes of
    staggered_kind_bytes over the HIP-event durations."""
    import numpy as np

    from caustica_amd import _native as N
    from caustica_amd import workloads
    from caustica_amd.solvers.kspace.native import KSpaceSession
    from caustica_amd.solvers.kspace.staggered import plan_staggered

    log("staggered: building the 256^3 workload")
    wl = workloads.cfg1_hetero_linear()
    plan = plan_staggered(wl.grid, wl.medium, wl.source, wl.spec, harmonics=wl.harmonics,
                          reference_point=wl.reference_point, alpha_power=1.1)
    spp = plan.spp
    rec = wl.spec.n_record_periods * spp
    total = wl.periods_total * spp
    absorbing, nonlinear = plan.stg["absorbing"], plan.stg["nonlinear"]
    nb = staggered_kind_bytes(plan.padded3, absorbing, nonlinear)
    with KSpaceSession(plan.padded3, 1, scheme="staggered") as sess:
        slot = sess.add(plan, max_periods=wl.periods_total + 4)

        def one_solve():
            sess.reset(slot)
            sess.schedule(slot, total - rec)
            sess.advance([slot], total)
            N.check(N.lib().cst_member_result(sess.handle, slot, 0, 2.0 / rec, None, None))

        one_solve()
        sess.sync()
        t0 = time.perf_counter()
        one_solve()
        sess.sync()
        dt = time.perf_counter() - t0
        peaks = sess.peaks(slot, 0, wl.periods_total - wl.spec.n_record_periods)
        if not np.isfinite(peaks).all():
            raise RuntimeError("staggered benchmark produced non-finite values")
        sess.set_profiling(True)
        one_solve()
        sess.sync()
        kt = sess.kernel_times()
        sess.set_profiling(False)
    kinds = {}
    step_ms = 0.0
    for name, (ms, cnt) in kt.items():
        if cnt == 0 or name not in nb:
            continue
        per_step = ms / total
        step_ms += per_step
        gbs = nb[name] / (per_step * 1e-3) / 1e9
        kinds[name] = {"launches": cnt, "us_per_step": round(per_step * 1e3, 2),
                       "bytes_per_step": nb[name], "GB_s": round(gbs, 1),
                       "frac": round(gbs / HBM_PEAK_GBS, 4)}
    step_bytes = sum(nb.values())
    gbs = step_bytes / (step_ms * 1e-3) / 1e9
    return {"workload": f"staggered_powerlaw_{wl.name}", "grid": list(wl.grid.shape), "spp": spp,
            "steps_per_solve": total, "alpha_power": 1.1, "absorbing": absorbing,
            "solves_per_s": round(1.0 / dt, 4), "ms_per_solve": round(dt * 1e3, 1),
            "ms_per_time_step": round(dt * 1e3 / total, 4),
            "step": {"alg_bytes": step_bytes, "kernel_ms": round(step_ms, 4), "GB_s": round(gbs, 1),
                     "frac": round(gbs / HBM_PEAK_GBS, 4)},
            "kinds": kinds}


def thermal_leg(sess, slot, plan, wl, rec_steps: int, n_heat: int, n_cool: int, log) -> dict:
    """Pennes + CEM43 on the 256^3 grid, heated by the resident 256^3 phasor: the heat
    source is formed on the device (alpha |P|^2 / (rho c)), scaled to 150 W/cm^2 focal
    intensity; dt 0.05 s, n_heat heating steps then n_cool cooling steps, perfused tissue
    (k 0.5, rho C 3.6e6, w_b 0.5 kg/m^3/s), T0 = T_art = 37 C. The tissue's k, perfusion
    and dt/(rho C) maps are uniform, so the solve takes the kernels that hold those values
    in registers; ``map_kernels`` times the same solve reading the maps (the kernels of a
    heterogeneous medium; the same bits). Roofline: each launch (two time steps, or one at
    the end of a 25-step guard interval of odd length) reads and writes every array it
    uses once: theta read and written, Q read (heating only), theta_max and the float64
    dose read and written = 36 P bytes heating and 32 P cooling, plus k, perfusion and
    dt/(rho C) read (12 P) by the map kernels. ``one_step_schedule_GB_s`` charges those
    bytes per time step instead, as a one-step-per-launch schedule would move them."""
    import numpy as np

    from caustica_amd import workloads
    from caustica_amd.sensors import DeviceHeatingSource
    from caustica_amd.thermal import PennesSolver, ThermalResult

    shape = plan.rec_shape
    P = int(np.prod(shape))
    focus = tuple(int(v) for v in wl.reference_point)
    ph = sess.result(slot, 0, 2.0 / rec_steps, with_pmax=False)[0].reshape(shape)
    scale = workloads.focal_intensity_scale(ph[focus], float(wl.medium.rho[focus]),
                                            float(wl.medium.c[focus]))
    src = DeviceHeatingSource.from_member(sess, slot, plan, wl.medium, wl.grid.dx,
                                          rec_steps=rec_steps).scaled(scale)
    med = workloads.pennes_tissue(shape, wl.grid.dx)
    t0 = np.full(shape, 37.0, np.float32)

    def run(uniform_maps: bool) -> dict:
        ps = PennesSolver("hip", profile=True, uniform_maps=uniform_maps)
        w0 = time.perf_counter()
        hot = ps.solve(t0, src, med, dt=0.05, n_steps=n_heat, dose=True)
        cool = ps.solve(hot.temperature, None, med, dt=0.05, n_steps=n_cool, dose=True,
                        dose0=hot.dose_cem43)
        wall = time.perf_counter() - w0
        res = ThermalResult.chain([hot, cool])
        kms = hot.meta["profile"]["kernel_ms"] + cool.meta["profile"]["kernel_ms"]
        launches = hot.meta["profile"]["launches"] + cool.meta["profile"]["launches"]
        maps = 0 if hot.meta["uniform_map_kernels"] else 12
        b_heat, b_cool = 36 + maps, 32 + maps
        one_step = (b_heat * hot.substeps * n_heat + b_cool * cool.substeps * n_cool) * P
        moved = (b_heat * hot.meta["profile"]["launches"] + b_cool * cool.meta["profile"]["launches"]) * P
        gbs = moved / (kms * 1e-3) / 1e9
        if not (np.isfinite(res.temperature_max).all() and np.isfinite(res.dose_cem43).all()):
            raise RuntimeError("thermal benchmark produced non-finite values")
        return {"res": res, "substeps": hot.substeps, "wall_s": round(wall, 3),
                "kernel_ms": round(kms, 3), "launches": launches,
                "uniform_map_kernels": bool(hot.meta["uniform_map_kernels"]),
                "roofline": {"bound": "hbm", "kernel": "k_pennes2", "achieved": round(gbs, 1),
                             "peak": HBM_PEAK_GBS, "unit": "GB/s",
                             "frac": round(gbs / HBM_PEAK_GBS, 4),
                             "alg_bytes_per_launch": {"heating": b_heat * P, "cooling": b_cool * P}},
                "one_step_schedule_GB_s": round(one_step / (kms * 1e-3) / 1e9, 1)}

    log(f"thermal: {n_heat} heating + {n_cool} cooling steps on {shape}")
    main = run(True)
    log("thermal: the same solve on the map-reading kernels")
    alt = run(False)
    res, alt_res = main.pop("res"), alt.pop("res")
    same = all(np.array_equal(getattr(res, n), getattr(alt_res, n))
               for n in ("temperature", "temperature_max", "dose_cem43"))
    return {
        "grid": list(shape), "dt_s": 0.05, "steps": n_heat + n_cool, "substeps": main.pop("substeps"),
        "focal_intensity_w_cm2": workloads.FOCAL_INTENSITY_W_M2 / 1e4, "q_scale": scale,
        "peak_temperature_c": round(float(res.temperature_max.max()), 3),
        "peak_cem43_min": float(res.dose_cem43.max()),
        **main,
        "map_kernels": {**{k: v for k, v in alt.items() if k != "substeps"},
                        "results_bitwise_equal": same},
    }


def serving_leg(n_req: int, rates, seed: int, serial_mode: str, log) -> dict:
    """The mixed-size serving queue: n_req requests drawn as the benchmark queue does
    (sizes uniform over 128^3, 192^3, 256^3, 128 x 128 x 256 at dx 0.375 mm; 70 % linear /
    30 % Westervelt with harmonics 1-2; fresh random-field media; bowl focus steered by
    +-10 mm lateral / +-15 mm axial with delay-and-sum phases), with Poisson arrivals at
    each rate (the same requests, arrival times rescaled). The continuous batcher serves
    them live. The serial plain form (one solve at a time through the solver API: host
    planning, upload, solve, read-back) runs live at every rate (``serial_mode="live"``),
    or once back to back, its measured per-request service times then replayed at each
    rate's arrivals as the single FIFO server it is (``"replay"``: the plain form's
    timeline without waiting out the idle gaps). Results are compared by digest: every
    served result must be bitwise the plain form's."""
    import dataclasses

    import numpy as np

    from caustica_amd import workloads
    from caustica_amd.serving import ContinuousBatcher, replay_fifo, result_digest, serve_serial, summarize
    from caustica_amd.solvers.kspace.native import KSpaceSession
    from caustica_amd.solvers.kspace.operators import plan_member
    from caustica_amd.tune import autotune

    t_gen = time.perf_counter()
    log(f"serving: generating {n_req} requests")
    base = workloads.serving_queue(
        n_req, rates[0], seed=seed,
        progress=lambda i, n: log(f"serving: {i} of {n} requests built") if i % 64 == 0 else None)
    log(f"serving: generated in {time.perf_counter() - t_gen:.1f} s")
    seen = {}
    for r in base:
        seen.setdefault(r.grid.shape, r)
    for r in seen.values():  # tune every shape once; both forms read the same choice
        plan = plan_member(r.grid, r.medium, r.source, r.spec, harmonics=r.harmonics,
                           reference_point=r.reference_point, nonlinear=r.solver == "westervelt")
        with KSpaceSession(plan.padded3, 2) as sess:
            autotune(sess, plan)
        del plan

    from concurrent.futures import ThreadPoolExecutor

    hasher = ThreadPoolExecutor(max_workers=2)  # digests off the serving loops' critical path

    def keep_digest(store, what):
        count = [0]

        def cb(resp):
            count[0] += 1
            if count[0] % 64 == 0:
                log(f"serving: {what}: {count[0]} solves done")
            if resp.result is not None:
                store[resp.rid] = hasher.submit(result_digest, resp.result)
                resp.result = None  # the digest is all the comparison needs
        return cb

    def digests(store):
        return {k: f.result() for k, f in store.items()}

    def pick(d):
        return {"solves_per_s": round(d["solves_per_s"], 3),
                "p50_ms": round(d["latency_p50_s"] * 1e3, 1),
                "p99_ms": round(d["latency_p99_s"] * 1e3, 1),
                "makespan_s": round(d["makespan_s"], 2), "errors": d["errors"]}

    plain_digest: dict = {}
    back = None
    if serial_mode == "replay":
        log(f"serving: serial plain form, {n_req} solves back to back")
        back = serve_serial([dataclasses.replace(r, arrival_s=0.0) for r in base],
                            on_done=keep_digest(plain_digest, "serial back to back"), wait_for_arrivals=False)
    service = None
    runs = []
    for rate in rates:
        reqs = [dataclasses.replace(r, arrival_s=r.arrival_s * rates[0] / rate) for r in base]
        if serial_mode == "replay":
            serial = replay_fifo(back, {r.rid: r.arrival_s for r in reqs})
        else:
            log(f"serving at {rate} req/s: serial plain form")
            plain_digest = {}
            serial = serve_serial(reqs, on_done=keep_digest(plain_digest, f"serial at {rate} req/s"))
        if service is None:
            service = {r.rid: r.done_s - r.admit_s for r in serial if r.error is None}
        log(f"serving at {rate} req/s: continuous batcher")
        served_digest: dict = {}
        b = ContinuousBatcher()
        try:
            served = b.serve(reqs, on_done=keep_digest(served_digest, f"served at {rate} req/s"))
        finally:
            b.close()
        s_ser, s_srv = summarize(serial), summarize(served)
        errs = sorted({r.error for r in served if r.error} | {r.error for r in serial if r.error})
        if errs:
            log(f"serving at {rate} req/s: errors: {errs[:3]}")
        single = [service[r.rid] for r in serial if r.shape == (256, 256, 256) and r.rid in service]
        single_ms = float(np.median(single)) * 1e3 if single else None
        slow = [r.latency_s / service[r.rid] for r in served if r.error is None and r.rid in service]
        pd, sd = digests(plain_digest), digests(served_digest)
        same = bool(pd) and len(sd) == len(pd) and all(sd.get(k) == v for k, v in pd.items())
        runs.append({
            "rate_hz": rate,
            "served": pick(s_srv) | {"max_batch_seen": s_srv["max_batch_seen"],
                                     "host_loop": {k: round(v, 3) for k, v in b.stats.items()},
                                     "p99_latency_over_own_plain_solve": round(float(np.percentile(slow, 99)), 2)
                                     if slow else None},
            "serial_plain": pick(s_ser) | {"timeline": "replayed from measured service times"
                                           if serial_mode == "replay" else "live"},
            "speedup_vs_serial": round(s_srv["solves_per_s"] / max(s_ser["solves_per_s"], 1e-12), 2),
            "single_solve_256_ms": None if single_ms is None else round(single_ms, 1),
            "p99_over_single_256": (None if single_ms is None
                                    else round(s_srv["latency_p99_s"] * 1e3 / single_ms, 2)),
            "results_bitwise_equal": same,
            "error_kinds": [e[:160] for e in errs[:3]],
        })
        log(f"serving: {json.dumps(runs[-1])}")
    hasher.shutdown(wait=True)
    shapes = {}
    for r in base:
        k = "x".join(map(str, r.grid.shape))
        shapes[k] = shapes.get(k, 0) + 1
    return {
        "queue": {"requests": n_req, "of": 512, "seed": seed, "shapes": shapes,
                  "westervelt": sum(r.solver == "westervelt" for r in base),
                  "serial_mode": serial_mode,
                  "mean_plain_service_ms": round(float(np.mean(list(service.values()))) * 1e3, 1)
                  if service else None},
        "runs": runs,
    }


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3, help="timed solves (bench steps)")
    ap.add_argument("--warmup", type=int, default=1, help="untimed solves")
    ap.add_argument("--workload", default="cfg1", choices=["cfg0", "cfg1", "cfg2"])
    ap.add_argument("--batch", type=int, default=1,
                    help="same-shape solves advanced together per launch (value = all solves/s; "
                         "per-pass profile only at 1)")
    ap.add_argument("--cpu-steps", type=int, default=20, help="timed CPU baseline steps (0: skip)")
    ap.add_argument("--cpu-warmup", type=int, default=3)
    ap.add_argument("--cpu-legs", default="headline", choices=["headline", "all"],
                    help="CPU baseline of the headline workload only, or of every configuration")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch every pass instead of replaying captured HIP graphs (the same "
                         "kernels and bits; for tools that cannot trace graph launches)")
    ap.add_argument("--no-westervelt", dest="westervelt", action="store_false",
                    help="skip the 256x256x384 Westervelt leg")
    ap.add_argument("--no-staggered", dest="staggered", action="store_false",
                    help="skip the leg of the opt-in staggered / power-law scheme (256^3)")
    ap.add_argument("--thermal-steps", type=int, default=600,
                    help="Pennes steps heated by the 256^3 phasor (1/3 heating, 2/3 cooling; 0: skip)")
    ap.add_argument("--serving-requests", type=int, default=48,
                    help="requests of the mixed-size serving queue (512 = the whole benchmark "
                         "queue; 0: skip the serving leg)")
    ap.add_argument("--serving-rates", default="32,2", help="Poisson arrival rates [req/s]")
    ap.add_argument("--serving-serial", default="auto", choices=["auto", "live", "replay"],
                    help="serial plain form live at every rate, or its service times measured "
                         "once and replayed (auto: live up to 64 requests)")
    ap.add_argument("--serving-only", action="store_true",
                    help="run the serving leg alone (prints its JSON object)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(
            "bench.py runs on one GPU only: this build targets a single MI355X; independent "
            "solves shard across GPUs with no exchange (one batcher per GPU), so no multi-GPU "
            "figure is produced.",
            file=sys.stderr,
        )
        return 2
    if not 1 <= args.batch <= 16:
        print("bench.py: --batch must be 1..16 (the members one launch advances)", file=sys.stderr)
        return 2

    import numpy as np

    from caustica_amd import _native as N
    from caustica_amd import workloads
    from caustica_amd.core.backend import get_backend
    from caustica_amd.solvers.kspace.native import KSpaceSession
    from caustica_amd.solvers.kspace.operators import plan_member

    def log(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    get_backend("hip")
    if args.no_graphs:
        KSpaceSession.graphs_default = False
    rates = [float(v) for v in args.serving_rates.split(",") if v]
    serial_mode = args.serving_serial
    if serial_mode == "auto":
        serial_mode = "live" if args.serving_requests <= 64 else "replay"
    if args.serving_only:
        out = serving_leg(args.serving_requests, rates, 4, serial_mode, log)
        print(json.dumps({"serving": out}))
        return 0

    log(f"building workload {args.workload}")
    wl = {"cfg0": workloads.cfg0_water_bowl, "cfg1": workloads.cfg1_hetero_linear,
          "cfg2": workloads.cfg2_westervelt}[args.workload]()
    plan = plan_member(wl.grid, wl.medium, wl.source, wl.spec, harmonics=wl.harmonics,
                       reference_point=wl.reference_point, nonlinear=wl.solver == "westervelt")
    spp = plan.spp
    rec_steps = wl.spec.n_record_periods * spp
    total_steps = wl.periods_total * spp
    record_from = total_steps - rec_steps
    sess = KSpaceSession(plan.padded3, max(2, args.batch))
    from caustica_amd.tune import autotune

    choice = autotune(sess, plan, use_cache=False)
    log(f"autotuned variants {choice}")
    slots = [sess.add(plan, max_periods=wl.periods_total + 4) for _ in range(args.batch)]
    slot = slots[0]

    def one_solve():
        """Fresh solve of the resident member(s), all advanced by the same launches; the
        float64 phasor finalization runs on the device, the result stays in HBM (no host
        copy inside the timed region)."""
        for s_ in slots:
            sess.reset(s_)
            sess.schedule(s_, record_from)
        sess.advance(slots, total_steps)
        for s_ in slots:
            for hi in range(len(plan.harmonics)):
                N.check(N.lib().cst_member_result(sess.handle, s_, hi, 2.0 / rec_steps, None, None))

    log(f"{wl.name}: spp {spp}, {total_steps} steps/solve; warmup {args.warmup}, timed {args.steps}")
    for _ in range(args.warmup):
        one_solve()
    sess.sync()
    lat = []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        t0 = time.perf_counter()
        one_solve()
        sess.sync()
        lat.append(time.perf_counter() - t0)
    sess.sync()
    elapsed = time.perf_counter() - t_start
    value = args.batch * args.steps / elapsed
    peaks = sess.peaks(slot, 0, wl.periods_total - wl.spec.n_record_periods)
    if not np.isfinite(peaks).all():
        print("bench: non-finite field", file=sys.stderr)
        return 3
    # the same solve with every pass launched directly (no graph replay), for the record
    sess.set_graphs(False)
    t0 = time.perf_counter()
    one_solve()
    sess.sync()
    direct_ms = (time.perf_counter() - t0) * 1e3
    sess.set_graphs(KSpaceSession.graphs_default)
    graphs = {"replayed": sess.graph_stats()["enabled"],
              "ms_per_solve_graphs": round(elapsed / args.steps * 1e3, 2),
              "ms_per_solve_direct": round(direct_ms, 2)}

    passes = {}
    roofline = None
    if not args.no_profile and args.batch == 1:
        passes, step = profile_passes(sess, one_solve, plan, total_steps)
        dom = max(passes, key=lambda k: passes[k]["avg_us"] * passes[k]["launches"])
        d = passes[dom]
        roofline = {
            "bound": "hbm",
            "kernel": dom,
            "achieved": d["GB_s"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": d["frac"],
            "traffic": read_traffic(dom, wl.name),
            "step": step,
        }

    log(f"gpu: {value:.3f} solves/s; passes {json.dumps(passes)}")
    westervelt = None
    if args.westervelt and args.workload == "cfg1":
        westervelt = westervelt_leg(log)
        log(f"westervelt: {json.dumps(westervelt)}")
    staggered = None
    if args.staggered and args.workload == "cfg1":
        staggered = staggered_leg(log)
        log(f"staggered: {json.dumps(staggered)}")
    thermal = None
    if args.thermal_steps > 0 and args.workload == "cfg1":
        n_heat = max(1, args.thermal_steps // 3)
        thermal = thermal_leg(sess, slot, plan, wl, rec_steps, n_heat, args.thermal_steps - n_heat, log)
        log(f"thermal: {json.dumps(thermal)}")
    sess.close()
    serving = None
    if args.serving_requests > 0:
        serving = serving_leg(args.serving_requests, rates, 4, serial_mode, log)
    cpu = None
    if args.cpu_steps > 0:
        cpu = cpu_baseline(wl, args.cpu_warmup, args.cpu_steps, args.cpu_legs, log)

    lat_ms = sorted(x * 1e3 for x in lat)
    ms_solve = elapsed / args.steps * 1e3
    out = {
        "metric": "CW solves/s + p50/p99 latency (mixed-size queue) vs serial plain; % HBM peak",
        "value": round(value, 4),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_solve, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded layered + Gaussian-random-field medium, voxelized bowl)",
        "config": {
            "workload": wl.name,
            "bench_step": (f"one whole CW solve of {total_steps} time steps + float64 phasor "
                           f"finalization, maps and source resident in HBM"),
            "solver": wl.solver,
            "grid": list(wl.grid.shape),
            "periods": wl.periods_total,
            "spp": spp,
            "steps_per_solve": total_steps,
            "ms_per_time_step": round(ms_solve / total_steps, 4),
            "harmonics": list(wl.harmonics),
            "batch": args.batch,
            "serving_queue": (None if serving is None else
                              f"{serving['queue']['requests']} of the 512 benchmark requests at "
                              f"{','.join(str(r) for r in rates)} req/s"),
        },
        "variants": {str(k): v for k, v in choice.items()},
        "latency_ms": {"p50": round(float(np.percentile(lat_ms, 50)), 2),
                       "p99": round(float(np.percentile(lat_ms, 99)), 2)},
        "graphs": graphs,
        "roofline": roofline,
        "passes": passes,
        "westervelt": westervelt,
        "staggered": staggered,
        "thermal": thermal,
        "serving": serving,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
