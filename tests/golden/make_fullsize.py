"""Write the full-size float64 oracle fixtures (tests/golden/full_*.npz).

Run in the build container (CPU only, no upstream import), as

    python tests/golden/make_fullsize.py cfg1 [--workers 4]
    python tests/golden/make_fullsize.py cfg2 [--workers 4]
    python tests/golden/make_fullsize.py serving_w128x128x256 [--workers 4]
    python tests/golden/make_fullsize.py serving_l192 [--workers 4]

cfg1: the 256^3 heterogeneous linear solve (seeded layered + 2 % random-field medium,
      900 steps at spp 15), fundamental phasor;
cfg2: the 256 x 256 x 384 Westervelt solve (2720 steps at spp 34), harmonics 1-3.
serving_*: two requests of the mixed-size serving queue (steered bowl, fresh random-field
      seed, pinned schedule): Westervelt 128 x 128 x 256 with harmonics 1-2 (975 steps)
      and linear 192^3 (780 steps).
The cfg1 run also writes full_cfg3.npz: the Pennes + CEM43 stage on the 256^3 grid,
heated by the float64 fundamental phasor (Q = alpha |P|^2 / (rho c) scaled to 150 W/cm^2
focal intensity, perfused tissue, dt 0.05 s, 200 heating + 400 cooling steps), by the
float64 Pennes oracle: samples of the final temperature, its maximum and the dose.

Each fixture holds the medium's digest and strided samples of c and rho (the GPU test
rebuilds the medium from its seed and checks these first), the schedule the oracle ran
(spp, dt, step count, time of flight, period peaks) and samples of every phasor and of
p_max (tests/_full.py: focal planes at stride 2, the volume at stride 8, a full-resolution
box around the focus; global and focal peak values). The float64
oracle takes about 1.7 s per step at 256^3 and 2.8 s at 256 x 256 x 384 on four cores.
"""

from __future__ import annotations

import argparse
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from caustica_amd import workloads  # noqa: E402
from tests import _full  # noqa: E402


# two requests of the mixed-size serving queue (workloads.serving_workload): a steered
# Westervelt 128 x 128 x 256 solve with harmonics 1-2 and a steered linear 192^3 solve
SERVING_CASES = {
    "serving_w128x128x256": (41, (128, 128, 256), "westervelt"),
    "serving_l192": (42, (192, 192, 192), "linear"),
}


def build(case: str):
    if case == "cfg1":
        return workloads.cfg1_hetero_linear(), dict(seed=1)
    if case == "cfg2":
        return workloads.cfg2_westervelt(), dict(seed=2)
    if case in SERVING_CASES:
        seed, shape, solver = SERVING_CASES[case]
        return workloads.serving_workload(seed, shape, solver), dict(seed=seed)
    raise SystemExit(f"unknown case {case!r}")


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("case", choices=("cfg1", "cfg2", *SERVING_CASES))
    ap.add_argument("--workers", type=int, default=4)
    args = ap.parse_args()
    t0 = time.time()
    wl, seed = build(args.case)
    orc = _full.oracle_for(wl, "fp64", args.workers)
    spp = orc.spp
    total = (orc.eff_max + wl.spec.n_record_periods) * spp
    print(f"[{args.case}] {wl.name}: spp {spp}, {total} steps, setup {time.time() - t0:.1f} s",
          flush=True)
    step = orc.step

    def step_logged(n):
        step(n)
        if (n + 1) % spp == 0:
            el = time.time() - t0
            print(f"[{args.case}] period {(n + 1) // spp} of {total // spp}  "
                  f"max|p| {np.abs(orc.p).max():.4e}  {el:.0f} s", flush=True)

    orc.step = step_logged
    res = orc.run()
    assert res.steps_total == total
    meta = dict(
        workload=wl.name, solver=wl.solver, shape=list(wl.grid.shape), dx=wl.grid.dx, **seed,
        focus=[int(v) for v in wl.reference_point], harmonics=list(wl.harmonics),
        periods_total=wl.periods_total, spp=res.spp, dt=res.dt, steps_total=res.steps_total,
        tof_periods=res.tof_periods, converged_period=res.converged_period,
        settle_capped=res.settle_capped, c_ref=res.c_ref,
        medium_sha256=_full.medium_digest(wl.medium), plane_stride=_full.PLANE_STRIDE,
        volume_stride=_full.VOLUME_STRIDE, oracle="fp64", workers=args.workers,
        cpu_seconds_wall=round(time.time() - t0, 1),
    )
    arrays = dict(_full.medium_sample(wl.medium))
    arrays["history"] = np.array([(p, pk) for p, pk, _ in res.history], np.float64)
    meta["peaks"] = {}
    for h in wl.harmonics:
        arrays.update(_full.field_samples(res.phasors[h].astype(np.complex64), wl.reference_point,
                                          f"ph{h}"))
        meta["peaks"][f"ph{h}"] = _full.peak_info(res.phasors[h], wl.reference_point)
    arrays.update(_full.field_samples(res.p_max.astype(np.float32), wl.reference_point, "pmax"))
    meta["peaks"]["pmax"] = _full.peak_info(res.p_max, wl.reference_point)
    path = _full.save(f"full_{args.case}", meta, **arrays)
    print(f"wrote {path} ({path.stat().st_size / 1024:.0f} KiB) after {time.time() - t0:.0f} s",
          flush=True)
    if args.case == "cfg1":
        thermal_cfg3(wl, res, meta, t0)


def thermal_cfg3(wl, res, acoustic_meta: dict, t0: float) -> None:
    from oracle.heating import heat_source
    from oracle.pennes import PennesOracle

    focus = tuple(int(v) for v in wl.reference_point)
    p1 = res.phasors[1]
    m = wl.medium
    scale = workloads.focal_intensity_scale(p1[focus], float(m.rho[focus]), float(m.c[focus]))
    q = heat_source({1: p1}, m.alpha, m.rho, m.c) * scale
    tm = workloads.pennes_tissue(p1.shape, wl.grid.dx)
    orc = PennesOracle(tm.k, tm.rho, tm.specific_heat, tm.perfusion, wl.grid.dx, precision="fp64")
    n_heat, n_cool, dt = 200, 400, 0.05
    print(f"[cfg3] Pennes {n_heat} + {n_cool} steps, q scale {scale:.6e}", flush=True)
    t_init = np.full(p1.shape, 37.0, np.float32)
    hot = orc.solve(t_init, q, dt, n_heat, dose=True)
    print(f"[cfg3] heating done, T max {hot.temperature_max.max():.4f} C, "
          f"{time.time() - t0:.0f} s", flush=True)
    cool = orc.solve(hot.temperature, None, dt, n_cool, dose=True, dose0=hot.dose)
    tmax = np.maximum(hot.temperature_max, cool.temperature_max)
    meta = dict(workload="cfg3_pennes_cem43_256x256x256", acoustic=acoustic_meta["workload"],
                acoustic_medium_sha256=acoustic_meta["medium_sha256"], focus=list(focus),
                q_scale=scale, dt=dt, n_heat=n_heat, n_cool=n_cool, substeps=hot.substeps,
                oracle="fp64", cpu_seconds_wall=round(time.time() - t0, 1), peaks={})
    arrays = {}
    for tag, f in (("temp", cool.temperature), ("tmax", tmax), ("dose", cool.dose)):
        arrays.update(_full.field_samples(f.astype(np.float32), focus, tag))
        meta["peaks"][tag] = _full.peak_info(f, focus)
    path = _full.save("full_cfg3", meta, **arrays)
    print(f"wrote {path} ({path.stat().st_size / 1024:.0f} KiB) after {time.time() - t0:.0f} s",
          flush=True)


if __name__ == "__main__":
    main()
