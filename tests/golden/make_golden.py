"""Write the golden fixtures under tests/golden/ from the upstream caustica package.

Run once, in a container that holds the upstream source tree, as

    python tests/golden/make_golden.py /root/reference/src

The upstream package is pure Python (numpy + scipy.fft); this script imports it
from the given source directory, runs small cases of the functions whose exact
output nothing else pins (the float32 step loop of the CW k-space engine, the
bowl voxelizer, the Pennes/CEM43 loop, the operator arrays), and stores inputs
and outputs as small .npz files. Nothing under tests/ imports the upstream
package at test time: the tests read these arrays only.

Fixtures (all float32/complex64 as the upstream path produces them):
  ops_*.npz        derivative factors, kappa, sponge for small shapes
  bowl_*.npz       bowl voxel index sets for the three benchmark geometries
  cw_*.npz         whole CW solves (phasors, p_max, schedule metadata)
  pennes_*.npz     Pennes + CEM43 solves (final T, T_max, dose)
  heat_*.npz       HeatingSource.from_phasors Q maps
  array_*.npz      spiral array geometry, delay-and-sum phases and voxelized sources

``--only arrays`` writes the array fixtures alone.
"""

from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent


def _save(name: str, **arrays) -> None:
    path = HERE / f"{name}.npz"
    np.savez_compressed(path, **arrays)
    print(f"wrote {path.name}: {path.stat().st_size / 1024:.1f} KiB")


def _meta(d: dict) -> np.ndarray:
    return np.frombuffer(json.dumps(d, sort_keys=True).encode(), dtype=np.uint8)


def main(src_dir: str) -> None:
    sys.path.insert(0, src_dir)
    import caustica.solvers as solvers
    from caustica import Grid, Medium, PMLSpec
    from caustica.materials import water
    from caustica.sensors import HeatingSource
    from caustica.solvers import CWRunSpec
    from caustica.solvers.kspace import operators as ops
    from caustica.sources import CWSource, bowl_cw_source, plane_cw_source
    from caustica.thermal import PennesSolver, ThermalMedium

    # ---------------- operator arrays ----------------
    for shape, dx, c_ref, dt in (
        ((8, 8, 8), 1e-3, 1500.0, 1e-7),
        ((9, 8, 7), 0.5e-3, 1600.0, 5e-8),
        ((8, 8, 6), 0.375e-3, 1550.0, 6.25e-8),
        ((12,), 0.375e-3, 1500.0, 1.25e-7),
        ((10, 6), 0.375e-3, 1500.0, 1.25e-7),
    ):
        ks = ops.k_vectors(shape, dx, np)
        kappa = ops.kappa_sinc(ks, c_ref=c_ref, dt=dt, xp=np)
        deriv = ops.spectral_derivative_factors(ks, kappa, shape, np)
        sponge = ops.sponge_volume(shape, 2, 2.0, np)
        tag = "x".join(map(str, shape))
        _save(
            f"ops_{tag}",
            shape=np.array(shape),
            dx=np.float64(dx),
            c_ref=np.float64(c_ref),
            dt=np.float64(dt),
            kappa=np.ascontiguousarray(np.broadcast_to(kappa, kappa.shape)),
            **{f"deriv{i}": np.ascontiguousarray(d) for i, d in enumerate(deriv)},
            sponge=sponge,
        )

    # ---------------- bowl voxel sets (benchmark geometries) ----------------
    bowls = {
        # 64x64x96 @ 0.375 mm, ROC 20 mm, aperture 16 mm, apex on the z = PML plane
        "bowl_cfg0": ((64, 64, 96), 0.375e-3, 10, 8e-3, 20e-3),
        # 256^3 @ 0.375 mm, ROC 64 mm, aperture 64 mm
        "bowl_cfg1": ((256, 256, 256), 0.375e-3, 16, 32e-3, 64e-3),
        # 256x256x384 @ 0.25 mm, ROC 64 mm, aperture shrunk by one voxel to fit
        "bowl_cfg2": ((256, 256, 384), 0.25e-3, 20, 31.75e-3, 64e-3),
    }
    for name, (shape, dx, pml_vox, a, roc) in bowls.items():
        grid = Grid(shape, dx, PMLSpec(thickness=pml_vox * dx))
        apex = (shape[0] // 2, shape[1] // 2, grid.pml_vox)
        src = bowl_cw_source(grid, 1e6, 1e5, aperture_radius=a, roc=roc, apex_vox=apex)
        _save(
            name,
            indices=src.indices.astype(np.int32),
            shape=np.array(shape),
            dx=np.float64(dx),
            pml_vox=np.int64(grid.pml_vox),
            aperture_radius=np.float64(a),
            roc=np.float64(roc),
            apex=np.array(apex),
        )

    rng = np.random.default_rng(20261203)

    def run(name, solver, grid, medium, source, spec, **kw):
        res = solvers.get(solver)().run(grid, medium, source, spec, backend="numpy", **kw)
        meta = {
            "solver": solver,
            "shape": list(grid.shape),
            "dx": grid.dx,
            "pml_thickness": grid.pml.thickness if grid.pml is not None else None,
            "pml_edge": grid.pml.edge if grid.pml is not None else None,
            "spec": spec.model_dump(),
            "f0": source.f0,
            "amplitude": source.amplitude,
            "ramp_periods": source.ramp_periods,
            "harmonics": sorted(res.phasors),
            "record_region": [[s.start, s.stop] for s in res.region],
            "reference_point": list(kw["reference_point"]) if kw.get("reference_point") else None,
            "dt": res.dt,
            "spp": res.spp,
            "steps_total": res.steps_total,
            "tof_periods": res.tof_periods,
            "converged_period": res.converged_period,
            "settle_capped": res.settle_capped,
            "history": [[int(p), float(pk), float(r)] for p, pk, r in res.convergence_history],
            "nonlinear_active": res.meta["nonlinear_active"],
            "c_ref": res.meta["c_ref"],
            "padded_shape": list(res.meta["padded_shape"]),
        }
        arrays = {
            "alpha": medium.alpha,
            "rho": medium.rho,
            "c": medium.c,
            "beta": medium.beta,
            "src_indices": source.indices.astype(np.int32),
            "src_phases": source.phases,
            "p_max": res.p_max,
            "meta_json": _meta(meta),
        }
        for h, ph in res.phasors.items():
            arrays[f"phasor{h}"] = ph
        _save(name, **arrays)
        return res

    # F1: 1-D plane wave, linear, lossless, convergence-controlled schedule
    dx = 1500.0 / (1e6 * 4.0)
    g = Grid((96,), dx, PMLSpec(thickness=12 * dx))
    m = Medium.homogeneous(g.shape, water(c=1500.0, alpha_np_m=5.0))
    s = plane_cw_source(g, f0=1e6, amplitude=1e5, position_vox=20)
    run(
        "cw_1d_linear",
        "linear",
        g,
        m,
        s,
        CWRunSpec(min_settle_periods=4, max_settle_periods=30, convergence_tol=0.02),
    )

    # F2: 1-D Westervelt, harmonics (1, 2, 3), pinned schedule
    dx = 1500.0 / (1e6 * 16.0)
    g = Grid((128,), dx, PMLSpec(thickness=16 * dx))
    m = Medium.homogeneous(g.shape, water(c=1500.0, beta=3.5, alpha_np_m=2.0))
    s = plane_cw_source(g, f0=1e6, amplitude=2e6, position_vox=24)
    run(
        "cw_1d_westervelt",
        "westervelt",
        g,
        m,
        s,
        CWRunSpec(min_settle_periods=6, max_settle_periods=6, convergence_tol=0.0),
        harmonics=(1, 2, 3),
    )

    # F3: 2-D heterogeneous (odd and even axes), alpha != 0, fast inclusion, phases
    dx = 0.375e-3
    shape = (48, 40)
    g = Grid(shape, dx, PMLSpec(thickness=6 * dx))
    c = np.full(shape, 1500.0, np.float32) * (1 + 0.02 * rng.standard_normal(shape)).astype(
        np.float32
    )
    c[30:38, 10:20] = 1800.0
    rho = (np.full(shape, 1000.0) * (1 + 0.02 * rng.standard_normal(shape))).astype(np.float32)
    alpha = np.full(shape, 3.0, np.float32)
    alpha[:, 25:] = 9.0
    m = Medium(alpha=alpha, rho=rho, c=c, beta=np.zeros(shape, np.float32))
    idx = np.array([[10, j] for j in range(8, 32)], np.int64)
    ph = (0.1 * np.arange(idx.shape[0])).astype(np.float32)
    s = CWSource(indices=idx, phases=ph, amplitude=5e4, f0=1e6, ramp_periods=2.0)
    run(
        "cw_2d_hetero",
        "linear",
        g,
        m,
        s,
        CWRunSpec(cfl=0.3, min_settle_periods=3, max_settle_periods=3, convergence_tol=0.0),
        record_region=(slice(8, 40), slice(4, 36)),
    )

    # F4: 3-D bowl, linear water, record region + reference point
    dx = 0.375e-3
    shape = (24, 24, 32)
    g = Grid(shape, dx, PMLSpec(thickness=4 * dx))
    m = Medium.homogeneous(shape, water(c=1500.0, alpha_np_m=2.533))
    s = bowl_cw_source(g, 1e6, 1e5, aperture_radius=3e-3, roc=6e-3, apex_vox=(12, 12, 5))
    run(
        "cw_3d_bowl",
        "linear",
        g,
        m,
        s,
        CWRunSpec(cfl=0.3, min_settle_periods=2, max_settle_periods=2, n_record_periods=2,
                  convergence_tol=0.0),
        record_region=(slice(2, 22), slice(4, 20), slice(3, 30)),
        reference_point=(12, 12, 21),
    )

    # F5: 3-D heterogeneous Westervelt (layered + random), harmonics 1-3, full record
    shape = (20, 24, 32)
    g = Grid(shape, dx, PMLSpec(thickness=4 * dx))
    z = np.arange(shape[2])
    c = np.where(z < 14, 1500.0, np.where(z < 18, 1450.0, 1580.0)).astype(np.float32)
    c = np.broadcast_to(c, shape) * (1 + 0.02 * rng.standard_normal(shape))
    rho = np.where(z < 14, 1000.0, np.where(z < 18, 950.0, 1050.0))
    rho = np.broadcast_to(rho, shape) * (1 + 0.02 * rng.standard_normal(shape))
    alpha = np.broadcast_to(np.where(z < 14, 0.0253, np.where(z < 18, 6.9, 8.06)), shape)
    beta = np.broadcast_to(np.where(z < 14, 3.5, np.where(z < 18, 4.25, 4.75)), shape)
    m = Medium(
        alpha=alpha.astype(np.float32),
        rho=rho.astype(np.float32),
        c=c.astype(np.float32),
        beta=beta.astype(np.float32),
    )
    s = bowl_cw_source(g, 1e6, 1e6, aperture_radius=3e-3, roc=6e-3, apex_vox=(10, 12, 5))
    run(
        "cw_3d_westervelt",
        "westervelt",
        g,
        m,
        s,
        CWRunSpec(cfl=0.25, min_settle_periods=2, max_settle_periods=2, convergence_tol=0.0,
                  n_record_periods=1),
        harmonics=(1, 2, 3),
    )

    # F6: 3-D shape that needs padding (22 -> 24, 21 -> 24, 30 stays): edge replication
    shape = (22, 21, 30)
    g = Grid(shape, dx, PMLSpec(thickness=3 * dx))
    c = (1500.0 * (1 + 0.03 * rng.standard_normal(shape))).astype(np.float32)
    m = Medium(
        alpha=np.full(shape, 1.0, np.float32),
        rho=np.full(shape, 1000.0, np.float32),
        c=c,
        beta=np.zeros(shape, np.float32),
    )
    s = bowl_cw_source(g, 1e6, 1e5, aperture_radius=3e-3, roc=6e-3, apex_vox=(11, 10, 4))
    run(
        "cw_3d_padded",
        "linear",
        g,
        m,
        s,
        CWRunSpec(cfl=0.3, min_settle_periods=1, max_settle_periods=1, convergence_tol=0.0),
    )

    # ---------------- Pennes + CEM43 ----------------
    shape = (16, 16, 16)
    dxt = 0.5e-3
    k = np.full(shape, 0.5, np.float32)
    k[:, :, 8:] = 0.2
    rho_t = np.full(shape, 1050.0, np.float32)
    cp = np.full(shape, 3600.0, np.float32)
    perf = np.full(shape, 0.5 / 1050.0, np.float32)
    perf[:4] = 0.0
    tm = ThermalMedium(k=k, rho=rho_t, specific_heat=cp, perfusion=perf, dx=dxt)
    xx = np.indices(shape).astype(np.float64) - 7.5
    q = (4e7 * np.exp(-(xx**2).sum(0) / (2 * 2.0**2))).astype(np.float32)
    t0 = np.full(shape, 37.0, np.float32)
    for name, boundary, sub in (
        ("pennes_insulated", "insulated", "refuse"),
        ("pennes_dirichlet_substep", "dirichlet", "substep"),
    ):
        solver = PennesSolver("numpy", boundary=boundary, on_unstable=sub)
        dt = 0.05 if sub == "refuse" else 0.4
        hot = solver.solve(t0, q, tm, dt=dt, n_steps=40, dose=True)
        cool = solver.solve(
            hot.temperature, None, tm, dt=dt, n_steps=20, dose=True, dose0=hot.dose_cem43
        )
        chained = type(hot).chain([hot, cool])
        _save(
            name,
            k=k,
            rho=rho_t,
            specific_heat=cp,
            perfusion=perf,
            q=q,
            temperature0=t0,
            dt=np.float64(dt),
            dx=np.float64(dxt),
            hot_temperature=hot.temperature,
            hot_temperature_max=hot.temperature_max,
            hot_dose=hot.dose_cem43,
            cool_temperature=cool.temperature,
            cool_dose=cool.dose_cem43,
            chained_temperature_max=chained.temperature_max,
            substeps=np.int64(hot.substeps),
            dt_stable=np.float64(hot.meta["dt_stable_s"]),
            meta_json=_meta({"boundary": boundary, "on_unstable": sub, "n_hot": 40,
                             "n_cool": 20}),
        )

    # ---------------- HeatingSource.from_phasors ----------------
    shape = (6, 7, 8)
    p1 = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64) * 1e5
    p2 = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64) * 2e4
    alpha = rng.uniform(0.1, 9.0, shape).astype(np.float32)
    rho = rng.uniform(950, 1100, shape).astype(np.float32)
    cc = rng.uniform(1450, 1600, shape).astype(np.float32)
    hs = HeatingSource.from_phasors(
        {1: p1, 2: p2}, alpha=alpha, rho=rho, c=cc, dx=1e-3, alpha_power_law_y=1.1
    )
    _save("heat_power_law", p1=p1, p2=p2, alpha=alpha, rho=rho, c=cc, y=np.float64(1.1), q=hs.q)


def arrays_fixtures(src_dir: str) -> None:
    """Spiral geometry (production 128-element and a compact 32-element array), DAS
    phases towards steered targets, and voxelized sources (indices, owning element,
    per-voxel phase) on the grids the array tests use."""
    if src_dir not in sys.path:
        sys.path.insert(0, src_dir)
    from caustica import Grid, PMLSpec
    from caustica.arrays import archimedean_spiral

    f0, c0 = 1.0e6, 1500.0
    prod_arr = archimedean_spiral()
    small = archimedean_spiral(n_elements=32, d_outer=0.030, d_inner=0.010, roc=0.030,
                               active_fraction=0.60)
    odd = archimedean_spiral(n_elements=7, d_outer=0.024, d_inner=0.006, roc=0.020,
                             active_fraction=0.3, n_arc_samples=257)
    targets = np.array([[0.0, 0.0, 0.100], [3e-3, -2e-3, 0.096], [-7e-3, 5e-3, 0.112]])
    das = np.stack([prod_arr.das_phases(t, f0=f0, c0=c0) for t in targets])
    small_target = np.array([3e-3, -2e-3, small.focal_length - 4e-3])
    small_das = small.das_phases(small_target, f0=f0, c0=c0)

    g_small = Grid(shape=(96, 96, 96), dx=0.5e-3, pml=PMLSpec(thickness=5e-3))
    v0 = small.voxelize(g_small, apex_vox=(48, 48, 12), f0=f0, amplitude=1e5)
    v1 = small.voxelize(g_small, apex_vox=(47, 50, 10), f0=f0, amplitude=2e5, phases=small_das)
    dxp = 0.375e-3
    g_prod = Grid(shape=(288, 288, 64), dx=dxp, pml=PMLSpec(thickness=6 * dxp))
    v2 = prod_arr.voxelize(g_prod, apex_vox=(144, 144, 8), f0=f0, amplitude=1e5, phases=das[1])
    _save("array_spiral",
          prod_positions=prod_arr.positions, prod_normals=prod_arr.normals,
          prod_elem_radius=np.float64(prod_arr.elem_radius),
          small_positions=small.positions, small_normals=small.normals,
          small_elem_radius=np.float64(small.elem_radius),
          odd_positions=odd.positions, odd_normals=odd.normals,
          odd_elem_radius=np.float64(odd.elem_radius),
          targets=targets, prod_das=das, small_target=small_target, small_das=small_das)
    _save("array_voxelize",
          small_idx=v0.source.indices.astype(np.int16), small_elem=v0.element_of_voxel,
          small_phase=v0.source.phases,
          steer_idx=v1.source.indices.astype(np.int16), steer_elem=v1.element_of_voxel,
          steer_phase=v1.source.phases,
          prod_idx=v2.source.indices.astype(np.int16), prod_elem=v2.element_of_voxel,
          prod_phase=v2.source.phases,
          meta_json=_meta({"small": {"shape": [96, 96, 96], "dx": 0.5e-3, "apex": [48, 48, 12],
                                     "label": v0.source.label},
                           "steer": {"apex": [47, 50, 10]},
                           "prod": {"shape": [288, 288, 64], "dx": dxp, "apex": [144, 144, 8],
                                    "label": v2.source.label}}))


if __name__ == "__main__":
    args = [a for a in sys.argv[1:] if not a.startswith("--") and a != "arrays"]
    src = args[0] if args else "/root/reference/src"
    if "--only" in sys.argv and sys.argv[sys.argv.index("--only") + 1] == "arrays":
        arrays_fixtures(src)
    else:
        main(src)
        arrays_fixtures(src)
