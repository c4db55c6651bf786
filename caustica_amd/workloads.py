"""Synthetic CW workloads: the benchmark configurations as (grid, medium, source, spec).

Media are generated, not loaded: layered water / skin / fat / muscle along +z with a
seeded Gaussian random field on c and rho (unit-variance white noise from numpy's
PCG64 ``default_rng(seed)``, convolved with a Gaussian kernel of standard deviation
corr_len / sqrt(2) so that the field's correlation function is a Gaussian of
standard deviation corr_len, then renormalized to unit variance).

Layer boundaries (z measured from the grid origin): water < 40 mm, skin 40-42 mm,
fat 42-50 mm, muscle >= 50 mm. Absorption alpha0 f^y in dB/cm/MHz^y is evaluated at
f0 once (the solver's absorption is frequency independent, as upstream). Skin's B/A
is not given by the benchmark definition; 6.5 is used.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from caustica_amd.core.grid import Grid
from caustica_amd.core.pml import PMLSpec
from caustica_amd.materials import beta_from_bona, power_law_np_m
from caustica_amd.medium import Medium
from caustica_amd.solvers.base import CWRunSpec
from caustica_amd.solvers.kspace.operators import cw_tof_periods
from caustica_amd.sources import bowl_cw_source, focus_phases

F0 = 1.0e6

# name: (c, rho, alpha0 [dB/cm/MHz^y], B/A), boundaries in mm
LAYERS = {
    "water": (1500.0, 1000.0, 0.0022, 5.0),
    "skin": (1610.0, 1090.0, 0.8, 6.5),
    "fat": (1450.0, 950.0, 0.6, 6.5),
    "muscle": (1580.0, 1050.0, 0.7, 7.5),
}
LAYER_Z_MM = (40.0, 42.0, 50.0)


@dataclass
class Workload:
    name: str
    solver: str
    grid: Grid
    medium: Medium
    source: object
    spec: CWRunSpec
    harmonics: tuple
    periods_total: int
    record_region: tuple | None = None
    reference_point: tuple | None = None


def gaussian_random_field(shape, dx: float, corr_len: float, seed: int) -> np.ndarray:
    """Unit-variance GRF with Gaussian correlation of standard deviation corr_len (float32)."""
    import os

    import scipy.fft as sfft

    rng = np.random.default_rng(seed)
    noise = rng.standard_normal(shape, dtype=np.float32)
    sigma_vox = corr_len / np.sqrt(2.0) / dx
    workers = min(16, os.cpu_count() or 1)
    spec = sfft.rfftn(noise, workers=workers)
    for ax, n in enumerate(shape):
        k = np.fft.rfftfreq(n) if ax == len(shape) - 1 else np.fft.fftfreq(n)  # cycles / voxel
        filt = np.exp(-2.0 * (np.pi * k * sigma_vox) ** 2).astype(np.float32)
        sh = [1] * len(shape)
        sh[ax] = filt.size
        spec *= filt.reshape(sh)
    out = sfft.irfftn(spec, s=shape, workers=workers)
    out -= out.mean(dtype=np.float64)
    out /= out.std(dtype=np.float64)
    return out.astype(np.float32)


def layered_medium(shape, dx: float, *, seed: int, sigma: float = 0.02, corr_len: float = 2e-3,
                   y: float = 1.1, nonlinear: bool = False, f0: float = F0) -> Medium:
    z_mm = (np.arange(shape[2]) * dx) * 1e3
    lid = np.digitize(z_mm, LAYER_Z_MM)  # 0 water, 1 skin, 2 fat, 3 muscle
    names = ("water", "skin", "fat", "muscle")
    col = {k: np.array([LAYERS[n][i] for n in names])[lid] for i, k in enumerate(("c", "rho", "a0", "bona"))}
    g1 = gaussian_random_field(shape, dx, corr_len, seed)
    g2 = gaussian_random_field(shape, dx, corr_len, seed + 7919)
    c = (col["c"][None, None, :] * (1.0 + sigma * g1)).astype(np.float32)
    rho = (col["rho"][None, None, :] * (1.0 + sigma * g2)).astype(np.float32)
    alpha_z = np.array([power_law_np_m(a, y, f0) for a in col["a0"]], np.float32)
    alpha = np.broadcast_to(alpha_z[None, None, :], shape).astype(np.float32)
    if nonlinear:
        beta_z = np.array([beta_from_bona(b) for b in col["bona"]], np.float32)
        beta = np.broadcast_to(beta_z[None, None, :], shape).astype(np.float32)
    else:
        beta = np.zeros(shape, np.float32)
    return Medium(alpha=alpha, rho=rho, c=c, beta=beta)


def pinned_spec(grid, medium, source, *, cfl: float, periods_total: int, n_record: int,
                reference_point=None) -> CWRunSpec:
    """CWRunSpec whose schedule runs exactly ``periods_total`` periods (no early stop)."""
    tof = cw_tof_periods(grid.shape, grid.dx, medium.c_min, source.indices, source.f0, reference_point)
    settle = periods_total - n_record - tof
    floor_settle = int(np.ceil(source.ramp_periods)) + 1
    if settle < floor_settle:
        raise ValueError(f"{periods_total} periods cannot hold tof={tof} + ramp + {n_record} record")
    return CWRunSpec(cfl=cfl, cfl_hard_max=max(0.5, cfl), min_settle_periods=settle,
                     max_settle_periods=settle, convergence_tol=0.0, n_record_periods=n_record)


def cfg0_water_bowl() -> Workload:
    """64x64x96 homogeneous water, 0.375 mm, 10-cell PML, bowl ROC 20 / aperture 16 mm,
    100 kPa, CFL 0.3, 40 periods, phasor from the last 4."""
    dx = 0.375e-3
    shape = (64, 64, 96)
    grid = Grid(shape, dx, PMLSpec(thickness=10 * dx))
    alpha = power_law_np_m(0.0022, 2.0, F0)
    full = np.ones(shape, np.float32)
    medium = Medium(alpha=full * alpha, rho=full * 1000.0, c=full * 1500.0, beta=full * 0.0)
    src = bowl_cw_source(grid, F0, 1.0e5, aperture_radius=8e-3, roc=20e-3, apex_vox=(32, 32, 10))
    focus = (32, 32, 10 + grid.voxels_for(20e-3))
    spec = pinned_spec(grid, medium, src, cfl=0.3, periods_total=40, n_record=4,
                       reference_point=focus)
    return Workload("cfg0_water_bowl_64x64x96", "linear", grid, medium, src, spec, (1,), 40,
                    reference_point=focus)


def cfg1_hetero_linear(seed: int = 1, shape=(256, 256, 256), steer_vox=None) -> Workload:
    """256^3 layered + 2 % GRF medium, 0.375 mm, 16-cell PML, bowl ROC 64 / aperture 64 mm,
    200 kPa, CFL 0.3, 60 periods (2 recorded), fundamental phasor."""
    dx = 0.375e-3
    grid = Grid(shape, dx, PMLSpec(thickness=16 * dx))
    medium = layered_medium(shape, dx, seed=seed)
    apex = (shape[0] // 2, shape[1] // 2, grid.pml_vox)
    focus = (apex[0], apex[1], apex[2] + grid.voxels_for(64e-3))
    phases = None
    if steer_vox is not None:
        focus = tuple(int(v) for v in np.asarray(focus) + np.asarray(steer_vox))
        phases = lambda idx: focus_phases(idx, dx, focus, 1500.0, F0)  # noqa: E731
    src = bowl_cw_source(grid, F0, 2.0e5, aperture_radius=32e-3, roc=64e-3, apex_vox=apex,
                         phases=phases)
    spec = pinned_spec(grid, medium, src, cfl=0.3, periods_total=60, n_record=2,
                       reference_point=focus)
    return Workload(f"cfg1_hetero_linear_{'x'.join(map(str, shape))}", "linear", grid, medium,
                    src, spec, (1,), 60, reference_point=focus)


def cfg2_westervelt(seed: int = 2, shape=(256, 256, 384)) -> Workload:
    """256x256x384 layered + GRF Westervelt, 0.25 mm, 20-cell PML, bowl ROC 64 mm with the
    aperture shrunk from 64 to 63.5 mm so the rim stays on the grid, 1 MPa, CFL 0.2,
    80 periods (2 recorded), harmonics 1-3."""
    dx = 0.25e-3
    grid = Grid(shape, dx, PMLSpec(thickness=20 * dx))
    medium = layered_medium(shape, dx, seed=seed, nonlinear=True)
    apex = (shape[0] // 2, shape[1] // 2, grid.pml_vox)
    src = bowl_cw_source(grid, F0, 1.0e6, aperture_radius=31.75e-3, roc=64e-3, apex_vox=apex)
    focus = (apex[0], apex[1], apex[2] + grid.voxels_for(64e-3))
    spec = pinned_spec(grid, medium, src, cfl=0.2, periods_total=80, n_record=2,
                       reference_point=focus)
    return Workload(f"cfg2_westervelt_{'x'.join(map(str, shape))}", "westervelt", grid, medium,
                    src, spec, (1, 2, 3), 80, reference_point=focus)


SERVING_SHAPES = ((128, 128, 128), (192, 192, 192), (256, 256, 256), (128, 128, 256))


def serving_request(rid: int, rng: np.random.Generator, arrival_s: float, shape=None,
                    solver: str | None = None):
    """One request of the mixed-size queue: config-1 medium family at dx 0.375 mm,
    fresh random-field seed, 70 % linear / 30 % Westervelt (harmonics 1, 2), bowl focus
    steered by a random offset (+-10 mm lateral, +-15 mm axial) with delay-and-sum
    phases. Geometry per size: ROC = min(64 mm, 0.6 Lz), aperture radius =
    min(32 mm, 0.8 ROC, 0.4 L_lateral); the steered focus is clipped to stay at least
    PML + 4 voxels inside the grid. Pinned schedule: time of flight to the focus + 10
    settle + 2 record periods, CFL 0.3."""
    from caustica_amd.serving.batcher import SolveRequest

    dx = 0.375e-3
    if shape is None:
        shape = SERVING_SHAPES[int(rng.integers(len(SERVING_SHAPES)))]
    if solver is None:
        solver = "westervelt" if rng.random() < 0.3 else "linear"
    grid = Grid(shape, dx, PMLSpec(thickness=16 * dx))
    medium = layered_medium(shape, dx, seed=int(rng.integers(1 << 30)), nonlinear=solver == "westervelt")
    lz = shape[2] * dx
    lat = min(shape[0], shape[1]) * dx
    roc = min(64e-3, 0.6 * lz)
    a = min(32e-3, 0.8 * roc, 0.4 * lat)
    apex = (shape[0] // 2, shape[1] // 2, grid.pml_vox)
    off = np.array([rng.uniform(-10e-3, 10e-3), rng.uniform(-10e-3, 10e-3), rng.uniform(-15e-3, 15e-3)])
    focus = np.array(apex, np.float64) + np.array([0.0, 0.0, roc / dx]) + off / dx
    lo = grid.pml_vox + 4
    focus = tuple(int(round(np.clip(f, lo, n - 1 - lo))) for f, n in zip(focus, shape))
    src = bowl_cw_source(grid, F0, 2.0e5, aperture_radius=a, roc=roc, apex_vox=apex,
                         phases=lambda idx: focus_phases(idx, dx, focus, 1500.0, F0))
    tof = cw_tof_periods(shape, dx, medium.c_min, src.indices, F0, focus)
    spec = pinned_spec(grid, medium, src, cfl=0.3, periods_total=tof + 12, n_record=2,
                       reference_point=focus)
    harmonics = (1, 2) if solver == "westervelt" else (1,)
    return SolveRequest(rid=rid, arrival_s=arrival_s, solver=solver, grid=grid, medium=medium,
                        source=src, spec=spec, harmonics=harmonics, reference_point=focus,
                        label=f"{solver} {'x'.join(map(str, shape))}")


def serving_workload(seed: int, shape, solver: str) -> Workload:
    """One request of the mixed-size queue as a Workload: the first request that
    ``serving_request`` draws from numpy's generator seeded with ``seed``, at the given
    shape and solver (its medium seed, steering offset and pinned schedule follow from the
    seed). Used for the float64 parity fixtures of the queue's configurations."""
    req = serving_request(0, np.random.default_rng(seed), 0.0, shape=tuple(shape), solver=solver)
    tof = cw_tof_periods(req.grid.shape, req.grid.dx, req.medium.c_min, req.source.indices,
                         req.source.f0, req.reference_point)
    total = tof + req.spec.min_settle_periods + req.spec.n_record_periods
    return Workload(f"serving_{solver}_{'x'.join(map(str, shape))}_seed{seed}", solver, req.grid,
                    req.medium, req.source, req.spec, tuple(req.harmonics), total,
                    reference_point=tuple(int(v) for v in req.reference_point))


def serving_queue(n: int, rate_hz: float, seed: int = 4, shapes=None, progress=None):
    """n requests with Poisson arrivals at rate_hz (seeded exponential inter-arrival times);
    ``progress(done, n)`` (optional) is called after each request is built."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for rid in range(n):
        t += rng.exponential(1.0 / rate_hz)
        shape = None if shapes is None else shapes[int(rng.integers(len(shapes)))]
        out.append(serving_request(rid, rng, t, shape=shape))
        if progress is not None:
            progress(rid + 1, n)
    return out


# Pennes tissue of the thermal benchmark: k 0.5 W/m/K, rho C 3.6e6 J/m^3/K, perfusion
# w_b = 0.5 kg/m^3/s with blood c_b 3617 J/kg/K (perfusion_rate = w_b / rho_b, so that
# the solver's perfusion_rate * rho_b * C_b equals 0.5 * 3617 W/m^3/K), T_art = T0 = 37 C.
PENNES_K = 0.5
PENNES_RHO = 1000.0
PENNES_C = 3600.0
PENNES_PERFUSION_RATE = 0.5 / 1050.0
FOCAL_INTENSITY_W_M2 = 150.0e4  # 150 W/cm^2


def pennes_tissue(shape, dx: float):
    """Homogeneous perfused tissue of the thermal benchmark on ``shape``."""
    from caustica_amd.thermal import ThermalMedium

    one = np.ones(shape, np.float32)
    return ThermalMedium(k=one * PENNES_K, rho=one * PENNES_RHO, specific_heat=one * PENNES_C,
                         perfusion=one * np.float32(PENNES_PERFUSION_RATE), dx=dx)


def focal_intensity_scale(phasor_at_focus: complex, rho: float, c: float,
                          target_w_m2: float = FOCAL_INTENSITY_W_M2) -> float:
    """Factor on Q that brings the focal intensity |P|^2 / (2 rho c) to ``target_w_m2``
    (Q scales as |P|^2, i.e. linearly in intensity)."""
    i_focus = abs(phasor_at_focus) ** 2 / (2.0 * rho * c)
    if not i_focus > 0:
        raise ValueError("zero field at the focus: nothing to scale")
    return float(target_w_m2 / i_focus)
