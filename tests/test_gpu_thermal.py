"""Parity of the HIP Pennes + CEM43 path against the CPU oracle and the upstream fixtures.

Tolerances: temperature and temperature_max <= 1e-3 K absolute, CEM43 <= 1e-3 relative
(to the map's maximum) against the float64 oracle; the upstream float32 fixture agrees
with the oracle at ~1e-5 K, so the same bars hold against it.
"""

import numpy as np
import pytest

from oracle.pennes import PennesOracle
from tests._golden import load

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["pennes_insulated", "pennes_dirichlet_substep"])
def test_pennes_matches_oracle_and_upstream(name):
    from caustica_amd.thermal import PennesSolver, ThermalMedium, ThermalResult

    print(f"[pennes] {name}", flush=True)
    g = load(name)
    m = g["meta"]
    med = ThermalMedium(k=g["k"], rho=g["rho"], specific_heat=g["specific_heat"],
                        perfusion=g["perfusion"], dx=float(g["dx"]))
    solver = PennesSolver("hip", boundary=m["boundary"], on_unstable=m["on_unstable"])
    dt = float(g["dt"])
    hot = solver.solve(g["temperature0"], g["q"], med, dt=dt, n_steps=m["n_hot"], dose=True)
    cool = solver.solve(hot.temperature, None, med, dt=dt, n_steps=m["n_cool"], dose=True,
                        dose0=hot.dose_cem43)
    chained = ThermalResult.chain([hot, cool])
    assert hot.substeps == int(g["substeps"])
    assert hot.meta["backend"] == "hip"
    o = PennesOracle(g["k"], g["rho"], g["specific_heat"], g["perfusion"], float(g["dx"]),
                     boundary=m["boundary"], on_unstable=m["on_unstable"], precision="fp64")
    oh = o.solve(g["temperature0"], g["q"], dt, m["n_hot"], dose=True)
    oc = o.solve(oh.temperature.astype(np.float32), None, dt, m["n_cool"], dose=True,
                 dose0=oh.dose.astype(np.float32))
    for dev, ref in ((hot.temperature, oh.temperature), (hot.temperature_max, oh.temperature_max),
                     (cool.temperature, oc.temperature), (hot.temperature, g["hot_temperature"]),
                     (chained.temperature_max, g["chained_temperature_max"])):
        assert np.abs(dev - ref).max() < 1e-3
    for dev, ref in ((hot.dose_cem43, oh.dose), (cool.dose_cem43, oc.dose),
                     (cool.dose_cem43, g["cool_dose"])):
        assert np.abs(dev - ref).max() <= 1e-3 * np.abs(ref).max()


def test_pennes_finite_guard_raises():
    from caustica_amd.thermal import PennesSolver, ThermalDivergedError, ThermalMedium

    shape = (8, 8, 8)
    med = ThermalMedium(k=np.full(shape, 0.5), rho=np.full(shape, 1000.0),
                        specific_heat=np.full(shape, 3600.0), perfusion=np.zeros(shape), dx=1e-3)
    q = np.zeros(shape, np.float32)
    q[4, 4, 4] = 1e38  # overflows float32 within one step
    with pytest.raises(ThermalDivergedError):
        PennesSolver("hip", finite_check_every=1).solve(np.full(shape, 37.0), q, med, dt=0.1,
                                                        n_steps=3, dose=True)


@pytest.mark.parametrize("shape", [(37, 13, 70), (64, 6, 130)])
@pytest.mark.parametrize("boundary", ["insulated", "dirichlet"])
def test_pennes_ragged_tiles_match_oracle(shape, boundary):
    """Shapes that cut the kernel's 64 x 4 (z, y) tiles and 32-plane x chunks: a partial
    last tile in z and y and a partial last chunk in x (37), and whole x chunks (64) with
    a y extent below one tile row pair and z crossing three tiles. Heterogeneous
    conductivity, perfusion and heat source, so every halo value enters the result."""
    from caustica_amd.thermal import PennesSolver, ThermalMedium

    print(f"[pennes] ragged {shape} {boundary}", flush=True)
    rng = np.random.default_rng(sum(shape))
    k = (0.2 + 0.3 * rng.random(shape)).astype(np.float32)
    rho = (1000.0 + 80.0 * rng.random(shape)).astype(np.float32)
    cp = (3500.0 + 300.0 * rng.random(shape)).astype(np.float32)
    perf = (5e-4 * rng.random(shape)).astype(np.float32)
    q = (3e7 * rng.random(shape) ** 4).astype(np.float32)
    t0 = (37.0 + 0.5 * rng.standard_normal(shape)).astype(np.float32)
    dx, dt, n = 5e-4, 0.05, 25
    med = ThermalMedium(k=k, rho=rho, specific_heat=cp, perfusion=perf, dx=dx)
    dev = PennesSolver("hip", boundary=boundary).solve(t0, q, med, dt=dt, n_steps=n, dose=True)
    ref = PennesOracle(k, rho, cp, perf, dx, boundary=boundary, precision="fp64").solve(t0, q, dt, n, dose=True)
    assert float(ref.temperature.max()) > 38.0  # the source heats
    assert np.abs(dev.temperature - ref.temperature).max() < 1e-3
    assert np.abs(dev.temperature_max - ref.temperature_max).max() < 1e-3
    assert np.abs(dev.dose_cem43 - ref.dose).max() <= 1e-3 * np.abs(ref.dose).max()


@pytest.mark.parametrize("boundary", ["insulated", "dirichlet"])
def test_pennes_two_step_launch_equals_two_one_step_launches(boundary):
    """The two-step kernel against the one-step kernel launched twice. The host launches
    single steps when the non-finite check comes every step and pairs when it comes every
    second step; both kernels run the same update function and take theta_max and the
    dose in the same order, so the results agree bit for bit. Ragged in every axis
    (x = 70 cuts both kernels' x chunks)."""
    from caustica_amd.thermal import PennesSolver, ThermalMedium

    shape = (70, 13, 37)
    print(f"[pennes] two-step vs one-step {shape} {boundary}", flush=True)
    rng = np.random.default_rng(7)
    k = (0.2 + 0.3 * rng.random(shape)).astype(np.float32)
    rho = (1000.0 + 80.0 * rng.random(shape)).astype(np.float32)
    cp = (3500.0 + 300.0 * rng.random(shape)).astype(np.float32)
    perf = (5e-4 * rng.random(shape)).astype(np.float32)
    q = (3e7 * rng.random(shape) ** 4).astype(np.float32)
    t0 = (42.0 + 0.5 * rng.standard_normal(shape)).astype(np.float32)
    med = ThermalMedium(k=k, rho=rho, specific_heat=cp, perfusion=perf, dx=5e-4)
    out = []
    for every in (1, 2):
        s = PennesSolver("hip", boundary=boundary, finite_check_every=every)
        out.append(s.solve(t0, q, med, dt=0.05, n_steps=8, dose=True))
    one, two = out
    for name in ("temperature", "temperature_max", "dose_cem43"):
        a, b = getattr(one, name), getattr(two, name)
        print(f"[pennes] {name}: max |diff| {float(np.abs(a - b).max()):.3e}", flush=True)
        assert np.array_equal(a, b), name


@pytest.mark.parametrize("boundary", ["insulated", "dirichlet"])
def test_pennes_uniform_map_kernels_equal_map_kernels(boundary):
    """Uniform k, perfusion and rho C (the thermal benchmark's tissue) take the kernels that
    hold the three values in registers; with those kernels switched off the same solve
    reads the maps. Same arithmetic on the same values: the same bits, for a heated phase
    of odd length (two-step launches and one one-step launch) and a cooling phase chained
    onto it. Both also agree with the float64 oracle."""
    from caustica_amd.thermal import PennesSolver, ThermalMedium

    shape = (70, 13, 37)
    print(f"[pennes] uniform-map kernels {shape} {boundary}", flush=True)
    rng = np.random.default_rng(23)
    one = np.ones(shape, np.float32)
    k, rho, cp, perf = one * 0.5, one * 1050.0, one * 3600.0 / 1.05, one * np.float32(5e-4)
    q = (3e7 * rng.random(shape) ** 4).astype(np.float32)
    t0 = (37.0 + 0.5 * rng.standard_normal(shape)).astype(np.float32)
    med = ThermalMedium(k=k, rho=rho, specific_heat=cp, perfusion=perf, dx=5e-4)
    runs = {}
    for uni in (True, False):
        s = PennesSolver("hip", boundary=boundary, finite_check_every=25, uniform_maps=uni)
        hot = s.solve(t0, q, med, dt=0.05, n_steps=9, dose=True)
        cool = s.solve(hot.temperature, None, med, dt=0.05, n_steps=7, dose=True, dose0=hot.dose_cem43)
        assert hot.meta["uniform_map_kernels"] is uni
        runs[uni] = (hot, cool)
    for a, b in zip(runs[True], runs[False]):
        for name in ("temperature", "temperature_max", "dose_cem43"):
            assert np.array_equal(getattr(a, name), getattr(b, name)), name
    o = PennesOracle(k, rho, cp, perf, 5e-4, boundary=boundary, precision="fp64")
    ref = o.solve(t0, q, 0.05, 9, dose=True)
    hot = runs[True][0]
    assert np.abs(hot.temperature - ref.temperature).max() < 1e-3
    assert np.abs(hot.temperature_max - ref.temperature_max).max() < 1e-3
    assert np.abs(hot.dose_cem43 - ref.dose).max() <= 1e-3 * np.abs(ref.dose).max()


@pytest.mark.parametrize("boundary", ["insulated", "dirichlet"])
def test_pennes_chunk_length_does_not_change_results(boundary):
    """The two-step kernel's x-chunk length (planes per block, chosen per grid unless set)
    only moves where blocks recompute their border planes: every length gives the same
    bits. Ragged shape, heterogeneous coefficients, 10 steps = 5 two-step launches."""
    from caustica_amd.thermal import PennesSolver, ThermalMedium

    shape = (70, 13, 37)
    rng = np.random.default_rng(17)
    k = (0.2 + 0.3 * rng.random(shape)).astype(np.float32)
    rho = (1000.0 + 80.0 * rng.random(shape)).astype(np.float32)
    cp = (3500.0 + 300.0 * rng.random(shape)).astype(np.float32)
    perf = (5e-4 * rng.random(shape)).astype(np.float32)
    q = (3e7 * rng.random(shape) ** 4).astype(np.float32)
    t0 = (42.0 + 0.5 * rng.standard_normal(shape)).astype(np.float32)
    med = ThermalMedium(k=k, rho=rho, specific_heat=cp, perfusion=perf, dx=5e-4)
    out = {}
    for xch in (0, 8, 16, 32, 64):
        s = PennesSolver("hip", boundary=boundary, finite_check_every=10, x_chunk=xch)
        out[xch] = s.solve(t0, q, med, dt=0.05, n_steps=10, dose=True)
    for xch, r in out.items():
        for name in ("temperature", "temperature_max", "dose_cem43"):
            assert np.array_equal(getattr(r, name), getattr(out[0], name)), (xch, name)
