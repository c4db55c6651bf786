"""Product host functions the benchmark workloads use, against upstream-written fixtures.

* ``caustica_amd.sources.bowl_cw_source`` must give exactly the upstream voxel set of
  each benchmark bowl (tests/golden/bowl_cfg{0,1,2}.npz, written by the upstream
  ``sources.bowl_cw_source``, src/caustica/sources.py:143-175): same voxels, same
  order, zero phase.
* ``caustica_amd.sensors.HeatingSource.from_phasors`` must give the upstream Q map bit
  for bit (tests/golden/heat_power_law.npz: harmonics 1-2 under a power-law absorption
  model, upstream src/caustica/sensors.py:193-295).
"""

import numpy as np
import pytest

from tests._golden import load


@pytest.mark.parametrize("name", ["bowl_cfg0", "bowl_cfg1", "bowl_cfg2"])
def test_product_bowl_matches_upstream_voxels(name):
    from caustica_amd import Grid, PMLSpec
    from caustica_amd.sources import bowl_cw_source

    g = load(name)
    dx = float(g["dx"])
    grid = Grid(tuple(int(s) for s in g["shape"]), dx, PMLSpec(thickness=int(g["pml_vox"]) * dx))
    src = bowl_cw_source(grid, 1.0e6, 1.0e5, aperture_radius=float(g["aperture_radius"]),
                         roc=float(g["roc"]), apex_vox=tuple(int(v) for v in g["apex"]))
    np.testing.assert_array_equal(src.indices, g["indices"].astype(np.int64))
    assert src.phases.dtype == np.float32 and not src.phases.any()


def test_benchmark_workloads_use_the_upstream_bowls():
    """The workload generator's cfg0 bowl is the fixture's (cfg1/cfg2 build 256^3 media,
    which the product-bowl test above covers without the medium)."""
    from caustica_amd import workloads

    wl = workloads.cfg0_water_bowl()
    np.testing.assert_array_equal(wl.source.indices, load("bowl_cfg0")["indices"].astype(np.int64))


def test_product_heating_source_matches_upstream():
    from caustica_amd.sensors import HeatingSource

    g = load("heat_power_law")
    hs = HeatingSource.from_phasors({1: g["p1"], 2: g["p2"]}, alpha=g["alpha"], rho=g["rho"],
                                    c=g["c"], dx=1e-3, alpha_power_law_y=float(g["y"]))
    assert hs.q.dtype == np.float32
    np.testing.assert_array_equal(hs.q, g["q"])
    assert hs.harmonics == (1, 2)


def test_product_heating_source_fundamental_only_is_two_alpha_intensity():
    """Q = alpha |P|^2 / (rho c) = 2 alpha I for the fundamental (upstream sensors.py:193-295;
    tests/test_thermal.py:720-727 there)."""
    from caustica_amd.sensors import HeatingSource

    g = load("heat_power_law")
    hs = HeatingSource.from_phasors({1: g["p1"]}, alpha=g["alpha"], rho=g["rho"], c=g["c"], dx=1e-3)
    intensity = np.abs(g["p1"].astype(np.complex128)) ** 2 / (2.0 * g["rho"] * g["c"].astype(np.float64))
    np.testing.assert_allclose(hs.q, 2.0 * g["alpha"] * intensity, rtol=1e-6)
