"""Transducer arrays against upstream fixtures and closed forms.

Fixtures (tests/golden/array_*.npz, written by tests/golden/make_golden.py from the
upstream ``arrays/transducer.py``): spiral element positions/normals/radius of the
production 128-element array, a compact 32-element array and a 7-element one;
delay-and-sum phases towards three targets; voxelized sources (index set, owning
element, per-voxel phase) on a 96^3 grid (uniform and steered phases, two apex
positions) and the production array on a 288 x 288 x 64 grid. Closed forms: the
equal-area element radius (upstream tests/test_arrays.py:28-43), elements on the
focal sphere, unit normals towards the focus, the one-radius inset, DAS phases of the
geometric focus all zero.
"""


import numpy as np
import pytest

from caustica_amd import Grid, PMLSpec
from caustica_amd.arrays import TransducerArray, archimedean_spiral
from tests._golden import load

F0, C0 = 1.0e6, 1500.0


@pytest.fixture(scope="module")
def gold():
    return load("array_spiral")


@pytest.fixture(scope="module")
def vox():
    return load("array_voxelize")


def _small():
    return archimedean_spiral(n_elements=32, d_outer=0.030, d_inner=0.010, roc=0.030,
                              active_fraction=0.60)


@pytest.mark.parametrize("name,kw", [
    ("prod", {}),
    ("small", dict(n_elements=32, d_outer=0.030, d_inner=0.010, roc=0.030, active_fraction=0.60)),
    ("odd", dict(n_elements=7, d_outer=0.024, d_inner=0.006, roc=0.020, active_fraction=0.3,
                 n_arc_samples=257)),
])
def test_spiral_matches_fixture(gold, name, kw):
    arr = archimedean_spiral(**kw)
    assert arr.elem_radius == float(gold[f"{name}_elem_radius"])
    np.testing.assert_array_equal(arr.positions, gold[f"{name}_positions"])
    np.testing.assert_array_equal(arr.normals, gold[f"{name}_normals"])


def test_production_geometry_closed_forms():
    arr = archimedean_spiral()
    assert arr.n_elements == 128
    # equal-area radius: 2 pi R (h_out - h_in) * 0.6 / 128 = pi a^2
    h = lambda r: 0.1 - np.sqrt(0.1**2 - r**2)
    a = np.sqrt(2 * 0.1 * (h(0.05) - h(0.022)) * 0.6 / 128)
    assert arr.elem_radius == pytest.approx(a, rel=1e-14)
    assert arr.elem_radius == pytest.approx(3.205e-3, rel=1e-2)
    r = np.hypot(arr.positions[:, 0], arr.positions[:, 1])
    assert r.min() >= 0.022 + arr.elem_radius - 1e-9
    assert r.max() <= 0.050 - arr.elem_radius + 1e-9
    np.testing.assert_allclose(np.linalg.norm(arr.positions - arr.focus, axis=1), 0.1, rtol=1e-9)
    to_focus = arr.focus - arr.positions
    np.testing.assert_allclose(arr.normals, to_focus / np.linalg.norm(to_focus, axis=1)[:, None],
                               rtol=0, atol=1e-15)
    # equal arc length between consecutive elements along the spiral
    th = np.unwrap(np.arctan2(arr.positions[:, 1], arr.positions[:, 0]))
    b = (r[-1] - r[0]) / (th[-1] - th[0])
    fine = np.linspace(th[0], th[-1], 200001)
    sp = np.sqrt((r[0] + b * (fine - th[0])) ** 2 + b**2)
    s = np.concatenate(([0.0], np.cumsum(np.diff(fine) * (sp[1:] + sp[:-1]) / 2)))
    s_el = np.interp(th, fine, s)
    np.testing.assert_allclose(np.diff(s_el), s_el[-1] / 127, rtol=1e-4)


def test_spiral_validation():
    with pytest.raises(ValueError, match="d_inner"):
        archimedean_spiral(d_inner=0.2, d_outer=0.1)
    with pytest.raises(ValueError, match="active_fraction"):
        archimedean_spiral(active_fraction=0.0)
    with pytest.raises(ValueError, match="does not fit"):
        archimedean_spiral(n_elements=7, d_outer=0.012, d_inner=0.004, roc=0.015,
                           active_fraction=0.9)
    with pytest.raises(ValueError, match="matching"):
        TransducerArray(np.zeros((3, 2)), np.zeros((3, 2)), 1e-3, 0.1)


def test_das_phases_match_fixture(gold):
    arr = archimedean_spiral()
    for t, want in zip(gold["targets"], gold["prod_das"]):
        got = arr.das_phases(t, f0=F0, c0=C0)
        assert got.dtype == np.float32
        np.testing.assert_array_equal(got, want)
    small = _small()
    np.testing.assert_array_equal(small.das_phases(gold["small_target"], F0, C0), gold["small_das"])


def test_das_phases_closed_form():
    arr = _small()
    # every element is focal_length from the geometric focus: all phases equal modulo
    # 2 pi (rounding may put some just below 2 pi instead of at 0)
    ph = arr.das_phases(arr.focus, F0, C0)
    assert np.abs(np.angle(np.exp(1j * ph.astype(np.float64)))).max() < 1e-4
    # a target off the focus: phases are k * (d_max - d_j) mod 2 pi, min 0, below 2 pi
    t = np.array([2e-3, 1e-3, 0.028])
    ph = arr.das_phases(t, F0, C0).astype(np.float64)
    d = np.linalg.norm(arr.positions - t, axis=1)
    k = 2 * np.pi * F0 / C0
    assert ph.min() == 0.0 and ph.max() < 2 * np.pi
    # phase differences reproduce path differences modulo 2 pi
    diff = np.angle(np.exp(1j * (ph + k * d)) / np.exp(1j * (ph[0] + k * d[0])))
    assert np.abs(diff).max() < 1e-5


@pytest.mark.parametrize("case", ["small", "steer", "prod"])
def test_voxelize_matches_fixture(gold, vox, case):
    meta = vox["meta"]
    if case == "prod":
        arr = archimedean_spiral()
        dx = meta["prod"]["dx"]
        grid = Grid((288, 288, 64), dx, PMLSpec(thickness=6 * dx))
        asrc = arr.voxelize(grid, tuple(meta["prod"]["apex"]), f0=F0, amplitude=1e5,
                            phases=gold["prod_das"][1])
    else:
        arr = _small()
        grid = Grid((96, 96, 96), 0.5e-3, PMLSpec(thickness=5e-3))
        if case == "small":
            asrc = arr.voxelize(grid, (48, 48, 12), f0=F0, amplitude=1e5)
        else:
            asrc = arr.voxelize(grid, (47, 50, 10), f0=F0, amplitude=2e5, phases=gold["small_das"])
    np.testing.assert_array_equal(asrc.source.indices, vox[f"{case}_idx"].astype(np.int64))
    np.testing.assert_array_equal(asrc.element_of_voxel, vox[f"{case}_elem"])
    np.testing.assert_array_equal(asrc.source.phases, vox[f"{case}_phase"])
    assert asrc.n_elements_represented == arr.n_elements
    if case in meta and "label" in meta[case]:
        assert asrc.source.label == meta[case]["label"]


def test_voxelize_invariants_and_errors():
    arr = _small()
    grid = Grid((96, 96, 96), 0.5e-3, PMLSpec(thickness=5e-3))
    asrc = arr.voxelize(grid, (48, 48, 12), f0=F0, amplitude=1e5)
    idx = asrc.source.indices
    # lexicographic, unique
    key = (idx[:, 0] * 96 + idx[:, 1]) * 96 + idx[:, 2]
    assert (np.diff(key) > 0).all()
    r_mm = np.hypot(idx[:, 0] - 48, idx[:, 1] - 48) * 0.5
    assert r_mm.max() <= 15.0 + 1.0
    h_vox = (0.030 - np.sqrt(0.030**2 - 0.015**2)) / 0.5e-3
    assert 0.5 * h_vox <= idx[:, 2].max() - idx[:, 2].min() <= h_vox + 2
    phases = np.linspace(0, 2 * np.pi, 32, endpoint=False).astype(np.float32)
    asrc = arr.voxelize(grid, (48, 48, 12), f0=F0, amplitude=1e5, phases=phases)
    np.testing.assert_array_equal(asrc.source.phases, phases[asrc.element_of_voxel])
    with pytest.raises(ValueError, match="phases must be"):
        arr.voxelize(grid, (48, 48, 12), f0=F0, amplitude=1e5, phases=np.zeros(5))
    with pytest.raises(ValueError, match="3-D"):
        arr.voxelize(Grid((96, 96), 0.5e-3, PMLSpec(thickness=5e-3)), (48, 48, 12), F0, 1e5)
    with pytest.raises(ValueError, match="outside grid"):
        arr.voxelize(grid, (10, 48, 12), f0=F0, amplitude=1e5)
    with pytest.raises(ValueError, match="lost all voxels"):
        arr.voxelize(Grid((96, 96, 96), 6e-3, PMLSpec(thickness=12e-3)), (48, 48, 12), F0, 1e5)
