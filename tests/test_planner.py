"""Planner: upstream inventory (pinned to the hand count of upstream
tests/test_planner.py:40-56 and the padding case :58-64), settle-policy step counts
(:74-99), device table and advice (:160-180), and the hip inventory's structure. The
hip inventory's total is pinned on the device against the library's own allocation
(tests/test_gpu_planner.py)."""

import json
import math

import numpy as np
import pytest

from caustica_amd import Grid, Medium, PMLSpec, plane_cw_source
from caustica_amd import planner
from caustica_amd.materials import water
from caustica_amd.solvers import CWRunSpec


def tiny_setup(shape=(32, 32), dx=1e-3, f0=0.5e6):
    grid = Grid(shape, dx, PMLSpec(thickness=4e-3))
    med = Medium.homogeneous(shape, water())
    src = plane_cw_source(grid, f0=f0, amplitude=1e5)
    return grid, med, src


def test_reference_inventory_hand_count():
    mm = planner.kspace_memory((60, 50, 40), nonlinear=True, n_harmonics=2, rec_elems=1000)
    assert mm.padded_shape == (60, 50, 40)
    p, r = 60 * 50 * 40, 60 * 50 * 21
    b = mm.breakdown
    assert b["state (p + u)"] == 16 * p
    assert b["property maps"] == 16 * p
    assert b["sponge"] == 4 * p
    assert b["spectral factors (i*k*kappa)"] == 24 * r
    assert b["record buffers"] == 1000 * 20
    assert b["step temporaries"] == 24 * r + 16 * p
    assert b["fft workspace"] == 16 * r
    assert mm.total_bytes == math.ceil(sum(b.values()) * 1.15)


def test_reference_inventory_pads_to_smooth_length():
    small = planner.kspace_memory((60, 50, 40), nonlinear=False, n_harmonics=1)
    padded = planner.kspace_memory((61, 50, 40), nonlinear=False, n_harmonics=1)
    assert padded.padded_shape == (64, 50, 40)
    assert padded.total_bytes > small.total_bytes


def test_hip_inventory_structure():
    """Every piece is a 256-byte multiple; the dominant terms are 4 B/voxel of pressure,
    56 B per padded spectrum bin (seven spectra) and 4 B/voxel per map; linear drops one
    map; record buffers scale with harmonics as 16 B + 4 B per recorded voxel."""
    kw = dict(n_harmonics=1, rec_elems=256**3, n_src=5000, max_periods=64)
    lin = planner.hip_inventory((256, 256, 256), nonlinear=False, **kw)
    nl = planner.hip_inventory((256, 256, 256), nonlinear=True, **kw)
    assert all(v % 256 == 0 for v in lin.breakdown.values())
    P = 256**3
    assert lin.breakdown["pressure (p)"] == 4 * P
    assert lin.breakdown["half spectra (sp, sg0-2, velocity state su0-2)"] == 56 * (256 * (256 * 128 + 16) + 256 * 256)
    assert nl.total_bytes - lin.total_bytes == 4 * P
    h3 = planner.hip_inventory((256, 256, 256), nonlinear=False,
                               **{**kw, "n_harmonics": 3})
    assert h3.total_bytes - lin.total_bytes == 2 * 16 * 256**3
    # the hip member is smaller than the upstream inventory of the same solve
    ref = planner.kspace_memory((256, 256, 256), False, 1)
    assert lin.total_bytes < ref.total_bytes


def test_steps_follow_settle_policy_and_t_end_floor():
    grid, med, src = tiny_setup()
    spec = CWRunSpec(min_settle_periods=2, max_settle_periods=12, n_record_periods=1)
    est = planner.estimate(grid, med, src, spec, solver="linear")
    from caustica_amd.solvers.kspace.operators import cw_discretization, cw_tof_periods

    spp, dt = cw_discretization(grid.dx, med.c_max, spec, src.f0)
    tof = cw_tof_periods(grid.shape, grid.dx, med.c_min, src.indices, src.f0)
    assert (est.spp, est.dt) == (spp, dt)
    floor_ = max(2, math.ceil(src.ramp_periods) + 1)
    assert est.steps_expected == (tof + floor_ + 1) * spp
    assert est.steps_worst == (max(tof + 12, tof + floor_) + 1) * spp

    spec = CWRunSpec(min_settle_periods=1, max_settle_periods=4, n_record_periods=1,
                     t_end_min_us=200.0)
    est = planner.estimate(grid, med, src, spec)
    need = math.ceil(200e-6 * src.f0) - 1
    assert tof + 4 < need
    assert est.steps_expected == est.steps_worst == (need + 1) * est.spp


def test_mi355x_row_and_aliases():
    devs, aliases = planner.load_gpu_db()
    d = devs["MI355X"]
    assert (d.vram_gib, d.mem_bw_gbs, d.fp32_tflops) == (288, 8000, 157.3)
    assert d.runs_hip and not devs["A100-40GB"].runs_hip
    grid, med, src = tiny_setup()
    est = planner.estimate(grid, med, src, gpu="gfx950")
    assert est.gpu == "MI355X" and est.source == "db" and est.fits
    assert "half spectra (sp, sg0-2, velocity state su0-2)" in est.vram_breakdown
    assert est.vram_bytes_reference_model > 0
    other = planner.estimate(grid, med, src, gpu="A100")
    assert other.gpu == "A100-40GB" and "fft workspace" in other.vram_breakdown
    with pytest.raises(ValueError, match="unknown gpu"):
        planner.estimate(grid, med, src, gpu="RTX9999")
    with pytest.raises(ValueError, match="k-space engine"):
        planner.estimate(grid, med, src, solver="kwave")


def test_bandwidth_time_model():
    """db step time = step bytes / (datasheet bandwidth x bw_eff); 256^3 linear moves
    248 H + 24 P bytes (2.50 GB) per step."""
    b = planner.step_bytes((256, 256, 256), False)
    assert b == 248 * 256 * 256 * 129 + 24 * 256**3
    grid = Grid((256, 256, 256), 2e-4, PMLSpec(thickness=16 * 2e-4))
    med = Medium.homogeneous(grid.shape, water())
    src = plane_cw_source(grid, f0=1e6, amplitude=1e5, axis=2)
    est = planner.estimate(grid, med, src, solver="linear")
    assert est.t_step_s == pytest.approx(b / (8000e9 * 0.59))
    assert est.padded_shape == (256, 256, 256)


def test_calibration_entry_overrides_db(tmp_path):
    grid, med, src = tiny_setup()
    path = tmp_path / "cal.json"
    key = "1x32x32:lin"
    path.write_text(json.dumps({"version": 1, "devices": {"MI355X": {
        key: {"bytes_per_s": 1e12, "t_step_s": 0.0, "warmup_s": 0.25, "steps": 10}}}}))
    est = planner.estimate(grid, med, src, solver="linear", calibration_path=path)
    assert est.source == "calibrated"
    assert est.t_step_s == pytest.approx(planner.step_bytes((1, 32, 32), False) / 1e12)
    assert est.warmup_s == 0.25


def test_oom_advice_and_compare():
    shape = (216, 216, 216)
    grid = Grid(shape, 0.5e-3, PMLSpec(thickness=3e-3))
    med = Medium.homogeneous(shape, water(beta=3.5))
    src = plane_cw_source(grid, f0=1e6, amplitude=1e5)
    tiny = planner.GPUSpec(key="TINY", vram_gib=1.5, mem_bw_gbs=100, fp32_tflops=5.0, bw_eff=0.5)
    est = planner.estimate(grid, med, src, solver="westervelt", gpu=tiny, harmonics=(1, 2))
    assert not est.fits
    assert any("increase dx" in a for a in est.advice)
    assert any("harmonic" in a for a in est.advice)
    assert any("'linear' solver" in a for a in est.advice)
    assert any("MI355X" in a for a in est.advice)
    assert "DOES NOT FIT" in est.summary()
    comp = planner.compare(grid, med, src, gpus=("T4", "MI355X", "H100"))
    assert comp.estimates[0].gpu == "MI355X"
    assert len(str(comp).splitlines()) == 5
    assert np.isfinite([e.t_expected_s for e in comp.estimates]).all()
