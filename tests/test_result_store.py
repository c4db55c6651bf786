"""Float16 quantization contract and the caustica-result/1 layout.

The quantizer is pinned by what float16 arithmetic fixes: values scaled into [-1, 1]
round with an error of at most half a float16 ulp at 1 (2^-12 of the peak), an
all-zero field is exact, a contract tighter than float16 can meet falls back to
float32 unchanged, and the recorded error is the measured one. The layout is checked
name by name against upstream ``io/store.py:90-170`` (group, dataset and attribute
names, dtypes) and by a round trip through ``tree_to_result`` within the recorded
errors. The HDF5 container itself needs h5py, absent here: its test is skipped when
h5py cannot be imported.
"""

import importlib.util

import numpy as np
import pytest

from caustica_amd.io.quantize import restore, try_float16
from caustica_amd.io.store import (
    RESULT_FORMAT,
    load_result,
    result_tree,
    save_result,
    tree_to_result,
)
from caustica_amd.solvers.base import SolverResult
from caustica_amd.sources import CWSource

HAS_H5PY = importlib.util.find_spec("h5py") is not None


def test_quantizer_bound_and_exact_cases():
    rng = np.random.default_rng(0)
    a = (rng.standard_normal((17, 9, 5)) * 3e5).astype(np.float32)
    q = try_float16(a)
    assert q.dtype_name == "float16" and q.stored.dtype == np.float16
    peak = float(np.abs(a).max())
    assert q.scale == peak
    back = q.restore()
    assert back.dtype == np.float32
    err = float(np.abs(back - a).max() / peak)
    assert q.norm_err == pytest.approx(err, rel=1e-6)
    assert q.norm_err <= 2.0**-12 * 1.001
    np.testing.assert_array_equal(back, restore(q.stored, q.scale))
    z = try_float16(np.zeros((3, 4), np.float32))
    assert z.dtype_name == "float16" and z.norm_err == 0.0 and not z.restore().any()
    f = try_float16(a, max_norm_err=1e-6)
    assert f.dtype_name == "float32" and f.scale == 1.0 and f.norm_err == 0.0
    np.testing.assert_array_equal(f.stored, a)
    # the peak itself is exact after the round trip (1.0 is a float16 value)
    i = np.unravel_index(np.argmax(np.abs(a)), a.shape)
    assert back[i] == a[i]


def _result(harmonics=(1, 2)):
    rng = np.random.default_rng(1)
    shape = (6, 5, 4)
    ph = {h: (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
          * np.float32(1e5 / h) for h in harmonics}
    res = SolverResult(
        phasor=ph[1], p_max=np.abs(ph[1]).astype(np.float32) * 1.1,
        region=(slice(2, 8), slice(0, 5), slice(3, 7)), dt=4e-8, spp=25, steps_total=900,
        t_end_s=3.6e-5, tof_periods=12, converged_period=30, settle_capped=False,
        convergence_history=[(1, 10.0, 1.0), (2, 12.0, 0.2)], phasors=ph,
        meta={"solver": "westervelt", "backend": "hip"})
    src = CWSource(indices=np.array([[1, 2, 3], [4, 1, 0]]), phases=np.array([0.1, 2.0], np.float32),
                   amplitude=2e5, f0=1e6, label="bowl")
    return res, src


def test_layout_names_and_dtypes():
    res, src = _result()
    t = result_tree(res, src, dx=2.5e-4, grid_shape=(16, 16, 16), pml_vox=4,
                    extra_attrs={"job_name": "j1"})
    a = t["attrs"]
    assert a["format"] == RESULT_FORMAT
    for k in ("caustica_version", "solver", "backend", "f0_hz", "dt_s", "spp", "steps_total",
              "t_end_s", "tof_periods", "converged_period", "settle_capped", "harmonics",
              "dx_m", "grid_shape", "pml_vox", "region_start", "region_stop",
              "phase_convention", "absorption_model", "job_name"):
        assert k in a, k
    assert a["harmonics"] == [1, 2] and a["region_start"] == [2, 0, 3]
    assert a["backend"] == "hip"
    ds = t["output"]["datasets"]
    assert sorted(ds) == ["p_imag_h1", "p_imag_h2", "p_max", "p_real_h1", "p_real_h2"]
    for arr, at in ds.values():
        assert arr.dtype == np.float16 and arr.shape == (6, 5, 4)
        assert set(at) == {"scale", "stored_dtype", "quant_norm_err", "reload"}
    assert set(t["output"]["attrs"]) >= {"stored", "derived", "phase_convention",
                                         "absorption_model", "pmax_window"}
    ind, _ = t["input"]["datasets"]["source_indices"]
    assert ind.dtype == np.int32 and ind.shape == (2, 3)
    assert t["input"]["datasets"]["source_phases_rad"][0].dtype == np.float32
    assert t["input"]["attrs"]["n_points"] == 2 and t["input"]["attrs"]["label"] == "bowl"
    hist, _ = t["convergence"]["datasets"]["history"]
    assert hist.dtype == np.float64 and hist.shape == (2, 3)


def test_round_trip_within_recorded_error():
    res, src = _result()
    t = result_tree(res, src, dx=2.5e-4, grid_shape=(16, 16, 16))
    back = tree_to_result(t)
    ds = t["output"]["datasets"]
    for h in (1, 2):
        for part, name in ((np.real, "p_real"), (np.imag, "p_imag")):
            arr, at = ds[f"{name}_h{h}"]
            err = np.abs(part(back.phasors[h]) - part(res.phasors[h])).max()
            assert err <= at["quant_norm_err"] * at["scale"] * (1 + 1e-6)
    assert back.region == res.region and back.steps_total == 900 and back.spp == 25
    assert back.convergence_history == res.convergence_history
    np.testing.assert_allclose(back.p_max, res.p_max, rtol=0, atol=ds["p_max"][1]["quant_norm_err"]
                               * ds["p_max"][1]["scale"] * (1 + 1e-6))
    raw = tree_to_result(result_tree(res, src, dx=1e-3, grid_shape=(8, 8, 8), quantize=False))
    np.testing.assert_array_equal(raw.phasors[2], res.phasors[2])
    bad = dict(t, attrs=dict(t["attrs"], format="other/1"))
    with pytest.raises(ValueError, match="format"):
        tree_to_result(bad)


@pytest.mark.skipif(HAS_H5PY, reason="h5py is installed: the ImportError path does not apply")
def test_save_without_h5py_names_it(tmp_path):
    res, src = _result()
    with pytest.raises(ImportError, match="h5py"):
        save_result(tmp_path / "r.h5", res, src, dx=1e-3, grid_shape=(8, 8, 8))
    assert not list(tmp_path.iterdir())


@pytest.mark.skipif(not HAS_H5PY, reason="h5py is not installed in this environment")
def test_h5_round_trip(tmp_path):  # pragma: no cover - needs h5py
    res, src = _result()
    p = save_result(tmp_path / "r.h5", res, src, dx=1e-3, grid_shape=(8, 8, 8))
    back = load_result(p)
    assert back.steps_total == res.steps_total
    np.testing.assert_allclose(back.phasors[1], res.phasors[1], rtol=0, atol=1e5 * 2.0**-11)
