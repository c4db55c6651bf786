"""A hip solve through the caustica-result/1 layout: every stored field of a
Westervelt solve with three harmonics is float16 within the 1e-3 contract, and the
result rebuilt from the stored tree matches the solve within each field's recorded
error (the storage floor upstream reports for its results)."""

import numpy as np
import pytest

import caustica_amd.solvers as solvers
from caustica_amd import workloads
from caustica_amd.io.store import result_tree, tree_to_result

pytestmark = pytest.mark.gpu


def test_hip_solve_round_trips_through_the_layout():
    req = workloads.serving_request(0, np.random.default_rng(21), 0.0, shape=(48, 48, 64),
                                    solver="westervelt")
    print("[result store] westervelt 48x48x64 h=(1,2,3)", flush=True)
    res = solvers.get("westervelt")().run(req.grid, req.medium, req.source, req.spec,
                                          backend="hip", harmonics=(1, 2, 3),
                                          reference_point=req.reference_point)
    tree = result_tree(res, req.source, dx=req.grid.dx, grid_shape=req.grid.shape,
                       pml_vox=req.grid.pml_vox)
    assert tree["attrs"]["backend"] == "hip" and tree["attrs"]["harmonics"] == [1, 2, 3]
    ds = tree["output"]["datasets"]
    for name, (arr, at) in ds.items():
        assert at["stored_dtype"] == "float16" and arr.dtype == np.float16, name
        assert 0.0 < at["quant_norm_err"] <= 1e-3, name
    back = tree_to_result(tree)
    for h in (1, 2, 3):
        for part, name in ((np.real, "p_real"), (np.imag, "p_imag")):
            at = ds[f"{name}_h{h}"][1]
            err = np.abs(part(back.phasors[h]) - part(res.phasors[h])).max()
            assert err <= at["quant_norm_err"] * at["scale"] * (1 + 1e-6)
    assert back.steps_total == res.steps_total and back.region == res.region
