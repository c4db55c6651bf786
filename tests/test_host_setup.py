"""CPU checks of the host side: library exports, host setup arrays, schedule logic.

No GPU is needed: the C ABI library is loaded (not called for compute), and the
float32 host setup (coefficient maps, k tables, sponge profiles, source scale,
schedule) is compared with the oracle's fp32 restatement, which is itself pinned
bit-for-bit to the upstream fixtures.
"""

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

from oracle.kspace import KSpaceOracle, RunSpec
from tests._cases import product_case

ROOT = Path(__file__).resolve().parents[1]
CASES = ["cw_1d_linear", "cw_1d_westervelt", "cw_2d_hetero", "cw_3d_bowl", "cw_3d_westervelt",
         "cw_3d_padded"]


def test_library_exports_every_declared_symbol():
    from caustica_amd import _native

    header = (ROOT / "include" / "caustica_hip.h").read_text()
    declared = set(re.findall(r"\bint\s+(cst_\w+)\s*\(", header))
    assert declared, "no declarations found"
    lib = ctypes.CDLL(str(_native.LIB_PATH))
    missing = [name for name in sorted(declared) if not hasattr(lib, name)]
    assert not missing, f"declared but not exported: {missing}"
    assert declared == set(_native.EXPORTS)
    assert _native.lib().cst_version() >= 100


def test_supported_lengths_cover_the_benchmark_shapes():
    from caustica_amd import _native

    lib = _native.lib()
    for n in (64, 96, 128, 192, 256, 384):
        assert lib.cst_supported_length(n, 0) == 1
        assert lib.cst_supported_length(n, 2) == 1
    assert lib.cst_supported_length(7, 0) == 0
    assert lib.cst_supported_length(75, 2) == 0  # odd real axis


@pytest.mark.parametrize("name", CASES)
def test_host_setup_matches_oracle_fp32(name):
    from caustica_amd.solvers.kspace.operators import plan_member

    g, grid, medium, source, spec, kw, solver, args = product_case(name)
    plan = plan_member(grid, medium, source, spec, nonlinear=solver == "westervelt", **kw)
    a = dict(args)
    a["spec"] = RunSpec(**a["spec"])
    orc = KSpaceOracle(precision="fp32", **a)
    m = g["meta"]
    assert (plan.spp, plan.dt) == (m["spp"], m["dt"])
    assert plan.tof_periods == m["tof_periods"]
    assert list(plan.padded) == m["padded_shape"]
    np.testing.assert_array_equal(plan.dt_over_rho, orc.dt_over_rho)
    np.testing.assert_array_equal(plan.rhoc2_dt, orc.rhoc2_dt)
    np.testing.assert_array_equal(plan.absorb, orc.absorb)
    if orc.beta2_dt is None:
        assert plan.beta2_dt is None
    else:
        np.testing.assert_array_equal(plan.beta2_dt, orc.beta2_dt)
    np.testing.assert_array_equal(plan.src_scale, orc.src_scale)
    # the dense sponge of upstream == ((sx sy) sz) of the device profiles
    nx, ny, nz = plan.padded3
    sz, sx, sy = plan.sponge[:nz], plan.sponge[nz:nz + nx], plan.sponge[nz + nx:]
    dense = (sx[:, None, None] * sy[None, :, None]) * sz[None, None, :]
    np.testing.assert_array_equal(dense.reshape(orc.sponge.shape), orc.sponge)
    # derivative factors: i k_d kappa with kappa from the full |k|
    kt = plan.ktab
    nzh = nz // 2 + 1
    kx, kxd = kt[:nx], kt[nx:2 * nx]
    ky, kyd = kt[2 * nx:2 * nx + ny], kt[2 * nx + ny:2 * nx + 2 * ny]
    kz, kzd = kt[2 * nx + 2 * ny:2 * nx + 2 * ny + nzh], kt[2 * nx + 2 * ny + nzh:]
    k2 = kx[:, None, None] ** 2 + ky[None, :, None] ** 2 + kz[None, None, :] ** 2
    kap = np.sinc(plan.c_ref * np.sqrt(k2) * (plan.dt / (2 * np.pi))).astype(np.float32)
    dz = (1j * kzd[None, None, :] * kap).astype(np.complex64)
    np.testing.assert_allclose(dz.reshape(orc.deriv[-1].shape), orc.deriv[-1], rtol=2e-6, atol=1e-3)
