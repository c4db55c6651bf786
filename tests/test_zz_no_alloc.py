"""cst_thermal_run_noalloc against the thermal session's entry points.

The no-allocation entry point runs the same kernels as cst_thermal_create / setup / run /
read on caller-owned device buffers (allocated by the test through the HIP runtime), with
its ping-pong buffer and finite flag in the library's fixed device array. Two calls in a
row (the second on a smaller grid, so the array holds the first call's contents) must
give bitwise the session's temperature, running maximum and dose; a grid above the
stated limit is refused before anything runs.
"""

import ctypes as C

import numpy as np
import pytest

from caustica_amd import _native as N

pytestmark = pytest.mark.gpu


def _case(shape, seed):
    rng = np.random.default_rng(seed)
    k = (0.2 + 0.3 * rng.random(shape)).astype(np.float32)
    perf = (1800.0 * rng.random(shape)).astype(np.float32)
    coeff = (0.05 / (3.6e6 * (0.9 + 0.2 * rng.random(shape)))).astype(np.float32)
    q = (3e7 * rng.random(shape) ** 4).astype(np.float32)
    theta0 = (5.0 + 0.5 * rng.standard_normal(shape)).astype(np.float32)
    dose0 = (0.01 * rng.random(shape)).astype(np.float32)
    return dict(k=k, perf=perf, coeff=coeff, q=q, theta0=theta0, dose0=dose0)


DX = 5e-4
HALF_DT_MIN = float(np.float32(0.5 * 0.05 / 60.0))
T_REF = 37.0


def _session(shape, c, n, boundary):
    lib = N.lib()
    desc = N.ThermalDesc()
    desc.dx = DX
    desc.k, desc.perf, desc.coeff = N.ptr(c["k"]), N.ptr(c["perf"]), N.ptr(c["coeff"])
    desc.q, desc.theta0, desc.dose0 = N.ptr(c["q"]), N.ptr(c["theta0"]), N.ptr(c["dose0"])
    desc.boundary = boundary
    desc.theta_wall = 0.0
    desc.t_ref = T_REF
    desc.half_dt_min = HALF_DT_MIN
    desc.dose = 1
    h = C.c_void_p()
    N.check(lib.cst_thermal_create(*shape, 0, C.byref(h)))
    try:
        N.check(lib.cst_thermal_setup(h, C.byref(desc)))
        N.check(lib.cst_thermal_run(h, n))
        out = [np.empty(shape, np.float32) for _ in range(3)]
        N.check(lib.cst_thermal_read(h, *[N.ptr(a) for a in out]))
    finally:
        lib.cst_thermal_destroy(h)
    return out


class _Hip:
    """The caller's side: device buffers from the HIP runtime the library itself links
    (hipMalloc / hipMemcpy / hipFree through ctypes)."""

    H2D, D2H = 1, 2

    def __init__(self):
        self.rt = C.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.rt.hipDeviceSynchronize.argtypes = []
        self.bufs = []

    def up(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(p), a.nbytes) == 0
        self.bufs.append(p)
        assert self.rt.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, self.H2D) == 0
        return p

    def down(self, p, like):
        out = np.empty_like(like)
        assert self.rt.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, self.D2H) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.rt.hipFree(p)
        self.bufs.clear()


def _noalloc(shape, c, n, boundary):
    hip = _Hip()
    try:
        theta, tmax = hip.up(c["theta0"]), hip.up(c["theta0"])
        dose = hip.up(c["dose0"].astype(np.float64))
        d = N.ThermalDevDesc()
        d.nx, d.ny, d.nz = shape
        d.dx = DX
        d.theta, d.theta_max, d.dose = theta, tmax, dose
        d.k, d.perf, d.coeff, d.q = (hip.up(c[k]) for k in ("k", "perf", "coeff", "q"))
        d.boundary = boundary
        d.theta_wall = 0.0
        d.t_ref = T_REF
        d.half_dt_min = HALF_DT_MIN
        d.dose_on = 1
        bad = C.c_int(-1)
        N.check(N.lib().cst_thermal_run_noalloc(C.byref(d), n, None, C.byref(bad)))
        assert hip.rt.hipDeviceSynchronize() == 0
        assert bad.value == 0
        f32 = c["theta0"]
        return [hip.down(theta, f32), hip.down(tmax, f32),
                hip.down(dose, f32.astype(np.float64)).astype(np.float32)]
    finally:
        hip.free()


@pytest.mark.parametrize("boundary", [0, 7])
def test_noalloc_equals_session_on_two_calls(boundary):
    for shape, n, seed in (((40, 24, 70), 7, 1), ((17, 13, 30), 4, 2)):
        print(f"[noalloc] {shape} {n} sub-steps, boundary {boundary}", flush=True)
        c = _case(shape, seed)
        ref = _session(shape, c, n, boundary)
        got = _noalloc(shape, c, n, boundary)
        for name, a, b in zip(("theta", "theta_max", "dose"), got, ref):
            assert np.array_equal(a, b), (shape, name, float(np.abs(a - b).max()))
        assert float(ref[0].max()) > float(c["theta0"].max())  # the source heats


def test_noalloc_refuses_grids_above_the_limit():
    d = N.ThermalDevDesc()
    d.nx, d.ny, d.nz = 257, 256, 256
    d.dx = DX
    code = N.lib().cst_thermal_run_noalloc(C.byref(d), 1, None, None)
    assert code == N.CST_ERR_ARG
    assert "no-allocation limit" in N.last_error()
