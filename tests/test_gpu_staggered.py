"""The opt-in staggered scheme on hip against its float64 oracle and closed forms.

The upstream solver has no staggered / power-law / split-PML scheme (power-law absorption
is its unstarted milestone M16, upstream MILESTONES.md:1201), so its parity anchor is the
float64 oracle of the same equations (oracle/staggered.py, itself pinned by closed forms
in tests/test_oracle_staggered.py) and the closed forms directly:
  one step from a random state (p, u_a, rho_a)    max |err| <= 2e-5 of max |field|
  full solves: fundamental phasor, p_max rel-L2   <= 1e-4; harmonics 2-3 <= 1e-3
  plane waves through hip: lossless phase speed c0 within 1e-4, PML ripple < 1e-4,
  power-law attenuation within 2 % of the exact dispersion relation, octave slope within
  0.03 of y, Kramers-Kronig phase speed within 6e-4; Fubini harmonics within 3 % / 5 %
and bit-identity where the scheme demands it (beta = 0 Westervelt == linear, a solve
alone == the same solve batched, a raw-state round trip resumes bit for bit).
"""

import numpy as np
import pytest

from oracle.analytic import fubini, powerlaw_wavenumber, shock_distance
from oracle.kspace import RunSpec, rel_l2
from oracle.staggered import StaggeredOracle
from tests._cases import product_case

pytestmark = pytest.mark.gpu

C0 = 1500.0


def _medium_case(shape, absorbing, nonlinear, seed=3):
    """Heterogeneous medium and a few source voxels on a grid of `shape`."""
    from caustica_amd import CWSource, Grid, Medium, PMLSpec

    rng = np.random.default_rng(seed)
    dx = 0.2e-3
    c = (1480.0 + 120.0 * rng.random(shape)).astype(np.float32)
    rho = (980.0 + 80.0 * rng.random(shape)).astype(np.float32)
    alpha = ((4.0 + 6.0 * rng.random(shape)) if absorbing else np.zeros(shape)).astype(np.float32)
    beta = (3.2 + 0.6 * rng.random(shape)).astype(np.float32)
    beta.reshape(-1)[: beta.size // 5] = 0.0  # m = 0 there
    nd = len(shape)
    ctr = np.array([n // 2 for n in shape])
    idx = np.array([ctr, ctr + 1, ctr - 1] if nd > 1 else [ctr, ctr + 2], np.int64)
    if nd == 1:
        idx = idx.reshape(-1, 1)
    grid = Grid(shape, dx, PMLSpec(thickness=3 * dx))
    medium = Medium(alpha=alpha, rho=rho, c=c, beta=beta)
    source = CWSource(indices=idx, phases=np.linspace(0.0, 1.0, idx.shape[0]).astype(np.float32),
                      amplitude=2e5, f0=1.5e6, ramp_periods=2.0)
    return grid, medium, source


def _oracle_for(grid, medium, source, spec, nonlinear, y, harmonics=(1,), record_region=None,
                reference_point=None):
    return StaggeredOracle(
        shape=grid.shape, dx=grid.dx, pml_vox=grid.pml_vox, alpha=medium.alpha, rho=medium.rho,
        c=medium.c, beta=medium.beta, src_idx=source.indices, src_phase=source.phases,
        amplitude=source.amplitude, f0=source.f0, ramp_periods=source.ramp_periods,
        spec=RunSpec(**spec.model_dump()), harmonics=harmonics, record_region=record_region,
        reference_point=reference_point, nonlinear=nonlinear, alpha_power=y)


def _embed(a, padded3):
    return np.ascontiguousarray(np.reshape(a, padded3), np.float32)


STEP_CASES = [
    ((12, 10, 16), True, True, 1.5),
    ((15, 16, 20), True, False, 1.1),
    ((16, 12, 24), False, True, 1.5),
    ((15, 12, 16), False, False, 1.5),
    ((20, 24), True, True, 1.3),
    ((48,), True, False, 1.8),
]


@pytest.mark.parametrize("shape,absorbing,nonlinear,y", STEP_CASES)
def test_one_step_from_random_state(shape, absorbing, nonlinear, y):
    from caustica_amd.solvers import CWRunSpec
    from caustica_amd.solvers.kspace.native import KSpaceSession
    from caustica_amd.solvers.kspace.staggered import plan_staggered

    print(f"[staggered one-step] {shape} absorbing={absorbing} nonlinear={nonlinear} y={y}", flush=True)
    grid, medium, source = _medium_case(shape, absorbing, nonlinear)
    spec = CWRunSpec(cfl=0.3, min_settle_periods=4, max_settle_periods=8)
    plan = plan_staggered(grid, medium, source, spec, nonlinear=nonlinear, alpha_power=y)
    assert plan.stg["absorbing"] == absorbing
    orc = _oracle_for(grid, medium, source, spec, nonlinear, y)
    assert orc.spp == plan.spp and orc.dt == plan.dt
    nd = grid.ndim
    rng = np.random.default_rng(11)
    shp = plan.padded
    p0 = rng.standard_normal(shp).astype(np.float32) * 1e4
    u0 = [rng.standard_normal(shp).astype(np.float32) * 1e-2 for _ in range(nd)]
    r0 = [rng.standard_normal(shp).astype(np.float32) * 1e-3 for _ in range(nd)]
    n0 = 3 * plan.spp + 2
    orc.p = p0.astype(np.float64)
    for i in range(nd):
        orc.u[i] = u0[i].astype(np.float64)
        orc.rho_a[i] = r0[i].astype(np.float64)
    orc.step(n0)
    dev_u = [np.zeros(plan.padded3, np.float32) for _ in range(3)]
    dev_r = [np.zeros(plan.padded3, np.float32) for _ in range(3)]
    for i in range(nd):
        dev_u[3 - nd + i] = _embed(u0[i], plan.padded3)
        dev_r[3 - nd + i] = _embed(r0[i], plan.padded3)
    with KSpaceSession(plan.padded3, 1, scheme="staggered") as s:
        slot = s.add(plan)
        s.set_state(slot, n0, _embed(p0, plan.padded3), *dev_u)
        s.set_density(slot, *dev_r)
        s.advance([slot], 1)
        (p1, *u1), n1 = s.get_state(slot)
        r1 = s.get_density(slot)
    assert n1 == n0 + 1

    def close(got, ref, what):
        ref = np.reshape(ref, plan.padded3)
        err = np.abs(got - ref).max() / np.abs(ref).max()
        assert err < 2e-5, f"{what}: {err:.3e}"

    close(p1, orc.p, "p")
    for i in range(nd):
        close(u1[3 - nd + i], orc.u[i], f"u{i}")
        close(r1[3 - nd + i], orc.rho_a[i], f"rho{i}")
    for a in range(3 - nd):  # embedding axes carry nothing
        assert not u1[a].any() and not r1[a].any()


FULL_CASES = [("cw_3d_bowl", "linear", 1.5), ("cw_3d_westervelt", "westervelt", 1.2),
              ("cw_2d_hetero", "linear", 1.5)]


@pytest.mark.parametrize("name,solver,y", FULL_CASES)
def test_full_solve_against_oracle(name, solver, y):
    import caustica_amd.solvers as solvers

    print(f"[staggered solve] {name} {solver} y={y}", flush=True)
    g, grid, medium, source, spec, kw, _s, args = product_case(name)
    res = solvers.get(solver)().run(grid, medium, source, spec, backend="hip", scheme="staggered",
                                    alpha_power=y, **kw)
    ref = _oracle_for(grid, medium, source, spec, solver == "westervelt", y, **kw).run()
    assert res.spp == ref.spp and res.steps_total == ref.steps_total
    assert res.converged_period == ref.converged_period
    assert res.meta["scheme"] == "cw-kspace-pstd-staggered/1"
    e1 = rel_l2(res.phasors[1], ref.phasors[1])
    ep = rel_l2(res.p_max, ref.p_max)
    assert e1 < 1e-4, f"fundamental {e1:.3e}"
    assert ep < 1e-4, f"p_max {ep:.3e}"
    for h in kw["harmonics"][1:]:
        eh = rel_l2(res.phasors[h], ref.phasors[h])
        assert eh < 1e-3, f"harmonic {h}: {eh:.3e}"
    # the staggered solve is a different scheme: it must not coincide with the reference one
    base = solvers.get(solver)().run(grid, medium, source, spec, backend="hip", **kw)
    assert rel_l2(base.phasors[1], res.phasors[1]) > 1e-4


def _run_1d(f0=1e6, alpha=0.0, y=1.5, n=512, dx=0.1e-3, pml=20, src=40, amp=1e5, beta=0.0,
            nonlinear=False, harmonics=(1,), settle=None):
    import caustica_amd.solvers as solvers
    from caustica_amd import CWSource, Grid, Medium, PMLSpec
    from caustica_amd.solvers import CWRunSpec

    tof = int(np.ceil((n - src) * dx / C0 * f0)) + 2
    settle = settle if settle is not None else tof + 4
    spec = CWRunSpec(cfl=0.3, cfl_hard_max=0.5, min_settle_periods=settle, max_settle_periods=settle,
                     convergence_tol=0.0, n_record_periods=2)
    grid = Grid((n,), dx, PMLSpec(thickness=pml * dx))
    medium = Medium(alpha=np.full(n, alpha, np.float32), rho=np.full(n, 1000.0, np.float32),
                    c=np.full(n, C0, np.float32), beta=np.full(n, beta, np.float32))
    source = CWSource(indices=np.array([[src]]), phases=np.zeros(1, np.float32), amplitude=amp, f0=f0)
    solver = "westervelt" if nonlinear else "linear"
    return solvers.get(solver)().run(grid, medium, source, spec, backend="hip", harmonics=harmonics,
                                     scheme="staggered", alpha_power=y)


def _fit(ph, x0=150, x1=400, dx=0.1e-3):
    """(alpha, k) of a travelling wave P = A exp((i k - alpha) x) by least squares."""
    x = np.arange(x0, x1) * dx
    a = np.vstack([np.ones_like(x), x]).T
    lam = np.linalg.lstsq(a, np.log(np.abs(ph[x0:x1])), rcond=None)[0][1]
    k = np.linalg.lstsq(a, np.unwrap(np.angle(ph[x0:x1])), rcond=None)[0][1]
    return -lam, k


def test_lossless_plane_wave_speed_and_pml_reflection():
    print("[staggered] lossless 1-D plane wave", flush=True)
    res = _run_1d()
    alpha, k = _fit(res.phasors[1])
    assert abs(2 * np.pi * 1e6 / k / C0 - 1.0) < 1e-4
    assert abs(alpha) < 1e-3
    a = np.abs(res.phasors[1][150:400])
    assert (a.max() - a.min()) / (a.max() + a.min()) < 1e-4


@pytest.mark.parametrize("y", [1.5, 1.1])
def test_power_law_absorption_and_dispersion(y):
    print(f"[staggered] power law y={y}", flush=True)
    meas = {}
    for f0 in (1e6, 2e6):
        res = _run_1d(f0=f0, alpha=10.0 * (f0 / 1e6) ** y, y=y)
        meas[f0] = _fit(res.phasors[1])
        k = powerlaw_wavenumber(f0, 10.0, 1e6, y)
        assert abs(meas[f0][0] / k.imag - 1.0) < 0.02, (f0, meas[f0][0], k.imag)
        cp, cp_want = 2 * np.pi * f0 / meas[f0][1], 2 * np.pi * f0 / k.real
        assert abs(cp / cp_want - 1.0) < 6e-4, (f0, cp, cp_want)
    slope = np.log(meas[2e6][0] / meas[1e6][0]) / np.log(2.0)
    assert abs(slope - y) < 0.03, slope


def test_fubini_harmonics():
    print("[staggered] Fubini", flush=True)
    dx = C0 / (1e6 * 16.0)
    res = _run_1d(n=600, dx=dx, pml=40, src=60, amp=2e6, beta=3.5, nonlinear=True,
                  harmonics=(1, 2, 3), settle=45)
    a1, a2, a3 = (np.abs(res.phasors[h]) for h in (1, 2, 3))
    sig0 = (30 * dx) / shock_distance(a1[90], 1e6, C0, beta=3.5)
    x_sh = shock_distance(a1[90] / fubini(1, sig0), 1e6, C0, beta=3.5)
    for iv in (150, 250, 350, 450, 540):
        s = ((iv - 60) * dx) / x_sh
        assert abs(a2[iv] / a1[iv] / (fubini(2, s) / fubini(1, s)) - 1.0) < 0.03, s
        assert abs(a3[iv] / a1[iv] / (fubini(3, s) / fubini(1, s)) - 1.0) < 0.05, s


def test_beta_zero_is_the_linear_solve_bitwise():
    print("[staggered] beta = 0", flush=True)
    lin = _run_1d(n=128, src=30, settle=8)
    nl = _run_1d(n=128, src=30, settle=8, nonlinear=True, beta=0.0)
    assert np.array_equal(lin.phasors[1], nl.phasors[1])
    assert np.array_equal(lin.p_max, nl.p_max)


def test_batched_equals_alone_and_raw_state_resumes_bitwise():
    from caustica_amd.solvers import CWRunSpec
    from caustica_amd.solvers.kspace.native import KSpaceSession
    from caustica_amd.solvers.kspace.staggered import plan_staggered

    print("[staggered] batched == alone, raw resume", flush=True)
    spec = CWRunSpec(cfl=0.3, min_settle_periods=4, max_settle_periods=8)
    plans = []
    for seed in (1, 2, 3):
        grid, medium, source = _medium_case((16, 12, 24), True, True, seed=seed)
        plans.append(plan_staggered(grid, medium, source, spec, nonlinear=True, alpha_power=1.4))
    steps = 5 * plans[0].spp
    rec_from = steps - 2 * plans[0].spp
    alone = []
    for pl in plans:
        with KSpaceSession(pl.padded3, 1, scheme="staggered") as s:
            slot = s.add(pl)
            s.schedule(slot, rec_from)
            s.advance([slot], steps)
            alone.append(s.result(slot, 0, 0.5)[0])
    with KSpaceSession(plans[0].padded3, 3, scheme="staggered") as s:
        slots = [s.add(pl) for pl in plans]
        for sl in slots:
            s.schedule(sl, rec_from)
        s.advance(slots, rec_from - 7)
        raw, n = s.get_raw(slots[1])
        s.advance(slots, steps - (rec_from - 7))
        together = [s.result(sl, 0, 0.5)[0] for sl in slots]
        # resume member 1 from its raw state: the same record window bit for bit
        s.reset(slots[1])
        s.set_raw(slots[1], n, raw)
        s.schedule(slots[1], rec_from)
        s.advance([slots[1]], steps - n)
        resumed = s.result(slots[1], 0, 0.5)[0]
    for a, b in zip(alone, together):
        assert np.array_equal(a, b)
    assert np.array_equal(resumed, alone[1])
    assert all(np.isfinite(a).all() and np.abs(a).max() > 0 for a in alone)


def test_scheme_mismatch_is_refused():
    from caustica_amd import _native as N
    from caustica_amd.solvers import CWRunSpec
    from caustica_amd.solvers.kspace.native import KSpaceSession
    from caustica_amd.solvers.kspace.operators import plan_member
    from caustica_amd.solvers.kspace.staggered import plan_staggered

    grid, medium, source = _medium_case((16, 12, 24), True, False)
    spec = CWRunSpec(cfl=0.3)
    ps = plan_staggered(grid, medium, source, spec)
    pr = plan_member(grid, medium, source, spec)
    with KSpaceSession(ps.padded3, 1, scheme="staggered") as s:
        with pytest.raises(ValueError):
            s.add(pr)
    with KSpaceSession(pr.padded3, 1) as s:
        with pytest.raises(ValueError):
            s.add(ps)
        pr.scheme = "staggered"  # the library refuses it too
        pr.stg = ps.stg
        with pytest.raises(N.NativeError):
            s.scheme = "staggered"
            s.add(pr)


def _stg_lengths(axis):
    from caustica_amd import _native as N

    lib = N.lib()
    if axis < 2:
        return [n for n in range(4, 513) if lib.cst_supported_length(n, axis)]
    return [n for n in range(4, 1025, 2) if lib.cst_supported_length(n, 2)]


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_every_supported_length_one_step(axis):
    """One absorbing, nonlinear staggered step from a random state at every instantiated
    axis length (the other axes tiny), against the float64 oracle: every FFT split, the
    shift factors on odd and even lengths, the PML tables and the real-split z transform
    of the staggered kernels."""
    from caustica_amd import CWSource, Grid, Medium, PMLSpec
    from caustica_amd.solvers import CWRunSpec
    from caustica_amd.solvers.kspace.native import KSpaceSession
    from caustica_amd.solvers.kspace.staggered import plan_staggered

    bad = []
    for n in _stg_lengths(axis):
        shape = [4, 4, 8]
        shape[axis] = n
        shape = tuple(shape)
        print(f"[staggered lengths] axis {axis} n {n} shape {shape}", flush=True)
        rng = np.random.default_rng(n)
        c = (1500.0 + 80.0 * rng.standard_normal(shape)).astype(np.float32)
        rho = (1000.0 + 60.0 * rng.standard_normal(shape)).astype(np.float32)
        alpha = (5.0 + 3.0 * rng.random(shape)).astype(np.float32)
        beta = (3.5 + rng.random(shape)).astype(np.float32)
        flat = rng.choice(int(np.prod(shape)), size=5, replace=False)
        idx = np.stack(np.unravel_index(flat, shape), axis=1).astype(np.int64)
        grid = Grid(shape, 2e-4, PMLSpec(thickness=4e-4))
        medium = Medium(alpha=alpha, rho=rho, c=c, beta=beta)
        source = CWSource(indices=idx, phases=rng.uniform(0, 6.28, 5).astype(np.float32),
                          amplitude=1e5, f0=1e6, ramp_periods=3.0)
        spec = CWRunSpec(cfl=0.3, convergence_tol=0.0, min_settle_periods=2, max_settle_periods=2,
                         n_record_periods=1)
        plan = plan_staggered(grid, medium, source, spec, nonlinear=True, alpha_power=1.4)
        assert plan.padded == shape
        orc = _oracle_for(grid, medium, source, spec, True, 1.4)
        p0 = rng.standard_normal(shape).astype(np.float32) * 1e4
        u0 = [rng.standard_normal(shape).astype(np.float32) * 1e-2 for _ in range(3)]
        r0 = [rng.standard_normal(shape).astype(np.float32) * 1e-3 for _ in range(3)]
        n0 = plan.spp + 2  # inside the ramp
        orc.p = p0.astype(np.float64)
        for i in range(3):
            orc.u[i] = u0[i].astype(np.float64)
            orc.rho_a[i] = r0[i].astype(np.float64)
        orc.step(n0)
        with KSpaceSession(plan.padded3, 1, scheme="staggered") as s:
            slot = s.add(plan)
            s.set_state(slot, n0, p0, *u0)
            s.set_density(slot, *r0)
            s.advance([slot], 1)
            (p1, *u1), _ = s.get_state(slot)
            r1 = s.get_density(slot)
        errs = [np.abs(p1 - orc.p).max() / np.abs(orc.p).max()]
        for i in range(3):
            errs.append(np.abs(u1[i] - orc.u[i]).max() / np.abs(orc.u[i]).max())
            errs.append(np.abs(r1[i] - orc.rho_a[i]).max() / np.abs(orc.rho_a[i]).max())
        if not max(errs) < 2e-5:
            bad.append((n, max(errs)))
    assert not bad, f"axis {axis}: lengths over the bar: {bad}"
