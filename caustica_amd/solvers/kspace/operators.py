"""Host-side setup of one CW k-space solve: padding, coefficient maps, k tables, sponge.

Follows upstream ``src/caustica/solvers/kspace/operators.py:23-157`` and the setup
block of ``engine.py:233-301``. Everything here runs once per solve on the host and
produces the float32 arrays the device consumes; the dense kappa / derivative-factor
/ sponge volumes of upstream are NOT built: the device forms them per bin/voxel from
the 1-D tables below.

Grids of rank 1 and 2 are embedded as (1, 1, n) and (1, n0, n1): the real (rfft)
axis stays the last one, as upstream.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from itertools import product
from math import ceil, floor

import numpy as np

from caustica_amd.core.pml import sponge_profile_1d


def optimal_fft_size(n: int) -> int:
    """Smallest size >= n whose prime factors are all in {2, 3, 5}."""
    if n < 1:
        raise ValueError(f"n must be >= 1, got {n}")
    while True:
        m = n
        for p in (2, 3, 5):
            while m % p == 0:
                m //= p
        if m == 1:
            return n
        n += 1


def pad_shape(shape) -> tuple:
    return tuple(optimal_fft_size(int(n)) for n in shape)


def pad_volume(vol: np.ndarray, padded) -> np.ndarray:
    """Edge-replicate ``vol`` up to ``padded`` at the high end only."""
    pad = []
    for have, want in zip(vol.shape, padded):
        if want < have:
            raise ValueError(f"padded shape {padded} smaller than volume {vol.shape}")
        pad.append((0, want - have))
    if all(p == (0, 0) for p in pad):
        return vol
    return np.pad(vol, pad, mode="edge")


def embed3(shape) -> tuple:
    """(n,) -> (1, 1, n); (n0, n1) -> (1, n0, n1); 3-D unchanged."""
    shape = tuple(int(s) for s in shape)
    return (1,) * (3 - len(shape)) + shape


def k_tables(padded3, dx: float) -> np.ndarray:
    """float32 wavenumbers per axis: k, k_d (Nyquist zeroed on even axes), rfft on z.

    k = float32(2 pi fftfreq(n, dx)) as upstream k_vectors (operators.py:60-73); the
    derivative table zeroes index n/2 of every even axis (operators.py:134-142).
    """
    nx, ny, nz = padded3
    out = []
    for ax, n in enumerate((nx, ny, nz)):
        f = np.fft.rfftfreq(n, d=dx) if ax == 2 else np.fft.fftfreq(n, d=dx)
        k = (2.0 * np.pi * f).astype(np.float32)
        kd = k.copy()
        if n % 2 == 0:
            kd[n // 2] = 0.0
        out += [k, kd]
    return np.ascontiguousarray(np.concatenate(out), dtype=np.float32)


def sponge_tables(padded3, nd: int, width: int, edge: float) -> np.ndarray:
    """sz, sx, sy profiles (float32); embedded unit axes are all ones."""
    prof = []
    for ax, n in enumerate(padded3):
        if ax < 3 - nd:
            prof.append(np.ones(n, np.float32))
        else:
            prof.append(sponge_profile_1d(n, width, edge))
    return np.ascontiguousarray(np.concatenate([prof[2], prof[0], prof[1]]), dtype=np.float32)


def cw_discretization(dx: float, c_max: float, spec, f0: float, harmonics=(1,)):
    """Exact-period (spp, dt) under the CFL limit (engine.py:86-113)."""
    period = 1.0 / f0
    dt_cfl = spec.cfl * dx / c_max
    spp = max(2, int(floor(period / dt_cfl)))
    dt = period / spp
    while dt * c_max / dx > spec.cfl_hard_max + 1e-12 and spp < 1024:
        spp += 1
        dt = period / spp
    h_max = max(harmonics)
    if 2 * h_max >= spp:
        raise ValueError(
            f"harmonic {h_max} is at/above the temporal Nyquist for spp={spp} (need 2*h < spp); "
            f"refine dx or lower cfl to raise spp."
        )
    return spp, dt


def cw_tof_periods(shape, dx: float, c_min: float, indices: np.ndarray, f0: float,
                   reference_point=None) -> int:
    """Periods of flight from the source to the farthest reference (engine.py:116-134)."""
    period = 1.0 / f0
    src_pos = indices.astype(np.float64) * dx
    if reference_point is not None:
        refs = np.asarray(reference_point, np.float64)[None, :] * dx
    else:
        refs = np.array(list(product(*((0, n - 1) for n in shape))), np.float64) * dx
    dmax = max(float(np.sqrt(((src_pos - r) ** 2).sum(axis=1)).max()) for r in refs)
    return max(1, int(ceil(dmax / c_min / period)))


def normalize_record_region(region, shape) -> tuple:
    if len(region) != len(shape):
        raise ValueError(f"record_region rank {len(region)} != grid rank {len(shape)}")
    out = []
    for sl, n_ax in zip(region, shape):
        start, stop, step = sl.indices(n_ax)
        if step != 1:
            raise ValueError(f"record_region does not support strides, got {sl}")
        if stop <= start:
            raise ValueError(f"record_region slice {sl} is empty on an axis of {n_ax}")
        out.append(slice(start, stop))
    return tuple(out)


@dataclass
class MemberPlan:
    """Everything one solve needs on the device, plus its control schedule."""

    shape: tuple            # active grid shape (rank 1..3)
    padded: tuple           # padded shape (rank of the grid)
    padded3: tuple          # padded shape embedded in 3-D
    dx: float
    pml_vox: int
    pml_edge: float
    spp: int
    dt: float
    f0: float
    c_ref: float
    amplitude: float
    ramp_periods: float
    dt_over_rho: np.ndarray
    rhoc2_dt: np.ndarray
    absorb: np.ndarray
    beta2_dt: np.ndarray | None
    ktab: np.ndarray
    sponge: np.ndarray
    src_idx3: np.ndarray    # (n, 3) int32 in the embedded frame
    src_phase: np.ndarray
    src_scale: np.ndarray
    conv: tuple             # slices over the active grid
    rec: tuple
    conv3: tuple            # (x0, x1, y0, y1, z0, z1)
    rec3: tuple
    harmonics: tuple
    tof_periods: int
    eff_min: int
    eff_max: int
    label: str = ""
    meta: dict = field(default_factory=dict)
    #: "reference" (the upstream scheme) or "staggered" (the opt-in scheme, whose maps and
    #: tables are in ``stg``; see caustica_amd.solvers.kspace.staggered)
    scheme: str = "reference"
    stg: dict | None = None

    @property
    def rec_shape(self) -> tuple:
        return tuple(s.stop - s.start for s in self.rec)

    def device_bytes(self) -> int:
        """HBM held by the member on the device (admission footprint)."""
        nx, ny, nz = self.padded3
        P = nx * ny * nz
        # one spectrum: rows of nz/2 bins padded to 16, planes + 16, Nyquist side plane
        H = nx * (ny * ((nz // 2 + 15) // 16 * 16) + 16) + (nx * ny + 15) // 16 * 16
        R = int(np.prod(self.rec_shape))
        if self.scheme == "staggered":
            # p, split density rho_a, and the maps given (dt/rho0' x3, dt rho0, c^2, ...)
            nmaps = sum(m is not None for m in self.stg["maps"])
            return int(16 * P + 56 * H + 4 * nmaps * P + R * (16 * len(self.harmonics) + 4) + (1 << 20))
        nmaps = 3 + (self.beta2_dt is not None)
        return int(4 * P + 56 * H + 4 * nmaps * P + R * (16 * len(self.harmonics) + 4) + (1 << 20))


def plan_member(grid, medium, source, spec, *, harmonics=(1,), record_region=None,
                reference_point=None, nonlinear=False) -> MemberPlan:
    """Host setup of one solve, with upstream dtypes (engine.py:233-301)."""
    harmonics = tuple(dict.fromkeys(int(h) for h in harmonics))
    shape = tuple(grid.shape)
    nd = len(shape)
    dx = grid.dx
    c_max = medium.c_max
    spp, dt = cw_discretization(dx, c_max, spec, source.f0, harmonics)
    padded = pad_shape(shape)
    padded3 = embed3(padded)
    # engine.py:239-249 (float32 under NumPy-2 scalar promotion)
    dt_over_rho = np.asarray(dt / pad_volume(medium.rho, padded), dtype=np.float32)
    rhoc2_dt = np.asarray(
        pad_volume(medium.rho * medium.c.astype(np.float64) ** 2, padded) * dt, dtype=np.float32
    )
    absorb = np.asarray(np.exp(-pad_volume(medium.alpha * medium.c, padded) * dt), dtype=np.float32)
    beta2_dt = None
    if nonlinear and not medium.is_linear:
        beta2_dt = np.asarray(pad_volume(medium.beta, padded) * (2.0 * dt), dtype=np.float32)
    pml_edge = grid.pml.edge if grid.pml is not None else 2.0
    ktab = k_tables(padded3, dx)
    sponge = sponge_tables(padded3, nd, grid.pml_vox, pml_edge)
    idx = source.indices
    c_src = medium.c[tuple(idx[:, d] for d in range(nd))]
    src_scale = np.asarray(2.0 * c_src.astype(np.float64) * dt / dx, dtype=np.float32)
    pad_lead = 3 - nd
    src_idx3 = np.zeros((idx.shape[0], 3), np.int32)
    src_idx3[:, pad_lead:] = idx
    tof = cw_tof_periods(shape, dx, medium.c_min, idx, source.f0, reference_point)
    eff_min = tof + max(spec.min_settle_periods, int(ceil(source.ramp_periods)) + 1)
    eff_max = max(tof + spec.max_settle_periods, eff_min)
    margin = grid.pml_vox + 2
    if any(n <= 2 * margin for n in shape):
        margin = 0
    conv = tuple(slice(margin, n - margin) for n in shape)
    active = tuple(slice(0, n) for n in shape)
    rec = normalize_record_region(record_region, shape) if record_region is not None else active

    def box3(sl):
        full = [(0, 1)] * pad_lead + [(s.start, s.stop) for s in sl]
        return tuple(v for pair in full for v in pair)

    return MemberPlan(
        shape=shape, padded=padded, padded3=padded3, dx=dx, pml_vox=grid.pml_vox,
        pml_edge=pml_edge, spp=spp, dt=dt, f0=float(source.f0), c_ref=c_max,
        amplitude=float(source.amplitude), ramp_periods=float(source.ramp_periods),
        dt_over_rho=np.ascontiguousarray(dt_over_rho), rhoc2_dt=np.ascontiguousarray(rhoc2_dt),
        absorb=np.ascontiguousarray(absorb),
        beta2_dt=None if beta2_dt is None else np.ascontiguousarray(beta2_dt),
        ktab=ktab, sponge=sponge, src_idx3=src_idx3,
        src_phase=np.ascontiguousarray(source.phases, np.float32),
        src_scale=np.ascontiguousarray(src_scale), conv=conv, rec=rec,
        conv3=box3(conv), rec3=box3(rec), harmonics=harmonics, tof_periods=tof,
        eff_min=eff_min, eff_max=eff_max, label=source.label,
    )
