"""Python handle on a libcaustica_hip session (include/caustica_hip.h).

A session owns up to ``max_members`` same-shape solves on one HIP stream; a single
solve is a one-member session. Upstream counterpart: the per-solve device state of
``engine.py:225-278`` (maps, operators, state p/u, source, record buffers).

Volumes cross the C ABI as DLPack tensors: the coefficient maps, the solver state and
the record-window results may be host arrays (numpy) or device tensors (any DLPack
producer on the session's GPU, or ``caustica_amd.dlpack`` exports), borrowed for the
call; ``result_export`` hands out new tensors on the host or the device.
"""

from __future__ import annotations

import contextlib
import ctypes as C
import threading

import numpy as np

from caustica_amd import _native as N
from caustica_amd import dlpack as D
from caustica_amd.solvers.kspace.operators import MemberPlan

KERNEL_KINDS = ("xgrad", "y3", "zu", "xdiv", "y1", "zp", "finalize")


class KSpaceSession:
    """``scheme``: "reference" (the upstream collocated scheme) or "staggered" (the opt-in
    staggered / power-law / split-PML scheme); a session holds members of one scheme.
    ``graphs_default`` (class-wide) is whether new sessions replay captured HIP graphs of
    the step sequence or launch every pass; the results are the same bits either way.

    Threads: the calls that change the session's member table, block pool or graph cache
    (add, remove, advance, trim, pooled_bytes) hold ``lock`` (reentrant). A thread that
    retires finished members holds it across its read-back calls and ``remove``, so that
    no step sequence is being captured on the session meanwhile; another thread can
    admit and advance members of other sessions, and skip this one while it is taken
    (``lock.acquire(blocking=False)``)."""

    graphs_default = True

    def __init__(self, padded3, max_members: int = 1, device: int = 0, scheme: str = "reference"):
        self.padded3 = tuple(int(v) for v in padded3)
        self.device = int(device)
        if scheme not in ("reference", "staggered"):
            raise ValueError(f"scheme must be 'reference' or 'staggered', got {scheme!r}")
        self.scheme = scheme
        self._lib = N.lib()
        h = C.c_void_p()
        code = N.SCHEME_STAGGERED if scheme == "staggered" else N.SCHEME_REFERENCE
        N.check(self._lib.cst_session_create_scheme(*self.padded3, int(max_members), int(device), code,
                                                   C.byref(h)))
        self._h = h
        self.lock = threading.RLock()
        if not KSpaceSession.graphs_default:
            N.check(self._lib.cst_session_set_graphs(self._h, 0))
        self.max_members = int(max_members)
        self._plans: dict[int, MemberPlan] = {}

    # -------------------------------------------------------------- lifetime
    def close(self) -> None:
        if self._h is not None and self._h.value:
            self._lib.cst_session_destroy(self._h)
        self._h = None

    def __del__(self):  # pragma: no cover - best effort
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    def stream(self) -> int:
        s = C.c_void_p()
        N.check(self._lib.cst_session_stream(self._h, C.byref(s)))
        return int(s.value or 0)

    # --------------------------------------------------------------- members
    def add(self, plan: MemberPlan, max_periods: int | None = None, maps=None) -> int:
        """Upload a member. ``maps`` (optional) replaces the plan's host coefficient maps
        by DLPack tensors (dt/rho, rho c^2 dt, absorb[, 2 beta dt]) in host or device
        memory, copied device to device when they are on the GPU."""
        if plan.padded3 != self.padded3:
            raise ValueError(f"plan shape {plan.padded3} != session shape {self.padded3}")
        if plan.scheme != self.scheme:
            raise ValueError(f"a {plan.scheme} plan in a {self.scheme} session")
        d = N.MemberDesc()
        d.dx = plan.dx
        d.pml_vox = plan.pml_vox
        d.pml_edge = plan.pml_edge
        d.dt = plan.dt
        d.c_ref = plan.c_ref
        d.spp = plan.spp
        d.f0 = plan.f0
        d.amplitude = plan.amplitude
        d.ramp_periods = plan.ramp_periods
        keep = []

        def arr(a, dt=np.float32):
            if a is None:
                return None
            b = np.ascontiguousarray(a, dtype=dt)
            keep.append(b)
            return N.ptr(b)

        d.ktab = arr(plan.ktab)
        if plan.scheme == "reference":
            if maps is None:
                maps = [np.ascontiguousarray(a, np.float32) if a is not None else None
                        for a in (plan.dt_over_rho, plan.rhoc2_dt, plan.absorb, plan.beta2_dt)]
            maps = list(maps) + [None] * (4 - len(maps))
            d.sponge = arr(plan.sponge)
        d.n_src = int(plan.src_idx3.shape[0])
        d.src_idx = arr(plan.src_idx3, np.int32)
        d.src_phase = arr(plan.src_phase)
        d.src_scale = arr(plan.src_scale)
        for i in range(6):
            d.conv[i] = int(plan.conv3[i])
            d.rec[i] = int(plan.rec3[i])
        if not 1 <= len(plan.harmonics) <= N.MAX_HARMONICS:
            raise ValueError(f"1..{N.MAX_HARMONICS} harmonics per solve on hip")
        d.n_harmonics = len(plan.harmonics)
        for i, h in enumerate(plan.harmonics):
            d.harmonics[i] = int(h)
        if max_periods is None:
            max_periods = plan.eff_max + 4
        d.max_periods = int(max_periods)
        slot = C.c_int(-1)
        with self.lock:
            self._add_locked(plan, d, maps, arr, slot)
        self._plans[slot.value] = plan
        return slot.value

    def _add_locked(self, plan, d, maps, arr, slot) -> None:
        if plan.scheme == "staggered":
            g = plan.stg
            sd = N.StaggeredDesc()
            sd.kfac = arr(g["kfac"], np.complex64)
            sd.pml = arr(g["pml"])
            sd.alpha_power = float(g["alpha_power"])
            sd.absorbing = 1 if g["absorbing"] else 0
            sd.src_axes = int(g["src_axes"])
            smaps = list(maps) if maps is not None else list(g["maps"])
            with contextlib.ExitStack() as ex:
                ts = [ex.enter_context(D.borrow(a)) for a in smaps]
                tarr = (C.c_void_p * 9)(*[C.cast(t, C.c_void_p) if t is not None else None for t in ts])
                N.check(self._lib.cst_member_add_staggered_dl(self._h, C.byref(d), C.byref(sd), tarr,
                                                              C.byref(slot)))
        else:
            with D.borrow(maps[0]) as m0, D.borrow(maps[1]) as m1, D.borrow(maps[2]) as m2, \
                    D.borrow(maps[3]) as m3:
                N.check(self._lib.cst_member_add_dl(self._h, C.byref(d), m0, m1, m2, m3, C.byref(slot)))

    def remove(self, slot: int) -> None:
        with self.lock:
            N.check(self._lib.cst_member_remove(self._h, int(slot)))
            self._plans.pop(slot, None)

    def trim(self) -> int:
        n = C.c_size_t(0)
        with self.lock:
            N.check(self._lib.cst_session_trim(self._h, C.byref(n)))
        return int(n.value)

    def pooled_bytes(self) -> int:
        n = C.c_size_t(0)
        with self.lock:
            N.check(self._lib.cst_session_pooled_bytes(self._h, C.byref(n)))
        return int(n.value)

    def reset(self, slot: int) -> None:
        N.check(self._lib.cst_member_reset(self._h, int(slot)))

    def plan(self, slot: int) -> MemberPlan:
        return self._plans[slot]

    def schedule(self, slot: int, record_from: int) -> None:
        N.check(self._lib.cst_member_schedule(self._h, int(slot), C.c_int64(int(record_from))))

    def step(self, slot: int) -> int:
        n = C.c_int64()
        N.check(self._lib.cst_member_step(self._h, int(slot), C.byref(n)))
        return int(n.value)

    # --------------------------------------------------------------- running
    def advance(self, slots, n_steps: int) -> None:
        slots = [int(s) for s in slots]
        arr = (C.c_int * len(slots))(*slots)
        with self.lock:
            N.check(self._lib.cst_advance(self._h, len(slots), arr, int(n_steps)))

    def sync(self) -> None:
        N.check(self._lib.cst_sync(self._h))

    def idle(self) -> bool:
        """True when all work queued on the session's stream has completed (non-blocking)."""
        v = C.c_int()
        N.check(self._lib.cst_session_query(self._h, C.byref(v)))
        return bool(v.value)

    def peaks(self, slot: int, first: int, count: int) -> np.ndarray:
        out = np.zeros(max(count, 0), np.float32)
        N.check(self._lib.cst_member_peaks(self._h, int(slot), int(first), int(count), N.ptr(out)))
        return out

    def result(self, slot: int, hi: int, scale: float, with_pmax: bool = True, copy: bool = True):
        """Finalize harmonic ``hi`` (float64 reduction, once) and copy phasor / p_max out;
        ``copy=False`` only finalizes (the phasor stays in HBM for a device-side consumer)."""
        plan = self._plans[slot]
        R = int(np.prod([plan.rec3[1] - plan.rec3[0], plan.rec3[3] - plan.rec3[2],
                         plan.rec3[5] - plan.rec3[4]]))
        if not copy:
            N.check(self._lib.cst_member_result_into(self._h, int(slot), int(hi), float(scale), None, None))
            return None, None
        ph = np.zeros(R, np.complex64)
        pm = np.zeros(R, np.float32) if with_pmax else None
        self.result_into(slot, hi, scale, ph, pm)
        return ph, pm

    def result_into(self, slot: int, hi: int, scale: float, phasor=None, pmax=None) -> None:
        """Finalize harmonic ``hi`` into caller tensors (complex64 / float32 with one
        element per record-box voxel, host or device memory; either may be None)."""
        with D.borrow(phasor) as tp, D.borrow(pmax) as tm:
            N.check(self._lib.cst_member_result_into(self._h, int(slot), int(hi), float(scale), tp, tm))

    def result_export(self, slot: int, hi: int, scale: float, device: bool = True,
                      with_pmax: bool = True):
        """Finalize harmonic ``hi`` into new tensors of the record-box shape on the GPU
        (``device=True``) or the host: (phasor, p_max or None) as ``dlpack.Exported``."""
        a, b = C.c_void_p(), C.c_void_p()
        N.check(self._lib.cst_member_result_export(
            self._h, int(slot), int(hi), float(scale), D.kDLROCM if device else D.kDLCPU, C.byref(a),
            C.byref(b) if with_pmax else None))
        return D.Exported(a.value), (D.Exported(b.value) if with_pmax else None)

    def get_state(self, slot: int, with_u: bool = True):
        """([p, u0, u1, u2], n) as real volumes of the padded shape (u None unless
        ``with_u``); the velocity is formed from its device spectrum."""
        shp = self.padded3
        arrs = [np.empty(shp, np.float32)] + [np.empty(shp, np.float32) if with_u else None
                                              for _ in range(3)]
        n = self.get_state_into(slot, *arrs)
        return arrs, n

    def get_raw(self, slot: int):
        """(bytes of the exact device state, n): restoring it with ``set_raw`` resumes the
        trajectory bit for bit."""
        size = C.c_size_t()
        N.check(self._lib.cst_member_state_size(self._h, C.byref(size)))
        buf = np.empty(size.value, np.uint8)
        n = C.c_int64()
        N.check(self._lib.cst_member_get_raw(self._h, int(slot), N.ptr(buf), buf.nbytes, C.byref(n)))
        return buf, int(n.value)

    def set_raw(self, slot: int, n: int, buf) -> None:
        b = np.ascontiguousarray(buf, np.uint8)
        N.check(self._lib.cst_member_set_raw(self._h, int(slot), N.ptr(b), b.nbytes, C.c_int64(int(n))))

    def set_state(self, slot: int, n: int, p, u0, u1, u2) -> None:
        """Restore (p, u0, u1, u2) at step n: numpy arrays of the padded shape, or any
        DLPack tensors of nx ny nz float32 elements (host or device)."""
        arrs = [np.ascontiguousarray(np.reshape(a, self.padded3), np.float32) if isinstance(a, np.ndarray)
                else a for a in (p, u0, u1, u2)]
        with D.borrow(arrs[0]) as a0, D.borrow(arrs[1]) as a1, D.borrow(arrs[2]) as a2, \
                D.borrow(arrs[3]) as a3:
            N.check(self._lib.cst_member_set_state_dl(self._h, int(slot), C.c_int64(int(n)), a0, a1, a2, a3))

    def get_state_into(self, slot: int, p=None, u0=None, u1=None, u2=None) -> int:
        """Copy the state into caller tensors (host or device, any may be None); returns n."""
        n = C.c_int64()
        with D.borrow(p) as a0, D.borrow(u0) as a1, D.borrow(u1) as a2, D.borrow(u2) as a3:
            N.check(self._lib.cst_member_get_state_into(self._h, int(slot), a0, a1, a2, a3, C.byref(n)))
        return int(n.value)

    def get_density(self, slot: int):
        """[rho_0, rho_1, rho_2] of a staggered member (real volumes of the padded shape)."""
        out = [np.empty(self.padded3, np.float32) for _ in range(3)]
        with D.borrow(out[0]) as a0, D.borrow(out[1]) as a1, D.borrow(out[2]) as a2:
            N.check(self._lib.cst_member_get_density_into(self._h, int(slot), a0, a1, a2))
        return out

    def set_density(self, slot: int, rho0, rho1, rho2) -> None:
        """Restore a staggered member's split density (after ``set_state``)."""
        arrs = [np.ascontiguousarray(np.reshape(a, self.padded3), np.float32) if isinstance(a, np.ndarray)
                else a for a in (rho0, rho1, rho2)]
        with D.borrow(arrs[0]) as a0, D.borrow(arrs[1]) as a1, D.borrow(arrs[2]) as a2:
            N.check(self._lib.cst_member_set_density_dl(self._h, int(slot), a0, a1, a2))

    def set_variant(self, pass_: int, variant: int) -> None:
        N.check(self._lib.cst_session_set_variant(self._h, int(pass_), int(variant)))

    def set_graphs(self, on: bool) -> None:
        """Replay captured HIP graphs of the step sequence (default on) or launch each pass."""
        N.check(self._lib.cst_session_set_graphs(self._h, 1 if on else 0))

    def graph_stats(self) -> dict:
        en, la, ca = C.c_int(), C.c_int64(), C.c_int64()
        N.check(self._lib.cst_session_graph_stats(self._h, C.byref(en), C.byref(la), C.byref(ca)))
        return {"enabled": bool(en.value), "launches": int(la.value), "captures": int(ca.value)}

    # ---------------------------------------------------------------- timing
    def set_profiling(self, on: bool) -> None:
        N.check(self._lib.cst_session_set_profiling(self._h, 1 if on else 0))

    def kernel_times(self) -> dict:
        out = {}
        for k, name in enumerate(KERNEL_KINDS):
            ms = C.c_double()
            cnt = C.c_int64()
            N.check(self._lib.cst_session_kernel_time(self._h, k, C.byref(ms), C.byref(cnt)))
            out[name] = (float(ms.value), int(cnt.value))
        return out
