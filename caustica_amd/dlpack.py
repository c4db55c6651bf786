"""DLPack tensors across the C ABI (include/cst_dlpack.h).

``borrow(obj)`` turns anything that exports DLPack (``__dlpack__``: numpy arrays, torch
tensors on the CPU or the GPU, ...) into a ``DLTensor*`` the library may read or write
for the duration of one call; the producer keeps ownership (the capsule is held, not
consumed, and its own destructor releases the tensor later). ``Exported`` wraps a
``DLManagedTensor*`` the library returns: it is the caller's, exported onwards through
``__dlpack__`` (numpy ``np.from_dlpack`` for host tensors, any DLPack consumer for device
tensors) or released by ``close()``.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from caustica_amd import _native as N

kDLCPU, kDLCUDA, kDLCUDAHost, kDLROCM, kDLROCMHost = 1, 2, 3, 10, 11
kDLInt, kDLUInt, kDLFloat, kDLComplex = 0, 1, 2, 5


class DLDevice(C.Structure):
    _fields_ = [("device_type", C.c_int32), ("device_id", C.c_int32)]


class DLDataType(C.Structure):
    _fields_ = [("code", C.c_uint8), ("bits", C.c_uint8), ("lanes", C.c_uint16)]


class DLTensor(C.Structure):
    _fields_ = [
        ("data", C.c_void_p),
        ("device", DLDevice),
        ("ndim", C.c_int32),
        ("dtype", DLDataType),
        ("shape", C.POINTER(C.c_int64)),
        ("strides", C.POINTER(C.c_int64)),
        ("byte_offset", C.c_uint64),
    ]


class DLManagedTensor(C.Structure):
    pass


DLManagedTensor._fields_ = [
    ("dl_tensor", DLTensor),
    ("manager_ctx", C.c_void_p),
    ("deleter", C.CFUNCTYPE(None, C.POINTER(DLManagedTensor))),
]

_CAPSULE = b"dltensor"
_api = C.pythonapi
_api.PyCapsule_GetPointer.restype = C.c_void_p
_api.PyCapsule_GetPointer.argtypes = [C.py_object, C.c_char_p]
_Destructor = C.CFUNCTYPE(None, C.c_void_p)
_api.PyCapsule_New.restype = C.py_object
_api.PyCapsule_New.argtypes = [C.c_void_p, C.c_char_p, _Destructor]
# the destructor sees the capsule as a borrowed raw pointer: no reference counting
_is_valid_raw = C.PYFUNCTYPE(C.c_int, C.c_void_p, C.c_char_p)(("PyCapsule_IsValid", _api))
_get_raw = C.PYFUNCTYPE(C.c_void_p, C.c_void_p, C.c_char_p)(("PyCapsule_GetPointer", _api))


@_Destructor
def _capsule_destructor(capsule_ptr):
    # an unconsumed capsule still owns its tensor and releases it; a consumer renames
    # the capsule ("used_dltensor") and takes the deleter over
    if _is_valid_raw(capsule_ptr, _CAPSULE):
        p = _get_raw(capsule_ptr, _CAPSULE)
        if p:
            m = C.cast(p, C.POINTER(DLManagedTensor))
            if m.contents.deleter:
                m.contents.deleter(m)


class borrow:
    """Context manager: ``with borrow(arr) as t:`` gives a ``POINTER(DLTensor)`` valid
    inside the block; ``borrow(None)`` gives None."""

    def __init__(self, obj):
        self.obj = obj
        self.capsule = None

    def __enter__(self):
        if self.obj is None:
            return None
        if isinstance(self.obj, Exported):  # our own tensor: lend it without giving it away
            return C.pointer(self.obj.tensor)
        if not hasattr(self.obj, "__dlpack__"):
            raise TypeError(f"{type(self.obj).__name__} does not export DLPack (__dlpack__)")
        if isinstance(self.obj, np.ndarray) and not self.obj.flags.writeable:
            self.obj = self.obj.copy()  # the unversioned protocol cannot export read-only arrays
        self.capsule = self.obj.__dlpack__()
        p = _api.PyCapsule_GetPointer(self.capsule, _CAPSULE)
        m = C.cast(p, C.POINTER(DLManagedTensor))
        return C.pointer(m.contents.dl_tensor)

    def __exit__(self, *exc):
        self.capsule = None  # the producer's destructor releases the tensor
        return False


class Exported:
    """A ``DLManagedTensor*`` the library returned (owned by this object until it is
    exported onwards or closed)."""

    def __init__(self, ptr):
        self._m = C.cast(ptr, C.POINTER(DLManagedTensor)) if ptr else None

    @property
    def tensor(self) -> DLTensor:
        return self._m.contents.dl_tensor

    @property
    def shape(self) -> tuple:
        t = self.tensor
        return tuple(int(t.shape[i]) for i in range(t.ndim))

    @property
    def on_device(self) -> bool:
        return self.tensor.device.device_type == kDLROCM

    def __dlpack_device__(self):
        d = self.tensor.device
        return (int(d.device_type), int(d.device_id))

    def __dlpack__(self, stream=None, **_kw):
        if self._m is None:
            raise ValueError("tensor already exported or closed")
        cap = _api.PyCapsule_New(C.cast(self._m, C.c_void_p), _CAPSULE, _capsule_destructor)
        self._m = None  # the capsule (then its consumer) owns it now
        return cap

    def to_numpy(self) -> np.ndarray:
        """A host copy (device tensors through cst_dl_copy); the tensor stays owned here."""
        t = self.tensor
        dt = {(kDLFloat, 32): np.float32, (kDLFloat, 64): np.float64, (kDLComplex, 64): np.complex64,
              (kDLInt, 32): np.int32}[(t.dtype.code, t.dtype.bits)]
        out = np.empty(self.shape, dt)
        with borrow(out) as dst:
            N.check(N.lib().cst_dl_copy(C.pointer(t), dst))
        return out

    def close(self) -> None:
        if self._m is not None and self._m.contents.deleter:
            self._m.contents.deleter(self._m)
        self._m = None

    def __del__(self):  # pragma: no cover - best effort
        try:
            self.close()
        except Exception:
            pass


def empty(shape, dtype, device: int | None = None) -> Exported:
    """A new library-owned-then-caller-owned tensor on the host (device None) or a GPU."""
    code, bits = {np.dtype(np.float32): (kDLFloat, 32), np.dtype(np.float64): (kDLFloat, 64),
                  np.dtype(np.complex64): (kDLComplex, 64), np.dtype(np.int32): (kDLInt, 32)}[np.dtype(dtype)]
    shp = (C.c_int64 * max(len(shape), 1))(*[int(v) for v in shape])
    out = C.c_void_p()
    N.check(N.lib().cst_dl_alloc(kDLCPU if device is None else kDLROCM, 0 if device is None else int(device),
                                 code, bits, shp, len(shape), C.byref(out)))
    return Exported(out.value)


def to_device(arr: np.ndarray, device: int = 0) -> Exported:
    """Upload a host array into a new device tensor (for callers without a GPU array type)."""
    arr = np.ascontiguousarray(arr)
    t = empty(arr.shape, arr.dtype, device)
    with borrow(arr) as src:
        N.check(N.lib().cst_dl_copy(src, C.pointer(t.tensor)))
    return t
