"""DLPack exchange with the library on the host (no GPU needed for host tensors).

numpy arrays are borrowed through their ``__dlpack__`` capsule; tensors the library
allocates (``cst_dl_alloc``) come back as ``DLManagedTensor`` the caller owns and can
hand to numpy (``np.from_dlpack`` consumes them: numpy's array then owns the memory and
calls the deleter) or release. ``cst_dl_copy`` refuses what the entry points refuse:
strided views, another dtype, another element count.
"""

import ctypes as C
import gc

import numpy as np
import pytest

from caustica_amd import _native as N
from caustica_amd import dlpack as D


def test_host_export_round_trip_through_numpy():
    src = np.arange(2 * 3 * 4, dtype=np.float32).reshape(2, 3, 4)
    t = D.empty(src.shape, np.float32)
    assert t.shape == (2, 3, 4) and not t.on_device
    with D.borrow(src) as s:
        N.check(N.lib().cst_dl_copy(s, C.pointer(t.tensor)))
    np.testing.assert_array_equal(t.to_numpy(), src)
    got = np.from_dlpack(t)          # consumes the export: numpy owns it now
    np.testing.assert_array_equal(got, src)
    with pytest.raises(ValueError):
        t.__dlpack__()
    del got
    gc.collect()


def test_unconsumed_exports_are_released():
    for _ in range(50):
        t = D.empty((64, 64), np.complex64)
        cap = t.__dlpack__()         # capsule never consumed: its destructor frees it
        del cap
        u = D.empty((8,), np.float64)
        u.close()
        u.close()                    # idempotent
    gc.collect()


@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.complex64, np.int32])
def test_dtypes(dtype):
    src = (np.arange(10) * 3).astype(dtype)
    t = D.empty(src.shape, dtype)
    with D.borrow(src) as s:
        N.check(N.lib().cst_dl_copy(s, C.pointer(t.tensor)))
    np.testing.assert_array_equal(t.to_numpy(), src)


def test_refusals_name_the_problem():
    lib = N.lib()
    dst = D.empty((4, 4), np.float32)
    base = np.zeros((4, 8), np.float32)
    with D.borrow(base[:, ::2]) as s:           # strided view
        assert lib.cst_dl_copy(s, C.pointer(dst.tensor)) == N.CST_ERR_ARG
    assert "strided" in N.last_error()
    with D.borrow(np.zeros((4, 4), np.float64)) as s:
        assert lib.cst_dl_copy(s, C.pointer(dst.tensor)) == N.CST_ERR_ARG
    assert "float64" in N.last_error() and "float32" in N.last_error()
    with D.borrow(np.zeros(15, np.float32)) as s:
        assert lib.cst_dl_copy(s, C.pointer(dst.tensor)) == N.CST_ERR_ARG
    assert "elements" in N.last_error()


def test_read_only_arrays_are_borrowed_as_copies():
    ro = np.ones(6, np.float32)
    ro.flags.writeable = False
    t = D.empty((6,), np.float32)
    with D.borrow(ro) as s:
        N.check(N.lib().cst_dl_copy(s, C.pointer(t.tensor)))
    np.testing.assert_array_equal(t.to_numpy(), ro)


def test_borrowing_an_export_keeps_it():
    t = D.empty((5,), np.float32)
    for _ in range(3):
        with D.borrow(t) as p:
            assert p.contents.data == t.tensor.data
    assert t.to_numpy().shape == (5,)
