"""Device-memory plumbing for the GPU tests: buffers allocated and copied through
the library's own HIP runtime (phx_device_malloc / phx_memcpy_*), whose
addresses are passed to the kernels through the C ABI."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import phosphor_amd as P


def need_gpu():
    if P.lib().phx_device_count() < 1:
        pytest.fail("GPU test run without a HIP device")


class DevBuf:
    """One device allocation (freed when the object goes away)."""

    def __init__(self, nbytes: int, fill: int | None = 0):
        need_gpu()
        self.nbytes = max(int(nbytes), 16)
        p = C.c_void_p()
        P.check(P.lib().phx_device_malloc(C.byref(p), self.nbytes), "device_malloc")
        self.ptr = int(p.value)
        if fill is not None:
            P.check(P.lib().phx_memset_device(self.ptr, fill, self.nbytes), "memset")

    def __del__(self):
        try:
            if self.ptr:
                P.lib().phx_device_free(self.ptr)
                self.ptr = 0
        except Exception:
            pass


def to_dev(a: np.ndarray) -> DevBuf:
    a = np.ascontiguousarray(a)
    b = DevBuf(a.nbytes, fill=None)
    if a.nbytes:
        P.check(P.lib().phx_memcpy_htod(b.ptr, a.ctypes.data, a.nbytes), "htod")
    return b


def empty_dev(nbytes: int, fill: int | None = 0) -> DevBuf:
    return DevBuf(nbytes, fill)


def from_dev(b: DevBuf, dtype, shape) -> np.ndarray:
    P.check(P.lib().phx_device_synchronize(), "synchronize")
    out = np.zeros(shape, dtype)
    if out.nbytes:
        P.check(P.lib().phx_memcpy_dtoh(out.ctypes.data, b.ptr, out.nbytes), "dtoh")
    return out


def dptr(b: DevBuf) -> int:
    return b.ptr
