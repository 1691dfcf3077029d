"""Minimal HIP runtime access for device buffers, streams and events.

Binds the same libamdhip64 that libdel3d_amd.so links against, so buffers made
here are valid for the library. (PyTorch's ROCm wheel ships its own HIP
runtime under a different soname; loading both into one process would start
two HIP runtimes, so this package does not use torch for device memory.)
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

_HIP = None


def _lib() -> C.CDLL:
    global _HIP
    if _HIP is None:
        cands = [os.environ.get("DEL3D_HIP_RUNTIME", ""),
                 "/opt/rocm/lib/libamdhip64.so.7", "/opt/rocm/lib/libamdhip64.so"]
        for c in cands:
            if c and Path(c).exists():
                _HIP = C.CDLL(c, mode=C.RTLD_GLOBAL)
                break
        if _HIP is None:
            raise RuntimeError("HIP runtime (libamdhip64) not found under /opt/rocm/lib")
        h = _HIP
        h.hipGetErrorString.restype = C.c_char_p
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipHostMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t, C.c_uint]
        h.hipHostFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
        h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        h.hipStreamDestroy.argtypes = [C.c_void_p]
        h.hipStreamSynchronize.argtypes = [C.c_void_p]
        h.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        h.hipEventDestroy.argtypes = [C.c_void_p]
        h.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        h.hipEventSynchronize.argtypes = [C.c_void_p]
        h.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        h.hipGetDeviceCount.argtypes = [C.POINTER(C.c_int)]
        h.hipSetDevice.argtypes = [C.c_int]
        h.hipDeviceGetAttribute.argtypes = [C.POINTER(C.c_int), C.c_int, C.c_int]
    return _HIP


H2D, D2H, D2D = 1, 2, 3


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = _lib().hipGetErrorString(rc).decode()
        raise RuntimeError(f"{what} failed: {msg} ({rc})")


def device_count() -> int:
    try:
        n = C.c_int(0)
        rc = _lib().hipGetDeviceCount(C.byref(n))
        return n.value if rc == 0 else 0
    except (RuntimeError, OSError):
        return 0


def synchronize() -> None:
    check(_lib().hipDeviceSynchronize(), "hipDeviceSynchronize")


def cu_count(dev: int = 0) -> int:
    v = C.c_int(0)
    # hipDeviceAttributeMultiprocessorCount
    check(_lib().hipDeviceGetAttribute(C.byref(v), 63, dev), "hipDeviceGetAttribute")
    return v.value


class DeviceArray:
    """A device allocation with a numpy dtype and shape."""

    def __init__(self, shape, dtype):
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        self.ptr = C.c_void_p(0)
        if self.nbytes:
            check(_lib().hipMalloc(C.byref(self.ptr), self.nbytes), "hipMalloc")

    @classmethod
    def from_host(cls, a: np.ndarray, stream=None) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a, stream)
        return d

    def upload(self, a: np.ndarray, stream=None) -> None:
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        if self.nbytes:
            if stream is None:
                check(_lib().hipMemcpy(self.ptr, a.ctypes.data, self.nbytes, H2D), "hipMemcpy H2D")
            else:
                check(_lib().hipMemcpyAsync(self.ptr, a.ctypes.data, self.nbytes, H2D, stream.handle),
                      "hipMemcpyAsync H2D")
                stream.synchronize()

    def to_host(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        if self.nbytes:
            check(_lib().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H), "hipMemcpy D2H")
        return out

    def zero(self) -> None:
        if self.nbytes:
            check(_lib().hipMemset(self.ptr, 0, self.nbytes), "hipMemset")

    @property
    def addr(self) -> int:
        return self.ptr.value or 0

    def free(self) -> None:
        if self.ptr.value:
            _lib().hipFree(self.ptr)
            self.ptr = C.c_void_p(0)

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class _PinnedBlock:
    """Owns one hipHostMalloc block; freed when the last array using it goes."""

    def __init__(self, nbytes: int):
        self.ptr = C.c_void_p(0)
        check(_lib().hipHostMalloc(C.byref(self.ptr), max(1, nbytes), 0), "hipHostMalloc")

    def __del__(self):
        try:
            if self.ptr.value:
                _lib().hipHostFree(self.ptr)
        except Exception:
            pass


class PinnedArray(np.ndarray):
    """numpy array over pinned host memory; views keep their parent (and so
    the pinned block) alive through numpy's base chain."""


def pinned_empty(shape, dtype) -> np.ndarray:
    """A numpy array in pinned (page-locked) host memory: copies between it
    and the device run as asynchronous DMA without staging."""
    dtype = np.dtype(dtype)
    shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
    nbytes = int(np.prod(shape, dtype=np.int64)) * dtype.itemsize
    blk = _PinnedBlock(nbytes)
    buf = (C.c_char * max(1, nbytes)).from_address(blk.ptr.value)
    arr = np.frombuffer(buf, dtype=np.uint8, count=nbytes).view(PinnedArray)
    arr._pinned_block = blk
    return arr.view(dtype).reshape(shape) if nbytes else arr[:0].view(dtype).reshape(shape)


def pinned_copy(a) -> np.ndarray:
    a = np.ascontiguousarray(a)
    out = pinned_empty(a.shape, a.dtype)
    out[...] = a
    return out


class Stream:
    def __init__(self):
        self.handle = C.c_void_p(0)
        check(_lib().hipStreamCreate(C.byref(self.handle)), "hipStreamCreate")

    def synchronize(self) -> None:
        check(_lib().hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    def __del__(self):
        try:
            if self.handle.value:
                _lib().hipStreamDestroy(self.handle)
        except Exception:
            pass


class Event:
    def __init__(self):
        self.handle = C.c_void_p(0)
        check(_lib().hipEventCreate(C.byref(self.handle)), "hipEventCreate")

    def record(self, stream: Stream | None = None) -> None:
        check(_lib().hipEventRecord(self.handle, stream.handle if stream else None), "hipEventRecord")

    def synchronize(self) -> None:
        check(_lib().hipEventSynchronize(self.handle), "hipEventSynchronize")

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float(0)
        check(_lib().hipEventElapsedTime(C.byref(ms), self.handle, end.handle), "hipEventElapsedTime")
        return float(ms.value)

    def __del__(self):
        try:
            if self.handle.value:
                _lib().hipEventDestroy(self.handle)
        except Exception:
            pass
