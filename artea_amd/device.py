"""Device memory, streams and events of libartea_hip.so's own HIP runtime.

Callers that also use a GPU framework must not mix its streams or device
buffers with these: a framework that bundles a different HIP runtime would
put two runtimes in one process.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, lib


class DeviceBuffer:
    def __init__(self, nbytes: int, dtype=np.uint8, shape=None):
        p = C.c_void_p()
        check(lib().artea_device_alloc(max(int(nbytes), 1), C.byref(p)))
        self.ptr = p.value
        self.nbytes = int(nbytes)
        self.dtype = np.dtype(dtype)
        self.shape = shape if shape is not None else (self.nbytes // self.dtype.itemsize,)

    @classmethod
    def from_numpy(cls, a: np.ndarray, stream=None):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes, a.dtype, a.shape)
        b.upload(a, stream)
        return b

    @classmethod
    def borrow(cls, ptr: int, shape, dtype):
        """Non-owning view of device memory owned elsewhere (e.g. an index's vectors)."""
        b = cls.__new__(cls)
        b.dtype = np.dtype(dtype)
        b.shape = tuple(shape)
        b.nbytes = int(np.prod(shape)) * b.dtype.itemsize
        b.ptr = ptr
        b._owned = False
        return b

    @classmethod
    def empty(cls, shape, dtype):
        dt = np.dtype(dtype)
        n = int(np.prod(shape)) * dt.itemsize
        return cls(n, dt, tuple(shape))

    def upload(self, a: np.ndarray, stream=None):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        check(lib().artea_memcpy_htod(C.c_void_p(self.ptr), a.ctypes.data_as(C.c_void_p), a.nbytes,
                                      C.c_void_p(stream.h if stream else None)))

    def numpy(self, stream=None) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        check(lib().artea_memcpy_dtoh(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), out.nbytes,
                                      C.c_void_p(stream.h if stream else None)))
        if stream:
            stream.synchronize()
        return out

    def zero(self):
        check(lib().artea_memset_device(C.c_void_p(self.ptr), 0, self.nbytes, None))

    def free(self):
        if getattr(self, "ptr", None) and getattr(self, "_owned", True):
            lib().artea_device_free(C.c_void_p(self.ptr))
        self.ptr = None

    def __del__(self):
        self.free()


class Stream:
    def __init__(self):
        p = C.c_void_p()
        check(lib().artea_stream_create(C.byref(p)))
        self.h = p.value

    def synchronize(self):
        check(lib().artea_stream_synchronize(C.c_void_p(self.h)))

    def __del__(self):
        if getattr(self, "h", None):
            lib().artea_stream_destroy(C.c_void_p(self.h))
            self.h = None


class Event:
    def __init__(self):
        p = C.c_void_p()
        check(lib().artea_event_create(C.byref(p)))
        self.h = p.value

    def record(self, stream: Stream):
        check(lib().artea_event_record(C.c_void_p(self.h), C.c_void_p(stream.h)))

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float()
        check(lib().artea_event_elapsed_ms(C.c_void_p(self.h), C.c_void_p(end.h), C.byref(ms)))
        return ms.value

    def __del__(self):
        if getattr(self, "h", None):
            lib().artea_event_destroy(C.c_void_p(self.h))
            self.h = None


def device_info():
    name = C.create_string_buffer(64)
    cus = C.c_int()
    total = C.c_size_t()
    check(lib().artea_device_info(name, 64, C.byref(cus), C.byref(total)))
    return {"arch": name.value.decode(), "compute_units": cus.value, "total_bytes": total.value}


def bruteforce_knn(metric: int, base: DeviceBuffer, n: int, queries: DeviceBuffer, nq: int, dim_pad: int, k: int,
                   stream: Stream = None, method: str = "fp64", stats: dict = None):
    """Exact top-k by (FP64 distance, vid); returns host (ids, distances).
    method "fp64": the FP64 kernel; "mfma": the BF16 MFMA filter with FP64
    rechecks (same result; ``stats`` receives its counters)."""
    ov = DeviceBuffer.empty((nq, k), np.uint32)
    od = DeviceBuffer.empty((nq, k), np.float64)
    sh = C.c_void_p(stream.h if stream else None)
    if method == "mfma":
        st = np.zeros(3, np.uint64)
        check(lib().artea_bruteforce_knn_mfma_device(metric, C.c_void_p(base.ptr), n, C.c_void_p(queries.ptr), nq,
                                                     dim_pad, k, C.c_void_p(ov.ptr), C.c_void_p(od.ptr), sh,
                                                     st.ctypes.data_as(C.c_void_p)))
        if stats is not None:
            stats.update(kept_pairs=int(st[0]), fp64_fallback_queries=int(st[1]), sample_rows=int(st[2]))
    elif method == "fp64":
        check(lib().artea_bruteforce_knn_device(metric, C.c_void_p(base.ptr), n, C.c_void_p(queries.ptr), nq, dim_pad,
                                                k, C.c_void_p(ov.ptr), C.c_void_p(od.ptr), sh))
    else:
        raise ValueError("method must be 'fp64' or 'mfma'")
    if stream:
        stream.synchronize()
    return ov.numpy(), od.numpy()
