"""Minimal HIP runtime plumbing (device buffers, streams, events) over the
same libamdhip64 that libdg_amd.so links, so that pointers and stream handles
passed through the C ABI belong to one runtime.  (PyTorch wheels bundle a
second HIP runtime with the same soname; mixing the two in one process is
avoided by using this module instead.)"""
import ctypes as C

import numpy as np

_h = None


def rt():
    global _h
    if _h is None:
        from . import lib
        lib()                                  # loads libamdhip64.so.7 as libdg_amd.so's dependency
        H = C.CDLL("libamdhip64.so.7")         # the already-loaded copy
        vp, sz = C.c_void_p, C.c_size_t
        for name, args in {
            "hipMalloc": [C.POINTER(vp), sz], "hipFree": [vp],
            "hipMemcpy": [vp, vp, sz, C.c_int], "hipMemset": [vp, C.c_int, sz],
            "hipMemcpy2D": [vp, sz, vp, sz, sz, sz, C.c_int],
            "hipDeviceSynchronize": [], "hipGetDeviceCount": [C.POINTER(C.c_int)],
            "hipStreamCreateWithFlags": [C.POINTER(vp), C.c_uint], "hipStreamDestroy": [vp],
            "hipStreamSynchronize": [vp], "hipEventCreate": [C.POINTER(vp)],
            "hipEventDestroy": [vp], "hipEventRecord": [vp, vp], "hipEventSynchronize": [vp],
            "hipEventElapsedTime": [C.POINTER(C.c_float), vp, vp],
        }.items():
            f = getattr(H, name)
            f.argtypes = args
            f.restype = C.c_int
        H.hipGetErrorString.argtypes = [C.c_int]
        H.hipGetErrorString.restype = C.c_char_p
        _h = H
    return _h


def check(e, what=""):
    if e != 0:
        raise RuntimeError(f"HIP error {e} ({rt().hipGetErrorString(e).decode()}) {what}")


def device_count():
    n = C.c_int(0)
    if rt().hipGetDeviceCount(C.byref(n)) != 0:
        return 0
    return n.value


def synchronize():
    check(rt().hipDeviceSynchronize(), "hipDeviceSynchronize")


class Buffer:
    """Device buffer of n elements of dtype (float64 by default)."""

    def __init__(self, n, dtype=np.float64):
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.nbytes = max(self.n, 1) * self.dtype.itemsize
        self.ptr = C.c_void_p()
        check(rt().hipMalloc(C.byref(self.ptr), self.nbytes), "hipMalloc")
        check(rt().hipMemset(self.ptr, 0, self.nbytes), "hipMemset")
        # hipMemset on device memory may return before the fill is done, and
        # it runs on the null stream, which the library's non-blocking
        # streams do not wait for: finish it before the buffer is handed out
        synchronize()

    @classmethod
    def from_array(cls, a, dtype=np.float64):
        a = np.ascontiguousarray(a, dtype=dtype)
        b = cls(a.size, dtype)
        b.upload(a)
        return b

    @property
    def addr(self):
        return self.ptr.value

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.n
        check(rt().hipMemcpy(self.ptr, a.ctypes.data_as(C.c_void_p), self.n * self.dtype.itemsize, 1),
              "hipMemcpy H2D")
        synchronize()                          # complete before work on other streams reads it

    def download(self, shape=None, count=None):
        """the first `count` elements (default all), optionally reshaped"""
        n = self.n if count is None else int(count)
        assert 0 <= n <= self.n
        a = np.empty(n, dtype=self.dtype)
        check(rt().hipMemcpy(a.ctypes.data_as(C.c_void_p), self.ptr, n * self.dtype.itemsize, 2),
              "hipMemcpy D2H")
        return a.reshape(shape) if shape is not None else a

    def download_range(self, start, count):
        """elements start .. start + count - 1"""
        start, count = int(start), int(count)
        assert 0 <= start and start + count <= self.n
        a = np.empty(count, dtype=self.dtype)
        src = C.c_void_p(self.ptr.value + start * self.dtype.itemsize)
        check(rt().hipMemcpy(a.ctypes.data_as(C.c_void_p), src, count * self.dtype.itemsize, 2), "hipMemcpy D2H")
        return a

    def copy_rows_from(self, dst_row, src, src_row, row_len, nrows, src_stride=1):
        """rows dst_row .. dst_row + nrows - 1 (row_len elements each) of this
        buffer <- rows src_row, src_row + src_stride, ... of buffer src (a
        device-to-device copy, finished before return)"""
        isz = self.dtype.itemsize
        assert src.dtype == self.dtype and nrows >= 0 and src_stride >= 1
        assert (dst_row + nrows) * row_len <= self.n
        assert nrows == 0 or (src_row + (nrows - 1) * src_stride + 1) * row_len <= src.n
        if nrows == 0:
            return
        check(rt().hipMemcpy2D(C.c_void_p(self.ptr.value + dst_row * row_len * isz), row_len * isz,
                               C.c_void_p(src.ptr.value + src_row * row_len * isz), src_stride * row_len * isz,
                               row_len * isz, nrows, 3), "hipMemcpy2D D2D")
        synchronize()

    def free(self):
        if self.ptr:
            rt().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    """A HIP stream at the default priority."""

    def __init__(self):
        self.h = C.c_void_p()
        check(rt().hipStreamCreateWithFlags(C.byref(self.h), 1), "hipStreamCreate")

    @property
    def handle(self):
        return self.h.value

    def synchronize(self):
        check(rt().hipStreamSynchronize(self.h), "hipStreamSynchronize")

    def destroy(self):
        if self.h:
            rt().hipStreamDestroy(self.h)
            self.h = C.c_void_p()


class Event:
    def __init__(self):
        self.h = C.c_void_p()
        check(rt().hipEventCreate(C.byref(self.h)), "hipEventCreate")

    def record(self, stream=None):
        check(rt().hipEventRecord(self.h, stream.h if stream else None), "hipEventRecord")

    def synchronize(self):
        check(rt().hipEventSynchronize(self.h), "hipEventSynchronize")

    def elapsed_ms(self, end):
        ms = C.c_float()
        check(rt().hipEventElapsedTime(C.byref(ms), self.h, end.h), "hipEventElapsedTime")
        return float(ms.value)

    def destroy(self):
        if self.h:
            rt().hipEventDestroy(self.h)
            self.h = C.c_void_p()
