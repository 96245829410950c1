"""Minimal ctypes binding of the HIP runtime: device buffers, copies, streams and events.

Plumbing for the Python API, tests and benchmark; the compute lives in libpentago_amd.so.
"""
import ctypes as C

_hip = None


def hip():
    global _hip
    if _hip is None:
        L = C.CDLL("libamdhip64.so")
        P = C.POINTER
        sig = {
            "hipGetDeviceCount": [P(C.c_int)],
            "hipSetDevice": [C.c_int],
            "hipMalloc": [P(C.c_void_p), C.c_size_t],
            "hipFree": [C.c_void_p],
            "hipMemcpy": [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int],
            "hipMemcpyAsync": [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p],
            "hipMemset": [C.c_void_p, C.c_int, C.c_size_t],
            "hipMemsetAsync": [C.c_void_p, C.c_int, C.c_size_t, C.c_void_p],
            "hipDeviceSynchronize": [],
            "hipStreamCreate": [P(C.c_void_p)],
            "hipStreamDestroy": [C.c_void_p],
            "hipStreamSynchronize": [C.c_void_p],
            "hipEventCreate": [P(C.c_void_p)],
            "hipEventDestroy": [C.c_void_p],
            "hipEventRecord": [C.c_void_p, C.c_void_p],
            "hipEventSynchronize": [C.c_void_p],
            "hipEventElapsedTime": [P(C.c_float), C.c_void_p, C.c_void_p],
            "hipMemGetInfo": [P(C.c_size_t), P(C.c_size_t)],
            "hipStreamWaitEvent": [C.c_void_p, C.c_void_p, C.c_uint],
            "hipLaunchHostFunc": [C.c_void_p, C.c_void_p, C.c_void_p],
            "hipGetErrorString": [C.c_int],
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = C.c_char_p if name == "hipGetErrorString" else C.c_int
        _hip = L
    return _hip


H2D, D2H, D2D = 1, 2, 3


def check(code, what):
    if code != 0:
        raise RuntimeError("%s failed: %s" % (what, hip().hipGetErrorString(code).decode()))


def device_count():
    n = C.c_int(0)
    try:
        if hip().hipGetDeviceCount(C.byref(n)) != 0:
            return 0
    except OSError:
        return 0
    return n.value


class Buffer:
    """A device allocation freed when the object is collected."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        check(hip().hipMalloc(C.byref(p), max(self.nbytes, 1)), "hipMalloc(%d)" % self.nbytes)
        self.ptr = p.value

    def __del__(self):
        if getattr(self, "ptr", None):
            hip().hipFree(C.c_void_p(self.ptr))
            self.ptr = None

    def upload(self, array, offset=0):
        check(hip().hipMemcpy(C.c_void_p(self.ptr + offset), C.c_void_p(array.ctypes.data), array.nbytes, H2D),
              "hipMemcpy H2D")

    def download(self, array, offset=0):
        check(hip().hipMemcpy(C.c_void_p(array.ctypes.data), C.c_void_p(self.ptr + offset), array.nbytes, D2H),
              "hipMemcpy D2H")
        return array

    def zero(self):
        check(hip().hipMemset(C.c_void_p(self.ptr), 0, self.nbytes), "hipMemset")


def copy_to_device(ptr, array):
    """Synchronous copy of a contiguous host array to device address ptr."""
    check(hip().hipMemcpy(C.c_void_p(ptr), C.c_void_p(array.ctypes.data), array.nbytes, H2D), "hipMemcpy H2D")


def synchronize():
    check(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


class Stream:
    def __init__(self):
        s = C.c_void_p()
        check(hip().hipStreamCreate(C.byref(s)), "hipStreamCreate")
        self.ptr = s.value

    def synchronize(self):
        check(hip().hipStreamSynchronize(C.c_void_p(self.ptr)), "hipStreamSynchronize")

    def __del__(self):
        if getattr(self, "ptr", None):
            hip().hipStreamDestroy(C.c_void_p(self.ptr))
            self.ptr = None


class Event:
    def __init__(self):
        e = C.c_void_p()
        check(hip().hipEventCreate(C.byref(e)), "hipEventCreate")
        self.ptr = e.value

    def record(self, stream=None):
        check(hip().hipEventRecord(C.c_void_p(self.ptr), C.c_void_p(stream.ptr if stream else None)), "hipEventRecord")

    def synchronize(self):
        check(hip().hipEventSynchronize(C.c_void_p(self.ptr)), "hipEventSynchronize")

    def elapsed_ms(self, end):
        t = C.c_float()
        check(hip().hipEventElapsedTime(C.byref(t), C.c_void_p(self.ptr), C.c_void_p(end.ptr)), "hipEventElapsedTime")
        return t.value

    def __del__(self):
        if getattr(self, "ptr", None):
            hip().hipEventDestroy(C.c_void_p(self.ptr))
            self.ptr = None


def mem_info():
    free, total = C.c_size_t(), C.c_size_t()
    check(hip().hipMemGetInfo(C.byref(free), C.byref(total)), "hipMemGetInfo")
    return free.value, total.value
