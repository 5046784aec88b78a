"""Minimal ctypes access to the HIP runtime that libvkzg.so itself links
(/opt/rocm/lib/libamdhip64.so.7): device buffers, streams and events.

torch's ROCm wheel bundles a second HIP/HSA runtime; loading both into one
process leaves the second unable to open the GPU, so code that calls
libvkzg.so with device pointers uses this module instead of torch.
"""
import ctypes

from ._lib import load as _load_vkzg

_hip = None

H2D, D2H, D2D = 1, 2, 3


def lib():
    global _hip
    if _hip is None:
        _load_vkzg()  # makes sure the same runtime instance is the one resolved
        H = ctypes.CDLL("libamdhip64.so.7")
        vp = ctypes.c_void_p
        H.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
        H.hipFree.argtypes = [vp]
        H.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
        H.hipMemcpyAsync.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int, vp]
        H.hipMemset.argtypes = [vp, ctypes.c_int, ctypes.c_size_t]
        H.hipStreamCreate.argtypes = [ctypes.POINTER(vp)]
        H.hipStreamDestroy.argtypes = [vp]
        H.hipStreamSynchronize.argtypes = [vp]
        H.hipEventCreate.argtypes = [ctypes.POINTER(vp)]
        H.hipEventDestroy.argtypes = [vp]
        H.hipEventRecord.argtypes = [vp, vp]
        H.hipEventSynchronize.argtypes = [vp]
        H.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), vp, vp]
        H.hipGetErrorString.restype = ctypes.c_char_p
        H.hipGetErrorString.argtypes = [ctypes.c_int]
        H.hipGetDeviceCount.argtypes = [ctypes.POINTER(ctypes.c_int)]
        _hip = H
    return _hip


def check(rc):
    if rc != 0:
        raise RuntimeError("HIP error %d: %s" % (rc, lib().hipGetErrorString(rc).decode()))


def device_count():
    n = ctypes.c_int(0)
    if lib().hipGetDeviceCount(ctypes.byref(n)) != 0:
        return 0
    return n.value


class DeviceBuffer:
    def __init__(self, nbytes):
        self.nbytes = nbytes
        self.ptr = ctypes.c_void_p()
        check(lib().hipMalloc(ctypes.byref(self.ptr), max(1, nbytes)))

    @classmethod
    def from_bytes(cls, data):
        b = cls(len(data))
        src = ctypes.create_string_buffer(bytes(data), max(1, len(data)))
        check(lib().hipMemcpy(b.ptr, src, len(data), H2D))
        return b

    def to_bytes(self, nbytes=None):
        n = self.nbytes if nbytes is None else nbytes
        dst = ctypes.create_string_buffer(max(1, n))
        check(lib().hipMemcpy(dst, self.ptr, n, D2H))
        return dst.raw[:n]

    def zero(self):
        check(lib().hipMemset(self.ptr, 0, self.nbytes))

    def free(self):
        if self.ptr:
            lib().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        self.handle = ctypes.c_void_p()
        check(lib().hipStreamCreate(ctypes.byref(self.handle)))

    def synchronize(self):
        check(lib().hipStreamSynchronize(self.handle))

    def destroy(self):
        if self.handle:
            lib().hipStreamDestroy(self.handle)
            self.handle = ctypes.c_void_p()


class Event:
    def __init__(self):
        self.handle = ctypes.c_void_p()
        check(lib().hipEventCreate(ctypes.byref(self.handle)))

    def record(self, stream=None):
        check(lib().hipEventRecord(self.handle, stream.handle if stream else None))

    def synchronize(self):
        check(lib().hipEventSynchronize(self.handle))

    def elapsed_ms(self, later):
        ms = ctypes.c_float()
        check(lib().hipEventElapsedTime(ctypes.byref(ms), self.handle, later.handle))
        return ms.value
