"""Device memory for tests that hand device addresses to libvpmx: hipMalloc / hipMemcpy of the
HIP runtime that libvpmx.so itself links (ROCm's libamdhip64 under /opt/rocm), so the
addresses belong to the library's runtime (torch ships a runtime of its own)."""
import ctypes as C
import os

HOST_TO_DEVICE, DEVICE_TO_HOST = 1, 2
_rt = None


def runtime():
    global _rt
    if _rt is None:
        for path in ("/opt/rocm/lib/libamdhip64.so.7", "/opt/rocm/lib/libamdhip64.so"):
            if os.path.exists(path):
                _rt = C.CDLL(path)
                break
        else:
            raise RuntimeError("ROCm's libamdhip64 not found")
    return _rt


def malloc(nbytes):
    p = C.c_void_p()
    r = runtime().hipMalloc(C.byref(p), C.c_size_t(int(nbytes)))
    if r != 0:
        raise RuntimeError("hipMalloc failed: %d" % r)
    return p.value


def free(p):
    runtime().hipFree(C.c_void_p(p))


def to_device(arr):
    p = malloc(arr.nbytes)
    r = runtime().hipMemcpy(C.c_void_p(p), arr.ctypes.data_as(C.c_void_p), C.c_size_t(arr.nbytes), HOST_TO_DEVICE)
    if r != 0:
        raise RuntimeError("hipMemcpy failed: %d" % r)
    return p


def to_host(arr, p):
    r = runtime().hipMemcpy(arr.ctypes.data_as(C.c_void_p), C.c_void_p(p), C.c_size_t(arr.nbytes), DEVICE_TO_HOST)
    if r != 0:
        raise RuntimeError("hipMemcpy failed: %d" % r)


def synchronize():
    r = runtime().hipDeviceSynchronize()
    if r != 0:
        raise RuntimeError("hipDeviceSynchronize failed: %d" % r)
