"""Measurement probes of the benchmark (liblkprobe.so, built beside
liblkgpu.so; not part of the Kriging library): a plain device copy and the
FP64 MFMA issue-rate probe with the shader clock it ran at."""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblkprobe.so")
_lib = None
_d = ctypes.POINTER(ctypes.c_double)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m libkriging_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        L.lkprobe_last_error.restype = ctypes.c_char_p
        L.lkprobe_device_copy.argtypes = [ctypes.c_int, ctypes.c_size_t, ctypes.c_int, _d]
        L.lkprobe_mfma.argtypes = [ctypes.c_int] * 4 + [_d, _d, _d]
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise RuntimeError(lib().lkprobe_last_error().decode())


def device_copy(nbytes, reps=3, device=0):
    """mean ms of a plain device copy kernel over nbytes (read + write)"""
    ms = ctypes.c_double()
    _check(lib().lkprobe_device_copy(int(device), int(nbytes), int(reps), ctypes.byref(ms)))
    return ms.value


def mfma(blocks, iters, chains=8, device=0):
    """(ms, TFLOP/s, shader clock MHz) of the FP64 MFMA issue-rate probe"""
    ms, tf, mhz = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    _check(lib().lkprobe_mfma(int(device), int(blocks), int(iters), int(chains), ctypes.byref(ms), ctypes.byref(tf),
                              ctypes.byref(mhz)))
    return ms.value, tf.value, mhz.value
