"""ctypes binding of the CPU baseline restatement (oracle/cpu_fast.c).

Timed by bench.py's cpu_baseline leg and checked against the oracle by
tests/test_cpu_fast.py; the product library never loads it.  The AVX-512
build (8 cells per batch) is used when the host supports it, else the AVX2
build (4 cells per batch).
"""
import ctypes
import os

import numpy as np

import pyoracle as po

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def _has_avx512():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    fl = line.split()
                    return all(x in fl for x in ("avx512f", "avx512dq", "avx512vl"))
    except OSError:
        pass
    return False


def lib():
    global _LIB
    if _LIB is None:
        name = "libcpufast_avx512.so" if _has_avx512() else "libcpufast_avx2.so"
        path = os.path.join(_HERE, name)
        if not os.path.exists(path):
            raise RuntimeError("CPU baseline not built: run `make -C oracle`")
        L = ctypes.CDLL(path)
        P = ctypes.POINTER(po.OrcParams)
        dp = ctypes.POINTER(ctypes.c_double)
        L.cfb_create.argtypes = [P]
        L.cfb_create.restype = ctypes.c_void_p
        L.cfb_destroy.argtypes = [ctypes.c_void_p]
        L.cfb_fill.argtypes = [ctypes.c_void_p, dp]
        L.cfb_vmult.argtypes = [ctypes.c_void_p, dp, dp]
        L.cfb_simd_width.restype = ctypes.c_int
        L.cfb_threads.restype = ctypes.c_int
        _LIB = L
    return _LIB


class TangentVmult:
    """the QP-cache tangent vmult of a 3D Q1 scalar-mobility operator"""

    def __init__(self, p):
        self.p = p
        self.h = lib().cfb_create(ctypes.byref(p))
        if not self.h:
            raise ValueError("CPU baseline: 3D, Q1, scalar mobility, 3..16 components only")

    def __del__(self):
        if getattr(self, "h", None):
            lib().cfb_destroy(self.h)
            self.h = None

    def fill(self, lin):
        lin = np.ascontiguousarray(lin, dtype=np.float64).ravel()
        lib().cfb_fill(self.h, po._dp(lin))

    def vmult(self, src, dst=None):
        src = np.ascontiguousarray(src, dtype=np.float64).ravel()
        if dst is None:
            dst = np.empty_like(src)
        lib().cfb_vmult(self.h, po._dp(src), po._dp(dst))
        return dst


def simd_width():
    return int(lib().cfb_simd_width())


def threads():
    return int(lib().cfb_threads())
