"""ctypes loader for the oracle's C restatement (oracle/_build/liboracle.so)."""

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_lp = ctypes.POINTER(ctypes.c_int64)
_i64 = ctypes.c_int64


def build():
    """Compile lean_geometry.c with gcc (via the Makefile next to it)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = ctypes.CDLL(_SO)
        L.oc_area_perimeter.argtypes = [_dp, _ip, _i64, _dp, _dp, _dp, _dp, ctypes.c_double, _dp, _i64,
                                        _dp, _dp, _dp, _lp]
        L.oc_area_perimeter.restype = ctypes.c_int
        L.oc_solve_radius.argtypes = [_dp, _ip, _i64, _dp, ctypes.c_double, _dp, _i64, ctypes.c_double,
                                      ctypes.c_int, _dp, _lp]
        L.oc_solve_radius.restype = ctypes.c_int
        L.oc_per_source_geometry.argtypes = [_dp, _i64, _i64, _ip, _i64, _dp, _dp, _dp, _dp, ctypes.c_double,
                                             _dp, _i64, _dp, _i64, ctypes.c_double, ctypes.c_int,
                                             _dp, _dp, _dp, _dp, _lp]
        L.oc_per_source_geometry.restype = ctypes.c_int64
        L.oc_fused_normalized_divergence.argtypes = [_dp, _i64, _i64, _ip, _i64, _dp, _dp, _dp, _dp, _dp]
        L.oc_fused_normalized_divergence.restype = None
        _lib = L
    return _lib


def dptr(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def iptr(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


def lptr(a):
    assert a.dtype == np.int64 and a.flags.c_contiguous
    return a.ctypes.data_as(_lp)
