"""Loader for tests/hostcheck/libhostcheck.so: the device column code of
the library (hycom-src_amd/csrc/hyb_*.h) compiled for the host, run in the
kernels' phase order.  Lets the CPU suite check the kernels' own arithmetic
against the oracle.  Test infrastructure only."""
import ctypes
import os

import hybgen_abi as abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SO = os.path.join(ROOT, "tests", "hostcheck", "libhostcheck.so")
FAMILIES = {"lane-arrays": "hostcheck_hybgen",
            "streaming": "hostcheck_hybgen_stream",
            "streaming-recompute": "hostcheck_hybgen_stream_rc"}
_lib = None


def lib():
    global _lib
    if _lib is None:
        h = ctypes.CDLL(SO)
        for fn in FAMILIES.values():
            getattr(h, fn).argtypes = [ctypes.POINTER(abi.HybParams),
                                       ctypes.POINTER(abi.HybFields),
                                       ctypes.c_int, ctypes.c_int, ctypes.c_int]
        h.hostcheck_stream_unsafe_columns.restype = ctypes.c_int64
        h.hostcheck_stokes.argtypes = [ctypes.POINTER(abi.HybStokesParams),
                                       ctypes.POINTER(abi.HybStokesFields), ctypes.c_int]
        h.hostcheck_trc.argtypes = [ctypes.POINTER(abi.HybTrcParams),
                                    ctypes.POINTER(abi.HybTrcFields)]
        h.hostcheck_stream_uv_fallback_points.restype = ctypes.c_int64
        _lib = h
    return _lib


def covers(family, P):
    ok = P.nhybrd == P.kdm
    if family in ("streaming", "streaming-recompute"):
        return ok
    return True


def run(family, dom, m=1, n=2, raflag=0):
    F = dom.fields()
    rc = getattr(lib(), FAMILIES[family])(ctypes.byref(dom.P), ctypes.byref(F), m, n, raflag)
    if rc != 0:
        raise RuntimeError(f"{FAMILIES[family]} rc={rc}")
    return dom


def unsafe_columns():
    return lib().hostcheck_stream_unsafe_columns()


def uv_fallback_points():
    return lib().hostcheck_stream_uv_fallback_points()
