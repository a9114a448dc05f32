"""Loaders for the checkers: the C restatement (oracle/liboracle_hybgen.so)
and the flang build of the reference hybgen (oracle/_ref/*.so).  Test
infrastructure only."""
import ctypes
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "hycom-src_amd"))
import hybgen_abi as abi  # noqa: E402

ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle_hybgen.so")
REF_DIR = os.path.join(ROOT, "oracle", "_ref")

_P = ctypes.POINTER
_oracle = None
_refs = {}


def oracle():
    global _oracle
    if _oracle is None:
        lib = ctypes.CDLL(ORACLE_SO)
        lib.oracle_hybgen.argtypes = [_P(abi.HybParams), _P(abi.HybFields),
                                      ctypes.c_int, ctypes.c_int, ctypes.c_int]
        lib.oracle_hybgen_omp.argtypes = [_P(abi.HybParams), _P(abi.HybFields),
                                          ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int]
        lib.oracle_hybgen_decisions.argtypes = [_P(abi.HybParams), _P(abi.HybFields),
                                                ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        lib.oracle_sig.restype = ctypes.c_double
        lib.oracle_sig.argtypes = [ctypes.c_double, ctypes.c_double]
        lib.oracle_remap_column.argtypes = [
            ctypes.c_int, ctypes.c_int, _P(ctypes.c_double), _P(ctypes.c_double),
            _P(ctypes.c_double), _P(ctypes.c_int), ctypes.c_double, _P(ctypes.c_double)]
        lib.oracle_stokes_vertical.argtypes = [_P(abi.HybStokesParams),
                                               _P(abi.HybStokesFields), ctypes.c_int]
        lib.oracle_plm_dpi_column.argtypes = [
            ctypes.c_int, ctypes.c_int, ctypes.c_int, _P(ctypes.c_double),
            _P(ctypes.c_double), _P(ctypes.c_double), ctypes.c_double, _P(ctypes.c_double)]
        lib.oracle_plm_dpi_column.restype = None
        lib.oracle_trc_column.argtypes = [
            ctypes.c_int, ctypes.c_int, ctypes.c_int, _P(ctypes.c_double), _P(ctypes.c_double),
            _P(ctypes.c_double), ctypes.c_int, _P(ctypes.c_double)]
        lib.oracle_trc_remap.argtypes = [_P(abi.HybTrcParams), _P(abi.HybTrcFields)]
        _oracle = lib
    return _oracle


def ref_available(denfix=False):
    return os.path.exists(ref_path(denfix))


def ref_path(denfix=False):
    return os.path.join(REF_DIR, "libhybgen_ref_denfix.so" if denfix else "libhybgen_ref.so")


def ref(denfix=False):
    if denfix not in _refs:
        # lazy binding: dpthuv's calls of latbdt/latbdtf (lbflag 2 and 4,
        # never taken here) stay unresolved
        lib = ctypes.CDLL(ref_path(denfix), mode=os.RTLD_LAZY)
        lib.href_hybgen.restype = ctypes.c_int
        lib.href_hybgen.argtypes = [_P(abi.HybParams), _P(abi.HybFields), ctypes.c_int,
                                    ctypes.c_int, ctypes.c_int, _P(ctypes.c_double)]
        lib.href_sig.restype = ctypes.c_double
        lib.href_sig.argtypes = [ctypes.c_double, ctypes.c_double]
        lib.href_consts.argtypes = [_P(ctypes.c_double)]
        lib.href_remap.argtypes = [
            ctypes.c_int, ctypes.c_int, _P(ctypes.c_double), _P(ctypes.c_double),
            _P(ctypes.c_double), _P(ctypes.c_int), ctypes.c_double, _P(ctypes.c_double)]
        lib.href_plm_dpi_column.argtypes = [
            ctypes.c_int, ctypes.c_int, ctypes.c_int, _P(ctypes.c_double),
            _P(ctypes.c_double), _P(ctypes.c_double), ctypes.c_double, _P(ctypes.c_double)]
        lib.href_plm_dpi_column.restype = None
        lib.href_za.restype = ctypes.c_int
        lib.href_za.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.c_int, ctypes.c_int,
                                ctypes.c_int, ctypes.c_int, _P(ctypes.c_double),
                                _P(ctypes.c_double), _P(ctypes.c_double)]
        _refs[denfix] = lib
    return _refs[denfix]


def run_oracle(dom, m=1, n=2, raflag=0, threads=1):
    F = dom.fields()
    rc = oracle().oracle_hybgen_omp(ctypes.byref(dom.P), ctypes.byref(F), m, n, raflag,
                                    threads)
    if rc != 0:
        raise RuntimeError(f"oracle_hybgen rc={rc}")
    return dom


def run_oracle_decisions(dom, m=1, n=2, raflag=0, threads=1):
    """run_oracle that also returns the per-column decisions of hybgenaj:
    (fixlay, kp) as (idm, jdm) int32 arrays, 0 on land."""
    import numpy as np
    P = dom.P
    fixlay = np.zeros((P.idm, P.jdm), dtype=np.int32, order="F")
    kp = np.zeros((P.idm, P.jdm), dtype=np.int32, order="F")
    F = dom.fields()
    rc = oracle().oracle_hybgen_decisions(ctypes.byref(P), ctypes.byref(F), m, n, raflag,
                                          threads, fixlay.ctypes.data, kp.ctypes.data)
    if rc != 0:
        raise RuntimeError(f"oracle_hybgen_decisions rc={rc}")
    return fixlay, kp


def run_ref(dom, m=1, n=2, raflag=0):
    F = dom.fields()
    secs = ctypes.c_double(0.0)
    rc = ref(bool(dom.P.ldenfix)).href_hybgen(ctypes.byref(dom.P), ctypes.byref(F), m, n,
                                              raflag, ctypes.byref(secs))
    if rc != 0:
        raise RuntimeError(f"href_hybgen rc={rc}")
    return secs.value


def diff_domains(a, b, names=abi.OUTPUTS):
    """Names of arrays that differ bitwise (NaN-aware), with counts."""
    import numpy as np
    out = {}
    for n in names:
        x, y = a.a[n], b.a[n]
        if x.dtype.kind == "f":
            bad = ~((x == y) | (np.isnan(x) & np.isnan(y)))
        else:
            bad = x != y
        c = int(bad.sum())
        if c:
            out[n] = c
    return out


def run_oracle_stokes(sd):
    F = sd.fields()
    rc = oracle().oracle_stokes_vertical(ctypes.byref(sd.P), ctypes.byref(F), sd.n)
    if rc != 0:
        raise RuntimeError(f"oracle_stokes_vertical rc={rc}")
    return sd


def start_ref(arctic=False):
    """href_start of the reference build (the ARCTIC build of mod_xc for a
    global grid), opened with lazy binding as ref() is."""
    key = "start-arctic" if arctic else "start"
    if key not in _refs:
        path = os.path.join(REF_DIR, "libhybgen_ref_arctic.so") if arctic else ref_path(False)
        lib = ctypes.CDLL(path, mode=os.RTLD_LAZY) if arctic else ref(False)
        lib.href_start.restype = ctypes.c_int
        lib.href_start.argtypes = [_P(abi.HybParams)] + [ctypes.c_void_p] * 6 + [
            _P(ctypes.c_int32)] + [ctypes.c_void_p] * 5
        _refs[key] = lib
    return _refs[key]


def start_ref_available(arctic=False):
    return os.path.exists(os.path.join(REF_DIR, "libhybgen_ref_arctic.so" if arctic
                                       else "libhybgen_ref.so"))


TRC_REF_SO = os.path.join(REF_DIR, "libtrcupd_ref.so")
_trc_ref = None


def trc_ref_available():
    return os.path.exists(TRC_REF_SO)


def trc_ref():
    """trcupd.F90 as built by oracle/ref: opened with lazy binding, since
    routines of it that the tests never call use routines of other files
    (pcmtrc_ and plmtrc_ need none of them)."""
    global _trc_ref
    if _trc_ref is None:
        lib = ctypes.CDLL(TRC_REF_SO, mode=os.RTLD_LAZY)
        for fn in (lib.pcmtrc_, lib.plmtrc_):
            fn.restype = None
            fn.argtypes = [_P(ctypes.c_double), _P(ctypes.c_double), _P(ctypes.c_int),
                           _P(ctypes.c_int), _P(ctypes.c_double), _P(ctypes.c_double),
                           _P(ctypes.c_int)]
        _trc_ref = lib
    return _trc_ref
