"""The drop-in Fortran shim (hycom-src_amd/fortran/hybgen_hip.F90), built
against HYCOM-src's own modules (oracle/ref/Makefile target `shim`).

CPU: the shim packs the model's module state into exactly the parameter
struct the C ABI expects and hands over pointers to the right module
arrays (the oracle run through those pointers equals the oracle run on the
caller's arrays).  GPU: the reference harness calls `hybgen(m,n,raflag)`,
which resolves to the shim, which runs the HIP library in copy mode."""
import ctypes
import os

import numpy as np
import pytest

import hybgen_abi as abi
import oracle_lib as ol

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SHIM_SO = os.path.join(ROOT, "oracle", "_ref", "libhybgen_shim.so")
_P = ctypes.POINTER
pytestmark = pytest.mark.skipif(not os.path.exists(SHIM_SO),
                                reason="shim check library not built (make -C oracle/ref shim)")
_lib = None


def shim():
    global _lib
    if _lib is None:
        abi.load()  # the shim library links the HIP library by path
        h = ctypes.CDLL(SHIM_SO)
        h.shimchk_pack.argtypes = [_P(abi.HybParams), _P(abi.HybFields), _P(abi.HybParams),
                                   _P(abi.HybFields)]
        h.shimchk_store.argtypes = [_P(abi.HybParams), _P(abi.HybFields)]
        h.href_hybgen.argtypes = [_P(abi.HybParams), _P(abi.HybFields), ctypes.c_int,
                                  ctypes.c_int, ctypes.c_int, _P(ctypes.c_double)]
        _lib = h
    return _lib


def _domain(**over):
    P, w = abi.workload_params("cfg1", idm=40, jdm=28, **over)
    return abi.HostDomain(P).synth(abi.make_synth(P, w, seed=17))


@pytest.mark.parametrize("over", [
    dict(hybmap=2), dict(hybmap=3, ntracr=4, trcflg=[2, 0, 0, 2], nreg=1),
    dict(hybmap=0, kdm=22, ntracr=0, trcflg=[], hybiso=0.0),
])
def test_shim_packs_parameters_and_fields(over):
    d = _domain(**over)
    P = d.P
    if over.get("ntracr", 2) >= 3:
        P.itracr801, P.itracr803 = 1, 3
    F = d.fields()
    P2, F2 = abi.HybParams(), abi.HybFields()
    assert shim().shimchk_pack(ctypes.byref(P), ctypes.byref(F), ctypes.byref(P2),
                               ctypes.byref(F2)) == 0
    assert bytes(P2) == bytes(P), "packed parameters differ from the caller's"
    # the oracle through the shim's pointers (the model's module arrays)
    want = d.copy()
    ol.run_oracle(want, threads=4)
    rc = ol.oracle().oracle_hybgen(ctypes.byref(P2), ctypes.byref(F2), 1, 2, 0)
    assert rc == 0
    got = d.copy()
    assert shim().shimchk_store(ctypes.byref(P), ctypes.byref(got.fields())) == 0
    df = ol.diff_domains(want, got)
    assert not df, f"fields reached through the shim differ: {df}"


@pytest.mark.gpu
@pytest.mark.parametrize("raflag", [0, 1])
def test_model_call_through_shim_on_gpu(gpu, raflag):
    """hybgen(m,n,raflag) as the model calls it (the reference harness, the
    shim in place of hybgen.F90), against the oracle."""
    print(f"\n[case] shim end-to-end raflag={raflag}", flush=True)
    d = _domain(hybmap=2, ntracr=3, trcflg=[2, 0, 0])
    if raflag:
        d.P.itracr801, d.P.itracr802 = 2, 3
        rng = np.random.default_rng(2)
        d.a["dpo"][...] = np.abs(d.a["dp"] * (1.0 + 0.01 * rng.standard_normal(d.a["dp"].shape)))
    want = d.copy()
    ol.run_oracle(want, raflag=raflag, threads=4)
    got = d.copy()
    secs = ctypes.c_double()
    rc = shim().href_hybgen(ctypes.byref(got.P), ctypes.byref(got.fields()), 1, 2, raflag,
                            ctypes.byref(secs))
    assert rc == 0
    df = ol.diff_domains(want, got)
    assert not df, f"model call through the shim differs from the oracle: {df}"
