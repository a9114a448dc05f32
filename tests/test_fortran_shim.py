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


def _dynamics(d, m, n, step):
    """A deterministic stand-in for one model step between hybgen calls,
    the same on both sides of the A/B: level n takes level m's state with
    the interfaces displaced (ordered, the bottom kept), T, S, tracers,
    u and v perturbed smoothly, th3d recomputed from T and S, and dpo set
    as the Robert-Asselin filter expects (dpo(n) = the new thickness,
    dpo(m) = the current one)."""
    P = d.P
    nb, K = P.nbdy, P.kdm
    rng = np.random.default_rng(1000 + step)
    a = d.a
    ip = a["ip"] != 0
    lm, ln = m - 1, n - 1
    dp = a["dp"][..., lm].copy()
    p = np.concatenate([np.zeros(dp.shape[:2] + (1,)), np.cumsum(dp, axis=2)], axis=2)
    # interior interfaces move by up to 30% of the thinner neighbouring layer
    lim = 0.3 * np.minimum(dp[..., :-1], dp[..., 1:])
    p[..., 1:-1] += lim * rng.uniform(-1.0, 1.0, lim.shape)
    p[..., 1:-1] = np.minimum(np.maximum.accumulate(p[..., 1:-1], axis=2), p[..., -1:])
    newdp = np.where(ip[..., None], np.diff(p, axis=2), dp)
    for name in ("temp", "saln", "u", "v"):
        x = a[name][..., lm]
        a[name][..., ln] = x + 0.01 * np.sin(0.3 * np.arange(x.shape[0])[:, None, None] + step)
    if P.ntracr:
        t = a["tracer"][..., lm, :]
        a["tracer"][..., ln, :] = np.maximum(t * (1.0 + 0.001 * step), 0.0)
    a["dp"][..., ln] = newdp
    a["dpu"][..., ln] = a["dpu"][..., lm]
    a["dpv"][..., ln] = a["dpv"][..., lm]
    sig = np.vectorize(lambda t, s: ol.ref(False).href_sig(t, s))
    th = sig(a["temp"][..., ln], a["saln"][..., ln]) - P.thbase
    a["th3d"][..., ln] = np.where(ip[..., None], th, a["th3d"][..., ln])
    a["dpo"][..., ln] = newdp
    a["dpo"][..., lm] = dp


@pytest.mark.gpu
def test_model_steps_through_shim_on_gpu(gpu):
    """An A/B over 8 model steps at the model's call site: one state steps
    with the reference hybgen (HYCOM-src's hybgen.F90, flang), the other
    with hybgen resolved to the shim (the HIP library in copy mode), each
    evolving on its own through the same stand-in dynamics, leapfrog time
    levels and the Robert-Asselin branch on, as mod_hycom.F90:2650 calls
    it; the two states are compared bit for bit after every call."""
    print("\n[case] 8-step A/B, reference hybgen vs shim", flush=True)
    d = _domain(hybmap=2, ntracr=3, trcflg=[2, 0, 0])
    ref_lib = ol.ref(False)
    ref_lib.href_sig.restype = ctypes.c_double
    ref_lib.href_sig.argtypes = [ctypes.c_double, ctypes.c_double]
    a, b = d.copy(), d.copy()
    secs = ctypes.c_double()
    for step in range(8):
        m, n = (1, 2) if step % 2 == 0 else (2, 1)
        _dynamics(a, m, n, step)
        _dynamics(b, m, n, step)
        assert not ol.diff_domains(a, b, list(abi.FIELD_NAMES))
        print(f"[step] {step} m={m} n={n}", flush=True)
        ol.run_ref(a, m=m, n=n, raflag=1)
        rc = shim().href_hybgen(ctypes.byref(b.P), ctypes.byref(b.fields()), m, n, 1,
                                ctypes.byref(secs))
        assert rc == 0
        df = ol.diff_domains(a, b)
        assert not df, f"step {step}: reference and shim differ: {df}"
