"""The mixed-precision form (HYB_PREC_MIXED): float reconstruction, FP64
interfaces, piece integrals and sums, each layer's polynomial keeping its
mean exactly.  Checked against the FP64 oracle: every remapped field
within 1e-5 of its scale, interfaces and thicknesses bitwise, column
integrals conserved to 1e-12 where no output is clamped or thin.  Host
build of the device code (tests/hostcheck)."""
import ctypes

import numpy as np
import pytest

import hybgen_abi as abi
import hostcheck_lib as hc
import oracle_lib as ol
from test_oracle_invariants import THIN, remap

D = ctypes.POINTER(ctypes.c_double)
I = ctypes.POINTER(ctypes.c_int)
TOL = 1e-5
ONEM = 9.806 / 0.001


def walk(mixed, method, si, pi, po):
    kk = len(si)
    so = np.zeros(kk)
    lc = np.zeros(kk, np.int32)
    si, pi, po = (np.ascontiguousarray(x, np.float64) for x in (si, pi, po))
    hc.lib().hostcheck_remap_column(mixed, method, kk, si.ctypes.data_as(D), pi.ctypes.data_as(D),
                                    po.ctypes.data_as(D), lc.ctypes.data_as(I),
                                    so.ctypes.data_as(D))
    return so


def _columns(rng, n, extreme):
    for _ in range(n):
        kk = int(rng.integers(4, 41))
        dp = rng.exponential(50.0 * ONEM, kk) * (rng.random(kk) > 0.15)
        pi = np.concatenate([[0.0], np.cumsum(dp)])
        if extreme:  # new layers of 1e-3 m next to thick ones
            dq = np.where(rng.random(kk) < 0.5, 1e-3 * ONEM, rng.exponential(80.0 * ONEM, kk))
        else:
            dq = rng.exponential(50.0 * ONEM, kk) + 1.0
        po = np.concatenate([[0.0], np.cumsum(dq)])
        po *= pi[-1] / po[-1]
        po[-1] = pi[-1]
        base = 40.0 if extreme else 10.0
        si = base + np.cumsum(rng.standard_normal(kk)) * (5.0 if extreme else 1.0)
        yield si, pi, po


@pytest.mark.parametrize("extreme", [0, 1])
@pytest.mark.parametrize("method", [0, 1, 2, 3])
def test_column_error_and_conservation(method, extreme):
    rng = np.random.default_rng(31 + method + 7 * extreme)
    worst, cons, ncons = 0.0, 0.0, 0
    for si, pi, po in _columns(rng, 300, extreme):
        ref = remap(method, si, pi, po)
        assert np.array_equal(walk(0, method, si, pi, po), ref)  # FP64 walk is bitwise
        got = walk(1, method, si, pi, po)
        worst = max(worst, np.max(np.abs(got - ref)) / np.max(np.abs(si)))
        dpo = np.diff(po)
        if not (((got == si.min()) | (got == si.max())).any() or (dpo <= THIN).any()):
            ncons += 1
            cons = max(cons, abs(np.sum(got * dpo) - np.sum(si * np.diff(pi)))
                       / np.sum(np.abs(si) * np.diff(pi)))
    assert worst <= TOL, f"max error {worst:.2e}"
    assert ncons > 20 and cons <= 1e-12, f"conservation {cons:.2e} over {ncons} columns"


@pytest.mark.parametrize("extreme", [0, 1])
@pytest.mark.parametrize("hybmap", [0, 1, 2, 3])
def test_domain_against_fp64_oracle(hybmap, extreme):
    P, w = abi.workload_params("cfg1", hybmap=hybmap, idm=70, jdm=44, ntracr=3,
                               trcflg=[2, 0, 0])
    d = abi.HostDomain(P).synth(abi.make_synth(P, w, seed=5, extreme=extreme))
    want = d.copy()
    ol.run_oracle(want, threads=4)
    P.precision = abi.PREC_MIXED
    got = d.copy()
    got.P = P
    hc.run("streaming", got)
    for f in ("dp", "p", "dpu", "dpv", "pu", "pv"):
        assert np.array_equal(want.a[f], got.a[f]), f"{f}: the regrid must stay FP64"
    for f in ("temp", "saln", "th3d", "tracer", "u", "v"):
        x, y = want.a[f], got.a[f]
        err = np.max(np.abs(x - y)) / np.max(np.abs(x))
        assert err <= TOL, f"{f}: error {err:.2e} of the field's scale"


def _colint(h, f, tr=None):
    P = h.P
    nb = P.nbdy
    x = h.a[f][..., 1] if tr is None else h.a[f][..., 1, tr]
    return (h.a["dp"][..., 1] * x).sum(axis=2)[nb:nb + P.ii, nb:nb + P.jj]


@pytest.mark.parametrize("hybmap", [1, 2, 3])
def test_domain_column_integrals_conserved(hybmap):
    """Column integrals sum_k dp*x of T, S and the tracers on a GLBb window,
    where the deep salinity is nearly uniform and float edge values often
    coincide: the mixed form changes them by at most 1e-12 wherever the
    FP64 oracle changes them by at most 1e-13."""
    w0 = abi.WORKLOADS["glbb"]
    P, w = abi.workload_params("glbb", hybmap=hybmap, idm=80, jdm=60)
    S = abi.make_synth(P, w, seed=200, itdm=w0["idm"], jtdm=w0["jdm"], i0=2000, j0=1600)
    d = abi.HostDomain(P).synth(S)
    want = d.copy()
    ol.run_oracle(want, threads=4)
    Pm, _ = abi.workload_params("glbb", hybmap=hybmap, idm=80, jdm=60)
    Pm.precision = abi.PREC_MIXED
    got = abi.HostDomain(Pm, {k: v.copy(order="F") for k, v in d.a.items()})
    hc.run("streaming", got)
    nb = P.nbdy
    sea = d.a["ip"][nb:nb + P.ii, nb:nb + P.jj] > 0
    fields = [("temp", None), ("saln", None)] + [("tracer", q) for q in range(P.ntracr)]
    checked = 0
    for f, tr in fields:
        i0 = _colint(d, f, tr)
        den = np.maximum(np.abs(i0), 1e-300)
        r64 = np.abs(_colint(want, f, tr) - i0) / den
        rmx = np.abs(_colint(got, f, tr) - i0) / den
        ok = sea & (r64 <= 1e-13)
        checked += int(ok.sum())
        worst = float(rmx[ok].max()) if ok.any() else 0.0
        assert worst <= 1e-12, f"{f}{'' if tr is None else tr + 1}: residual {worst:.2e}"
    assert checked > 1000
