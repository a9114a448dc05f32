"""The tracer column remaps of trcupd.F90: pcmtrc (808-909) and plmtrc
with plmtrcx/plmtrcs (911-1147), as initrc (trcupd.F90:167-199) and the
tracer relaxation of thermf (thermf.F90:770-806) call them per sea column.

Pinning: the oracle's column (oracle_trc_column) against trcupd.F90 itself,
built as it lies by oracle/ref into oracle/_ref/libtrcupd_ref.so, bitwise,
on columns that reach every branch: layers inside and across outputs,
thin layers (max(., 1e-6)), outputs below the bottom, and outputs whose
scan runs to the last layer and leaves the start lf where it was (the next
output then scans those layers again, each adding 1e-6/(zb-zt) of its
value).  The library's column code (hyb_trc.h compiled for the host) is
checked against the oracle on gridded domains; the GPU against the oracle
in test_gpu_trc.py.  Where the reference stops ('bad lf': new interfaces
out of order) the reference cannot be run; the oracle and the library
agree there on NaN outputs and an error."""
import ctypes

import numpy as np
import pytest

import hybgen_abi as abi
import hostcheck_lib as hc
import oracle_lib as ol

D = ctypes.POINTER(ctypes.c_double)
I = ctypes.POINTER(ctypes.c_int)
ONEM = abi.ONEM


def _interfaces(rng, n, depth, thin_frac=0.2):
    d = rng.exponential(1.0, n) * (rng.random(n) > thin_frac)
    d[rng.random(n) < 0.05] = 3e-7  # thinner than trcupd's thin = 1e-6
    d[0] = max(d[0], 0.1)  # the first output must start above the bottom
    p = np.concatenate([[0.0], np.cumsum(d)])
    return p * (depth / max(p[-1], 1e-300))


def trc_column(rng, ki, ko, ks):
    depth = rng.uniform(5.0, 4000.0) * ONEM
    pi = _interfaces(rng, ki, depth)
    po = _interfaces(rng, ko, depth * rng.choice([0.6, 1.0, 1.0, 1.4]))
    if po[1] - po[0] < 1e-6:
        po[1:] += 1e-3  # a thin first output makes the reference read so(0)
    si = rng.standard_normal((ks, ki)).cumsum(axis=1) + 5.0
    return si, pi, po


def oracle_col(method, si, pi, po):
    ks, ki = si.shape
    ko = len(po) - 1
    so = np.zeros((ks, ko))
    rc = ol.oracle().oracle_trc_column(method, ki, ks, np.ascontiguousarray(si).ctypes.data_as(D),
                                       pi.ctypes.data_as(D), po.ctypes.data_as(D), ko,
                                       so.ctypes.data_as(D))
    return rc, so


def ref_col(method, si, pi, po):
    ks, ki = si.shape
    ko = len(po) - 1
    so = np.zeros((ks, ko))
    fn = ol.trc_ref().plmtrc_ if method else ol.trc_ref().pcmtrc_
    a = [ctypes.c_int(x) for x in (ki, ks, ko)]
    fn(np.ascontiguousarray(si).ctypes.data_as(D), np.ascontiguousarray(pi).ctypes.data_as(D),
       ctypes.byref(a[0]), ctypes.byref(a[1]), so.ctypes.data_as(D),
       np.ascontiguousarray(po).ctypes.data_as(D), ctypes.byref(a[2]))
    return so


needs_ref = pytest.mark.skipif(not ol.trc_ref_available(), reason="trcupd.F90 build absent")


@needs_ref
@pytest.mark.parametrize("method", [0, 1])
@pytest.mark.parametrize("ki,ko,ks", [(2, 4, 1), (5, 41, 3), (22, 22, 2), (41, 32, 4),
                                      (64, 17, 2), (33, 64, 5)])
def test_oracle_column_matches_reference(method, ki, ko, ks):
    rng = np.random.default_rng(1000 * method + 31 * ki + ko)
    rescans = 0
    for _ in range(150):
        si, pi, po = trc_column(rng, ki, ko, ks)
        rc, got = oracle_col(method, si, pi, po)
        assert rc == 0
        assert np.array_equal(got, ref_col(method, si, pi, po))
        # outputs whose scan ran past the last layer without stopping early
        rescans += int(np.sum((po[1:] >= pi[ki - 1]) & (po[:-1] < pi[ki])))
    assert rescans > 0


def _grid(rng, method, ki, ko, ks, idm=41, jdm=19, nbdy=3):
    NI, NJ = idm + 2 * nbdy, jdm + 2 * nbdy
    P = abi.HybTrcParams(idm=idm, jdm=jdm, ii=idm, jj=jdm, nbdy=nbdy, ki=ki, ko=ko, ks=ks,
                         method=method, reserved=0)
    a = {"si": np.zeros((NI, NJ, ki, ks), order="F"), "pi": np.zeros((NI, NJ, ki + 1), order="F"),
         "po": np.zeros((NI, NJ, ko + 1), order="F"),
         "mask": np.asfortranarray((rng.random((NI, NJ)) < 0.7).astype(np.int32)),
         "so": np.asfortranarray(np.full((NI, NJ, ko, ks), -3.0))}
    for I in range(NI):
        for J in range(NJ):
            si, pi, po = trc_column(rng, ki, ko, ks)
            a["si"][I, J] = si.T
            a["pi"][I, J] = pi
            a["po"][I, J] = po
    return P, a


def _fields(a):
    F = abi.HybTrcFields()
    for k, v in a.items():
        setattr(F, k, v.ctypes.data)
    return F


@pytest.mark.parametrize("method", [0, 1])
@pytest.mark.parametrize("ki,ko,ks", [(2, 9, 1), (41, 41, 4), (17, 30, 6), (64, 12, 2)])
def test_library_column_code_matches_oracle(method, ki, ko, ks):
    rng = np.random.default_rng(7 * ki + ko + method)
    P, a = _grid(rng, method, ki, ko, ks)
    b = {k: v.copy(order="F") for k, v in a.items()}
    assert ol.oracle().oracle_trc_remap(ctypes.byref(P), ctypes.byref(_fields(a))) == 0
    assert hc.lib().hostcheck_trc(ctypes.byref(P), ctypes.byref(_fields(b))) == 0
    assert np.array_equal(a["so"], b["so"])
    land = a["mask"] == 0
    assert np.all(b["so"][land] == -3.0)


def test_out_of_order_interfaces_are_reported():
    """New interfaces out of order would stop the reference ('bad lf');
    the oracle and the library both give NaN from there and an error."""
    rng = np.random.default_rng(3)
    P, a = _grid(rng, 1, 12, 12, 2, idm=8, jdm=5)
    a["mask"][...] = 0
    a["mask"][4, 4] = 1
    po = a["po"][4, 4]
    po[6] = po[3] - 1.0  # out of order: a later output starts above lf
    po[7:] = np.maximum(po[7:], po[5])
    b = {k: v.copy(order="F") for k, v in a.items()}
    assert ol.oracle().oracle_trc_remap(ctypes.byref(P), ctypes.byref(_fields(a))) == abi.HYB_EINVAL
    assert hc.lib().hostcheck_trc(ctypes.byref(P), ctypes.byref(_fields(b))) == abi.HYB_EINVAL
    assert np.array_equal(a["so"], b["so"], equal_nan=True)
    assert np.isnan(a["so"][4, 4]).any()


def test_abi_rejects_bad_arguments():
    lib = abi.load()
    rng = np.random.default_rng(0)
    P, a = _grid(rng, 0, 5, 5, 1, idm=6, jdm=4)
    F = _fields(a)
    for field, value in (("ki", 1), ("ki", 65), ("ko", 0), ("ks", 0), ("method", 2),
                         ("reserved", 1)):
        Q = abi.HybTrcParams.from_buffer_copy(P)
        setattr(Q, field, value)
        assert lib.hycom_trc_remap(ctypes.byref(Q), ctypes.byref(F)) == abi.HYB_EINVAL, field
