"""HYCOM .a/.b files (hycom_za_read*, hycom_restart_read; README.src.mod_za:
46-60, mod_za_sm.h:678-1256, machi_c.c:50-67, mod_restart.F90:14-651).

Pinning: .a records written by the reference's own zaiowr3 (oracle/_ref,
href_za) are read back by the library; records the test writes are read
by the reference's zaiord3 and by the library; both give the same values
and the same per-record ranges, bitwise.  The restart reader is checked on
a pair written in restart_out's layout (header lines, one .b line per .a
record in the 4100 format of mod_restart.F90:1250)."""
import ctypes
import os

import numpy as np
import pytest

import hybgen_abi as abi
import oracle_lib as ol

D = ctypes.POINTER(ctypes.c_double)
I = ctypes.POINTER(ctypes.c_int32)
VOID = np.float32(2.0 ** 100)
IDM, JDM, NB = 37, 23, 6  # one extent for every reference call: zaiost sizes its buffer once


def write_a(path, recs):
    """records (idm, jdm) as big-endian float32, each padded with 2^100 to
    a multiple of 4096 words (README.src.mod_za:46-60)"""
    idm, jdm = recs[0].shape
    nw = ((idm * jdm + 4095) // 4096) * 4096
    with open(path, "wb") as f:
        for r in recs:
            w = np.full(nw, VOID, dtype=">f4")
            w[:idm * jdm] = np.asarray(r, dtype=np.float32).ravel(order="F")
            f.write(w.tobytes())


def halo(idm, jdm, nrec, fill=-9.0):
    return np.asfortranarray(np.full((idm + 2 * NB, jdm + 2 * NB, nrec), fill))


def lib_read(path, nrec, mask=None, rec0=1):
    out = halo(IDM, JDM, nrec)
    mn, mx = np.zeros(nrec), np.zeros(nrec)
    rc = abi.load().hycom_za_read(path.encode(), IDM, JDM, NB, rec0, nrec,
                                  None if mask is None else mask.ctypes.data_as(I),
                                  out.ctypes.data_as(D), mn.ctypes.data_as(D), mx.ctypes.data_as(D))
    return rc, out, mn, mx


def ref_za(mode, path, h):
    nrec = h.shape[2]
    mn, mx = np.zeros(nrec), np.zeros(nrec)
    p = path.encode()
    rc = ol.ref().href_za(mode, p, len(p), IDM, JDM, nrec, h.ctypes.data_as(D),
                          mn.ctypes.data_as(D), mx.ctypes.data_as(D))
    assert rc == 0
    return mn, mx


needs_ref = pytest.mark.skipif(not ol.ref_available(), reason="reference build absent")


def _records(rng, nrec):
    h = halo(IDM, JDM, nrec)
    h[NB:NB + IDM, NB:NB + JDM, :] = rng.standard_normal((IDM, JDM, nrec)) * 10.0 ** rng.integers(
        -3, 6, nrec)
    h[NB + 3, NB + 4, 0] = float(VOID)  # a data void
    return h


@needs_ref
def test_reference_written_records_read_back(tmp_path):
    rng = np.random.default_rng(1)
    h = _records(rng, 5)
    path = str(tmp_path / "ref.a")
    ref_mn, ref_mx = ref_za(1, path, h.copy(order="F"))
    rc, got, mn, mx = lib_read(path, 5)
    assert rc == 0
    want = h[NB:NB + IDM, NB:NB + JDM, :].astype(np.float32).astype(np.float64)
    assert np.array_equal(got[NB:NB + IDM, NB:NB + JDM, :], want)
    assert np.all(got[:NB] == -9.0) and np.all(got[:, :NB] == -9.0)  # halo untouched
    assert np.array_equal(mn, ref_mn) and np.array_equal(mx, ref_mx)
    assert mx[0] < 1e29  # the void is left out of the range


@needs_ref
def test_written_records_read_by_reference_and_library(tmp_path):
    rng = np.random.default_rng(2)
    recs = [rng.standard_normal((IDM, JDM)).astype(np.float32) for _ in range(4)]
    recs[2][5, 6] = VOID
    path = str(tmp_path / "t.a")
    write_a(path, recs)
    h = halo(IDM, JDM, 4)
    ref_mn, ref_mx = ref_za(0, path, h)
    rc, got, mn, mx = lib_read(path, 4)
    assert rc == 0
    assert np.array_equal(got[NB:NB + IDM, NB:NB + JDM], h[NB:NB + IDM, NB:NB + JDM])
    assert np.array_equal(mn, ref_mn) and np.array_equal(mx, ref_mx)
    rc, got2, _, _ = lib_read(path, 2, rec0=3)  # records 3..4 alone
    assert rc == 0 and np.array_equal(got2, got[:, :, 2:4])


def test_masked_read(tmp_path):
    rng = np.random.default_rng(3)
    recs = [rng.standard_normal((IDM, JDM)).astype(np.float32)]
    mask = np.asfortranarray((rng.random((IDM + 2 * NB, JDM + 2 * NB)) < 0.6).astype(np.int32))
    path = str(tmp_path / "m.a")
    write_a(path, recs)
    rc, got, mn, mx = lib_read(path, 1, mask=mask)
    assert rc == 0
    inner = mask[NB:NB + IDM, NB:NB + JDM] != 0
    g = got[NB:NB + IDM, NB:NB + JDM, 0]
    assert np.array_equal(g[inner], recs[0][inner].astype(np.float64))
    assert np.all(g[~inner] == -9.0)
    assert mn[0] == recs[0][inner].min() and mx[0] == recs[0][inner].max()
    recs[0][np.argwhere(inner)[0][0], np.argwhere(inner)[0][1]] = VOID
    write_a(path, recs)
    rc, _, _, _ = lib_read(path, 1, mask=mask)
    assert rc == abi.HYB_EINVAL  # a void inside the mask (mod_za_sm.h:842-851)


def write_restart(base, P, fields, names, nstep=120, dtime=365.25, old=False):
    """a restart pair in restart_out's layout: 2 header lines, then one .b
    line per .a record in the 4100 format (mod_restart.F90:752-760, 1250)"""
    recs, lines = [], []
    for name, arr, nl in zip(names, fields, [a.shape[2] for a in fields]):
        for r in range(nl):
            x = arr[NB:NB + P.idm, NB:NB + P.jdm, r].astype(np.float32)
            recs.append(x)
            ok = x != VOID
            k, t = (r % P.kdm) + 1, (r // P.kdm) + 1
            lines.append("%-8s: layer,tlevel,range = %3d%3d  %16.7E%16.7E" %
                         (name, k, t, x[ok].min(), x[ok].max()))
    write_a(base + ".a", recs)
    with open(base + ".b", "w") as f:
        kind = "RESTART: " if old else "RESTART2:"
        f.write(f"{kind} iexpt,iversn,yrflag,sigver =     10    23     3     2\n")
        f.write(f"{kind} nstep,dtime,thbase =  {nstep}  {dtime}  25.0\n")
        for ln in lines:
            f.write(ln + "\n")


def test_restart_read(tmp_path):
    P, _ = abi.workload_params("cfg1", idm=IDM, jdm=JDM, kdm=9, ntracr=2, nsigma=8)
    K = P.kdm
    rng = np.random.default_rng(4)
    NI, NJ = IDM + 2 * NB, JDM + 2 * NB

    def f(n, lo, hi):
        return np.asfortranarray(rng.uniform(lo, hi, (NI, NJ, n)))

    src = {"u": f(2 * K, -1, 1), "v": f(2 * K, -1, 1), "dp": f(2 * K, 0, 1e5),
           "temp": f(2 * K, -2, 30), "saln": f(2 * K, 30, 38), "ubavg": f(3, -1, 1),
           "vbavg": f(3, -1, 1), "pbavg": f(3, 0, 1), "pbot": f(1, 1e6, 5e7),
           "tracer": f(2 * K * 2, 0, 5)}
    names = ["u", "v", "dp", "temp", "saln", "ubavg", "vbavg", "pbavg", "pbot", "tracer"]
    base = str(tmp_path / "restart_in")
    write_restart(base, P, [src[n] for n in names], names)
    d = abi.HostDomain(P)
    d.a["ip"][NB:NB + IDM, NB:NB + JDM] = (rng.random((IDM, JDM)) < 0.7).astype(np.int32)
    F = d.fields()
    ns, dt = ctypes.c_int32(), ctypes.c_double()
    lib = abi.load()
    rc = lib.hycom_restart_read((base + ".a").encode(), (base + ".b").encode(), ctypes.byref(P),
                                ctypes.byref(F), ctypes.byref(ns), ctypes.byref(dt))
    assert rc == 0, lib.hycom_hybgen_last_error()
    assert ns.value == 120 and dt.value == 365.25
    inner = (slice(NB, NB + IDM), slice(NB, NB + JDM))
    for n in ("u", "v", "dp", "temp", "saln", "ubavg", "vbavg"):
        got = d.a[n].reshape(NI, NJ, -1, order="F")[inner]
        want = src[n][inner].astype(np.float32).astype(np.float64)
        assert np.array_equal(got, want), n
    got = d.a["tracer"].reshape(NI, NJ, -1, order="F")[inner]
    assert np.array_equal(got, src["tracer"][inner].astype(np.float32).astype(np.float64))
    # th3d as restart_in forms it for a RESTART2 file (mod_restart.F90:83-97)
    sig = np.vectorize(ol.oracle().oracle_sig)
    t = d.a["temp"].reshape(NI, NJ, -1, order="F")[inner]
    s = d.a["saln"].reshape(NI, NJ, -1, order="F")[inner]
    th = d.a["th3d"].reshape(NI, NJ, -1, order="F")[inner]
    ip = d.a["ip"][inner][:, :, None] == 1
    want = np.where(ip, sig(t, s) - P.thbase, 0.0)
    assert np.array_equal(th, want)
    # a .b range that does not match its record is refused (restart_in3d 604-616)
    with open(base + ".b") as fh:
        lines = fh.readlines()
    lines[5] = lines[5][:lines[5].index("=") + 1] + "   4  1   9.0000000E+03   9.5000000E+03\n"
    with open(base + ".b", "w") as fh:
        fh.writelines(lines)
    rc = lib.hycom_restart_read((base + ".a").encode(), (base + ".b").encode(), ctypes.byref(P),
                                ctypes.byref(F), None, None)
    assert rc == abi.HYB_EINVAL
    assert b"not consistent" in lib.hycom_hybgen_last_error()
