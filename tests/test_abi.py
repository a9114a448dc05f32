"""The C ABI (include/hycom_hybgen.h) on the CPU: the library loads, exports
every declared entry point, validates parameters like blkdat.F90, and fails
loudly without a gfx950 device (there is no CPU fallback)."""
import ctypes
import os
import re

import pytest

import hybgen_abi as abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hycom_hybgen.h")


def declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(hycom_hybgen\w*)\s*\(", txt)))


def test_header_declares_the_entry_points():
    names = declared()
    for need in ("hycom_hybgen", "hycom_hybgen_device", "hycom_hybgen_check_params",
                 "hycom_hybgen_derive_vgrid", "hycom_hybgen_strerror"):
        assert need in names


def test_library_exports_every_declared_symbol():
    lib = abi.load()
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"


def test_abi_version_and_strerror():
    lib = abi.load()
    txt = open(HEADER).read()
    ver = int(re.search(r"#define HYB_ABI_VERSION (\d+)", txt).group(1))
    assert lib.hycom_hybgen_abi_version() == ver
    lib.hycom_hybgen_strerror.restype = ctypes.c_char_p
    for code in (0, -1, -2, -3, -4, -5, -6, -7, -8):
        assert lib.hycom_hybgen_strerror(code)


def _params(**over):
    P, _ = abi.workload_params("cfg1", **over)
    return P


def _rc(P):
    return abi.load().hycom_hybgen_check_params(ctypes.byref(P))


def test_check_params_accepts_the_workloads():
    assert _rc(_params()) == 0
    P, _ = abi.workload_params("glbb")
    assert _rc(P) == 0


@pytest.mark.parametrize("field,value,code", [
    ("hybflg", 1, -1),        # blkdat.F90:1007-1011: 17-term EOS needs hybflg 0
    ("hybthk", 0.95, -1),     # blkdat.F90:966-974
    ("hybthn", 0.5, -1),      # blkdat.F90:956-964
    ("hybmap", 4, -1),
    ("mxlmy", 1, -2),         # MY2.5 q2/q2l not built
    ("ntracr", 65, -2),
    ("kdm", 3, -2),
    ("precision", 7, -1),
])
def test_check_params_rejects(field, value, code):
    P = _params()
    setattr(P, field, value)
    assert _rc(P) == code


def test_check_params_rejects_shared_diag_tracer():
    P = _params(ntracr=3, trcflg=[0, 0, 0])
    P.itracr801, P.itracr802 = 2, 2
    assert _rc(P) == -2


def test_check_params_rejects_lcm_carry_configuration():
    # !isopcm with nhybrd<kdm carries lcm across the columns of a row in the
    # reference (hybgen.F90:386-394 outside the i loop); not supported
    P = _params(nhybrd=30, hybiso=0.0)
    assert _rc(P) == -2


def test_device_entry_fails_without_gpu():
    lib = abi.load()
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is visible")
    except Exception:
        pass
    P = _params(idm=16, jdm=12)
    d = abi.HostDomain(P).synth(abi.make_synth(P, abi.WORKLOADS["cfg1"], seed=1))
    F = d.fields()
    before = d.a["temp"].copy()
    rc = lib.hycom_hybgen(ctypes.byref(P), ctypes.byref(F), 1, 2, 0)
    assert rc == -5  # HYB_ENODEVICE
    assert (d.a["temp"] == before).all()  # nothing computed on the CPU


def test_noalloc_entry_sizes_and_no_gpu():
    """The allocation-free entry point: its workspace size, the refusal of
    domains beyond the fixed array (checked on the host before any device
    work) and no CPU fallback."""
    lib = abi.load()
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is visible")
    except Exception:
        pass
    cap = lib.hycom_hybgen_noalloc_capacity()
    assert cap == 256 << 20
    P = _params(idm=16, jdm=12)
    NIJ = (16 + 12) * (12 + 12)
    r = lambda b: (b + 255) // 256 * 256  # noqa: E731
    want = r(8 * NIJ * (P.kdm + 1)) + r(4 * NIJ) + r(8 * NIJ) + r(8 * NIJ)
    assert lib.hycom_hybgen_noalloc_need(ctypes.byref(P), 0) == want
    assert lib.hycom_hybgen_noalloc_need(ctypes.byref(P), 1) == want + r(24 * NIJ * P.kdm)
    d = abi.HostDomain(P)
    F = d.fields()
    assert lib.hycom_hybgen_device_noalloc(ctypes.byref(P), ctypes.byref(F), 1, 2, 0,
                                           None) == abi.HYB_ENODEVICE
    # a domain whose workspace exceeds the array: refused by its shape alone
    big = _params(idm=1200, jdm=900, kdm=41)
    assert lib.hycom_hybgen_noalloc_need(ctypes.byref(big), 0) > cap
    assert lib.hycom_hybgen_device_noalloc(ctypes.byref(big), ctypes.byref(F), 1, 2, 0,
                                           None) == -1  # HYB_EINVAL
    assert b"fixed array" in lib.hycom_hybgen_last_error()
    P.check_finite = 1
    assert lib.hycom_hybgen_device_noalloc(ctypes.byref(P), ctypes.byref(F), 1, 2, 0,
                                           None) == -2  # HYB_EUNSUPPORTED


def _host_batch(specs):
    """hyb_domain array over host-side fields (only validated on the CPU);
    the caller keeps the returned domains alive while D is used."""
    doms = []
    for over in specs:
        P = _params(idm=16, jdm=12, **over)
        doms.append(abi.HostDomain(P))
    Fs = [d.fields() for d in doms]
    D = (abi.HybDomain * len(doms))()
    for q, (d, F) in enumerate(zip(doms, Fs)):
        D[q].P = ctypes.addressof(d.P)
        D[q].F = ctypes.addressof(F)
    return D, doms, Fs


def test_batch_rejects_bad_batches():
    lib = abi.load()
    D, keep, _ = _host_batch([{}])
    assert lib.hycom_hybgen_batch(ctypes.addressof(D), 0, 1, 2, None) == abi.HYB_EINVAL
    assert lib.hycom_hybgen_batch(None, 1, 1, 2, None) == abi.HYB_EINVAL
    assert lib.hycom_hybgen_batch(ctypes.addressof(D), 1, 3, 2, None) == abi.HYB_EINVAL
    # nhybrd < kdm has no streaming kernels: refused, not run another way
    D, keep, _ = _host_batch([{}, dict(nhybrd=30)])
    assert _rc(keep[1].P) == 0
    rc = lib.hycom_hybgen_batch(ctypes.addressof(D), 2, 1, 2, None)
    assert rc == abi.HYB_EUNSUPPORTED
    assert b"domain 1" in lib.hycom_hybgen_last_error()
    # two domains over the same p array
    D, doms, Fs = _host_batch([{}, {}])
    D[1].F = ctypes.addressof(Fs[0])
    assert lib.hycom_hybgen_batch(ctypes.addressof(D), 2, 1, 2, None) == abi.HYB_EINVAL


def test_batch_fails_without_gpu():
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is visible")
    except Exception:
        pass
    D, keep, _ = _host_batch([{}, dict(kdm=22)])
    assert abi.load().hycom_hybgen_batch(ctypes.addressof(D), 2, 1, 2, None) == abi.HYB_ENODEVICE


def test_server_config_checks_and_no_gpu():
    lib = abi.load()
    h = ctypes.c_void_p()
    for bad in (dict(max_active=-1), dict(max_active=257), dict(remap_cols=96),
                dict(regrid_cols=32), dict(hbm_budget=-5)):
        C = abi.HybServerConfig(**bad)
        assert lib.hycom_hybgen_server_create(ctypes.byref(C), ctypes.byref(h)) == abi.HYB_EINVAL
        assert not h.value
    try:
        import torch
        if torch.cuda.is_available():
            return
    except Exception:
        pass
    assert lib.hycom_hybgen_server_create(None, ctypes.byref(h)) == abi.HYB_ENODEVICE
    assert not h.value
    # NULL handles are refused, not dereferenced
    n = ctypes.c_int32()
    assert lib.hycom_hybgen_server_submit(None, 0, 1, 0, 0) == abi.HYB_ENULL
    assert lib.hycom_hybgen_server_poll(None, None, 0, 0, ctypes.byref(n)) == abi.HYB_ENULL
    assert lib.hycom_hybgen_server_destroy(None) == 0


def test_batch_tuning_table_without_a_batch():
    """The autotuner's candidate table is empty before any batch call in
    this process, and a negative capacity is refused."""
    lib = abi.load()
    if lib.hycom_hybgen_batch_tuning(None, None, None, None, 0) > 0:
        pytest.skip("a batch ran earlier in this process")
    assert lib.hycom_hybgen_batch_tuning(None, None, None, None, 8) == 0
    assert lib.hycom_hybgen_batch_tuning(None, None, None, None, -1) == abi.HYB_EINVAL
    lv, km = ctypes.c_int32(-5), ctypes.c_int32(-5)
    assert lib.hycom_hybgen_batch_staged_levels(ctypes.byref(lv), ctypes.byref(km)) == 0
    assert lv.value == 0 and km.value == 0
