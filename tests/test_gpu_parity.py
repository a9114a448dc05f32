"""GPU parity: the HIP path through the C ABI against the oracle, bitwise.

Every case prints its name, shape and seed before touching the GPU."""
import ctypes

import numpy as np
import pytest

import hybgen_abi as abi
import oracle_lib as ol

pytestmark = pytest.mark.gpu


FAMILIES = {1: "lane-arrays", 3: "streaming"}


@pytest.fixture(params=[1, 3], ids=list(FAMILIES.values()))
def kernel_path(request, gpu):
    """Runs the test once per kernel family.  A family that does not cover
    the case's parameters (streaming: nhybrd==kdm) is
    replaced by the lane-array kernels."""
    yield request.param
    abi.check(abi.load().hycom_hybgen_set_kernel_path(0))


def _select(family, P):
    ok = P.nhybrd == P.kdm
    if family == 3 and not ok:
        family = 1
    print(f"[kernels] {FAMILIES[family]}", flush=True)
    abi.check(abi.load().hycom_hybgen_set_kernel_path(family))


def _case(name, **over):
    print(f"\n[case] {name} {over}", flush=True)
    P, w = abi.workload_params(over.pop("workload", "cfg1"), **over)
    return P, w


def _oracle_out(d0, m=1, n=2, raflag=0):
    a = d0.copy()
    ol.run_oracle(a, m=m, n=n, raflag=raflag, threads=8)
    return a


def _assert_same(a, b, what):
    df = ol.diff_domains(a, b)
    assert not df, f"{what}: arrays differ bitwise: {df}"


@pytest.mark.parametrize("hybmap", [0, 1, 2, 3])
def test_device_path_matches_oracle_cfg1(gpu, kernel_path, hybmap):
    P, w = _case("cfg1-device", hybmap=hybmap)
    _select(kernel_path, P)
    d0 = abi.HostDomain(P).synth(abi.make_synth(P, w, seed=100))
    want = _oracle_out(d0)
    dd = abi.DeviceDomain(P).upload(d0)
    dd.hybgen(m=1, n=2)
    import torch
    torch.cuda.synchronize()
    got = dd.download()
    _assert_same(want, got, f"hybmap={hybmap}")


@pytest.mark.parametrize("variant", [
    dict(ldenfix=1), dict(hybthk=0.5, hybthn=5.0), dict(nsigma=0, hybrlx=1.0),
    dict(nreg=1, dx00=30.0, dx00x=300.0, dx00f=1.1), dict(nreg=3, hybmap=3),
    dict(kdm=22, ntracr=0, trcflg=[]), dict(kdm=41, ntracr=4, trcflg=[2, 0, 0, 0]),
    dict(hybmap=1, hybiso=3.0), dict(kdm=41, hybmap=3, ntracr=1, trcflg=[0]),
    dict(kdm=22, hybmap=0, nreg=1), dict(kdm=41, hybmap=1, hybthk=0.9),
])
def test_device_path_variants(gpu, kernel_path, variant):
    P, w = _case("cfg1-variant", idm=97, jdm=61, **variant)
    _select(kernel_path, P)
    d0 = abi.HostDomain(P).synth(abi.make_synth(P, w, seed=7))
    want = _oracle_out(d0)
    dd = abi.DeviceDomain(P).upload(d0)
    dd.hybgen(m=1, n=2)
    got = dd.download()
    _assert_same(want, got, str(variant))


@pytest.mark.parametrize("mn", [(1, 2), (2, 1)])
def test_robert_asselin_and_diag_tracers(gpu, kernel_path, mn):
    m, n = mn
    P, w = _case("raflag", idm=80, jdm=50, ntracr=4, trcflg=[2, 0, 0, 0], hybmap=2)
    P.itracr801, P.itracr802, P.itracr803 = 2, 3, 4
    _select(kernel_path, P)
    d0 = abi.HostDomain(P).synth(abi.make_synth(P, w, seed=11), m=m, n=n)
    rng = np.random.default_rng(1)
    d0.a["dpo"][...] = np.abs(d0.a["dp"] * (1.0 + 0.01 * rng.standard_normal(d0.a["dp"].shape)))
    want = _oracle_out(d0, m=m, n=n, raflag=1)
    dd = abi.DeviceDomain(P).upload(d0)
    dd.hybgen(m=m, n=n, raflag=1)
    got = dd.download()
    _assert_same(want, got, f"raflag m={m} n={n}")


def test_copy_mode_matches_oracle(gpu):
    P, w = _case("cfg1-copy-mode", hybmap=2)
    d0 = abi.HostDomain(P).synth(abi.make_synth(P, w, seed=3))
    want = _oracle_out(d0)
    got = d0.copy().hybgen(m=1, n=2)
    _assert_same(want, got, "copy mode")


def _copy_bytes(P, raflag):
    """Bytes copy mode moves per call (see hycom_hybgen_copy_stats): levels
    n (and m, dpo with raflag) of the 3-d fields both ways, ubavg/vbavg of
    those levels up, p, pu, pv both ways, theta, depths, topiso, depthu,
    depthv and the three masks up."""
    NIJ = (P.idm + 2 * P.nbdy) * (P.jdm + 2 * P.nbdy)
    lev = 8 * NIJ * P.kdm
    nl = 2 if raflag else 1
    pp = 3 * 8 * NIJ * (P.kdm + 1)
    const = lev + 4 * 8 * NIJ + 3 * 4 * NIJ
    up = nl * ((8 + P.ntracr) * lev + 2 * 8 * NIJ) + (2 * lev if raflag else 0) + pp + const
    down = nl * (8 + P.ntracr) * lev + pp
    return up, down


def test_copy_mode_sequence_matches_oracle(gpu):
    """Successive copy-mode calls on one host domain, as a model makes
    them: time levels alternating, a Robert-Asselin call, a second domain
    of the same shape (new host arrays), and calls after the caller changed
    theta, depthu/depthv and the halo of p in place (the model sets
    depthu/depthv after inicon's hybgen calls, mod_hycom.F90:1495,1684).
    Every call against the oracle on the same sequence, and the bytes each
    call moved."""
    lib = abi.load()
    P, w = _case("copy-sequence", idm=90, jdm=57, kdm=41, hybmap=2, ntracr=3,
                 trcflg=[2, 0, 0])
    d0 = abi.HostDomain(P).synth(abi.make_synth(P, w, seed=31))
    rng = np.random.default_rng(31)
    d0.a["dpo"][...] = np.abs(d0.a["dp"] * (1.0 + 0.01 * rng.standard_normal(d0.a["dp"].shape)))
    # land p, pu, pv hold values the call must keep
    for name in ("p", "pu", "pv"):
        d0.a[name][...] = rng.standard_normal(d0.a[name].shape)
    want, got = d0.copy(), d0.copy()
    up = ctypes.c_int64()
    down = ctypes.c_int64()
    for q, (m, n, ra) in enumerate([(1, 2, 0), (2, 1, 0), (1, 2, 1), (2, 1, 0)]):
        print(f"[call] {q} m={m} n={n} raflag={ra}", flush=True)
        ol.run_oracle(want, m=m, n=n, raflag=ra, threads=8)
        got.hybgen(m=m, n=n, raflag=ra)
        _assert_same(want, got, f"copy mode call {q}")
        abi.check(lib.hycom_hybgen_copy_stats(ctypes.byref(up), ctypes.byref(down)))
        assert (up.value, down.value) == _copy_bytes(P, ra), (q, up.value, down.value)
    # a second domain of the same shape: its own land values of p, pu, pv
    d1 = abi.HostDomain(P).synth(abi.make_synth(P, w, seed=32))
    for name in ("p", "pu", "pv"):
        d1.a[name][...] = -rng.random(d1.a[name].shape)
    w1, g1 = d1.copy(), d1.copy()
    ol.run_oracle(w1, threads=8)
    g1.hybgen(m=1, n=2)
    _assert_same(w1, g1, "copy mode, second domain")
    # fields changed in place between calls, with nothing said to the library
    want.a["theta"][...] += 0.01
    got.a["theta"][...] += 0.01
    ol.run_oracle(want, m=1, n=2, threads=8)
    got.hybgen(m=1, n=2)
    _assert_same(want, got, "copy mode after theta changed")
    for name in ("depthu", "depthv"):
        want.a[name][...] *= 0.75
        got.a[name][...] *= 0.75
    nb = P.nbdy
    for x in (want, got):
        x.a["p"][:nb] = 123.0
        x.a["p"][:, -nb:] = -7.0
    ol.run_oracle(want, m=2, n=1, threads=8)
    got.hybgen(m=2, n=1)
    _assert_same(want, got, "copy mode after depthu/depthv and the halo of p changed")


def test_device_synth_matches_host_synth(gpu):
    P, w = _case("synth", idm=130, jdm=77, kdm=41, ntracr=4, trcflg=[2, 0, 0, 0])
    S = abi.make_synth(P, w, seed=21)
    h = abi.HostDomain(P).synth(S)
    dd = abi.DeviceDomain(P).synth(S)
    got = dd.download()
    names = [n for n in abi.FIELD_NAMES]
    df = ol.diff_domains(h, got, names)
    assert not df, f"generator differs host vs device: {df}"


@pytest.mark.parametrize("hybmap", [2, 3], ids=["ppm", "weno"])
def test_larger_domain_matches_oracle(gpu, kernel_path, hybmap):
    """A 600x400 window of the GLBb field (kdm 41, 4 tracers) with both
    remappers the GLBb workload names, PPM and WENO."""
    P, w = _case("mid-size", idm=600, jdm=400, kdm=41, ntracr=4, trcflg=[2, 0, 0, 0],
                 workload="glbb", hybmap=hybmap)
    _select(kernel_path, P)
    S = abi.make_synth(P, w, seed=200, itdm=4500, jtdm=3298, i0=1800, j0=1200)
    dd = abi.DeviceDomain(P).synth(S)
    d0 = dd.download()
    want = _oracle_out(d0)
    dd.hybgen(m=1, n=2)
    got = dd.download()
    _assert_same(want, got, f"600x400x41 hybmap={hybmap}")


@pytest.mark.parametrize("kdm,hybmap,ntracr,extreme", [
    (5, 2, 0, 0), (17, 3, 3, 0), (50, 2, 7, 1), (64, 1, 8, 0), (64, 0, 1, 1),
    (29, 2, 9, 1), (41, 3, 4, 1), (17, 2, 2, 1),
])
def test_streaming_any_kdm(gpu, kdm, hybmap, ntracr, extreme):
    """Streaming kernels on shapes the fixed-K kernels do not cover: odd and
    maximal kdm, tracer groups beyond one walk (ntracr > 4), the in-place
    fallback columns and the deferred clamp (extreme inputs)."""
    dx = dict(dx00=30.0, dx00f=1.1, dx00x=300.0) if extreme else {}
    P, w = _case("streaming", idm=90, jdm=57, kdm=kdm, hybmap=hybmap, nsigma=min(8, kdm),
                 ntracr=ntracr, trcflg=[2, 0, 0, 2, 0, 0, 0, 0, 0][:ntracr],
                 hybiso=0.01 if kdm % 2 else 0.0, **dx)
    _select(3, P)
    d0 = abi.HostDomain(P).synth(abi.make_synth(P, w, seed=kdm, extreme=extreme))
    want = _oracle_out(d0)
    dd = abi.DeviceDomain(P).upload(d0)
    dd.hybgen(m=1, n=2)
    got = dd.download()
    abi.check(abi.load().hycom_hybgen_set_kernel_path(0))
    _assert_same(want, got, f"kdm={kdm} hybmap={hybmap} ntracr={ntracr}")


@pytest.mark.parametrize("hybmap", [1, 2, 3])
def test_mixed_precision_on_gpu(gpu, hybmap):
    """HYB_PREC_MIXED on the GPU: interfaces bitwise the FP64 oracle's, every
    remapped field within 1e-5 of its scale, and within 1e-5 of the host
    build of the same mixed code.  The device uses the hardware reciprocals
    and the host IEEE quotients; a limiter comparison that lands on the
    other side of its threshold moves a value by more than the rounding
    alone, so the two builds are held to the stated tolerance, not to
    the ulp."""
    import hostcheck_lib as hc
    P, w = _case("mixed", idm=90, jdm=57, kdm=41, hybmap=hybmap, ntracr=3, trcflg=[2, 0, 0])
    d0 = abi.HostDomain(P).synth(abi.make_synth(P, w, seed=9, extreme=1))
    want = _oracle_out(d0)
    P.precision = abi.PREC_MIXED
    host = d0.copy()
    host.P = P
    hc.run("streaming", host)
    dd = abi.DeviceDomain(P).upload(d0)
    dd.hybgen(m=1, n=2)
    got = dd.download()
    for f in ("dp", "p", "dpu", "dpv", "pu", "pv"):
        assert np.array_equal(want.a[f], got.a[f]), f
    for f in ("temp", "saln", "th3d", "tracer", "u", "v"):
        x = want.a[f]
        assert np.max(np.abs(x - got.a[f])) / np.max(np.abs(x)) <= 1e-5, f
    for f in ("dp", "p", "dpu", "dpv", "pu", "pv"):
        assert np.array_equal(host.a[f], got.a[f]), f
    for f in ("temp", "saln", "th3d", "tracer", "u", "v"):
        x = host.a[f]
        err = np.max(np.abs(x - got.a[f])) / np.max(np.abs(x))
        assert err <= 1e-5, f"{f}: GPU mixed vs host build of the same code {err:.2e}"


@pytest.mark.parametrize("kdm,hybmap,extreme", [(41, 2, 0), (22, 1, 1), (64, 0, 1), (17, 3, 1)])
def test_streaming_uv_with_lds(gpu, kdm, hybmap, extreme):
    """The u/v walk that keeps its new thicknesses in LDS (tuning knob 2 =
    0; the default forms them again from p) gives the oracle's bits."""
    dx = dict(dx00=30.0, dx00f=1.1, dx00x=300.0) if extreme else {}
    P, w = _case("uv-recompute", idm=90, jdm=57, kdm=kdm, hybmap=hybmap, nsigma=min(8, kdm),
                 hybiso=3.0 if hybmap == 1 else 0.01, **dx)
    _select(3, P)
    lib = abi.load()
    d0 = abi.HostDomain(P).synth(abi.make_synth(P, w, seed=kdm + 5, extreme=extreme))
    want = _oracle_out(d0)
    dd = abi.DeviceDomain(P).upload(d0)
    abi.check(lib.hycom_hybgen_tune(2, 0))
    try:
        dd.hybgen(m=1, n=2)
        got = dd.download()
    finally:
        abi.check(lib.hycom_hybgen_tune(2, 1))
        abi.check(lib.hycom_hybgen_set_kernel_path(0))
    _assert_same(want, got, f"uv recompute kdm={kdm} hybmap={hybmap}")


def test_uv_prefetch_depths(gpu):
    """The u/v walk at prefetch depths 0, 2 and 4 (knob 1; 4 is the
    default): bitwise the oracle."""
    P, w = _case("cfg1-uv-prefetch", hybmap=3)
    d0 = abi.HostDomain(P).synth(abi.make_synth(P, w, seed=106))
    want = _oracle_out(d0)
    lib = abi.load()
    import torch
    try:
        for pd in (0, 2, 4):
            print(f"[uv prefetch] {pd}", flush=True)
            abi.check(lib.hycom_hybgen_tune(1, pd))
            dd = abi.DeviceDomain(P).upload(d0)
            dd.hybgen(m=1, n=2)
            torch.cuda.synchronize()
            _assert_same(want, dd.download(), f"uv prefetch {pd}")
    finally:
        abi.check(lib.hycom_hybgen_tune(1, 4))


@pytest.mark.parametrize("hybmap,raflag", [(2, 0), (3, 1), (1, 0), (0, 1)])
def test_concurrent_uv_phase(gpu, hybmap, raflag):
    """The u/v phase on the library's second stream beside the p-column
    remap (tuning knob 4) gives the oracle's bits, the RA branch included."""
    P, w = _case("concurrent-uv", idm=90, jdm=57, kdm=41, hybmap=hybmap, ntracr=3,
                 trcflg=[2, 0, 0])
    _select(3, P)
    lib = abi.load()
    d0 = abi.HostDomain(P).synth(abi.make_synth(P, w, seed=70 + hybmap))
    if raflag:
        rng = np.random.default_rng(hybmap)
        d0.a["dpo"][...] = np.abs(d0.a["dp"] * (1.0 + 0.01 * rng.standard_normal(d0.a["dp"].shape)))
    want = _oracle_out(d0, raflag=raflag)
    dd = abi.DeviceDomain(P).upload(d0)
    abi.check(lib.hycom_hybgen_tune(4, 1))
    try:
        dd.hybgen(m=1, n=2, raflag=raflag)
        got = dd.download()
    finally:
        abi.check(lib.hycom_hybgen_tune(4, 0))
        abi.check(lib.hycom_hybgen_set_kernel_path(0))
    _assert_same(want, got, f"concurrent u/v hybmap={hybmap} raflag={raflag}")


@pytest.mark.parametrize("kdm,hybmap,extreme", [(41, 2, 1), (22, 3, 0), (64, 1, 1)])
def test_regrid_without_read_ahead(gpu, kdm, hybmap, extreme):
    """The regrid variant that loads its window as it takes it in (tuning
    knob 3 = 0; the default loads it one iteration ahead) gives the
    oracle's bits."""
    dx = dict(dx00=30.0, dx00f=1.1, dx00x=300.0) if extreme else {}
    P, w = _case("regrid-ahead", idm=90, jdm=57, kdm=kdm, hybmap=hybmap, nsigma=min(8, kdm),
                 **dx)
    _select(3, P)
    lib = abi.load()
    d0 = abi.HostDomain(P).synth(abi.make_synth(P, w, seed=kdm + 11, extreme=extreme))
    want = _oracle_out(d0)
    dd = abi.DeviceDomain(P).upload(d0)
    abi.check(lib.hycom_hybgen_tune(3, 0))
    try:
        dd.hybgen(m=1, n=2)
        got = dd.download()
    finally:
        abi.check(lib.hycom_hybgen_tune(3, 1))
        abi.check(lib.hycom_hybgen_set_kernel_path(0))
    _assert_same(want, got, f"regrid without read-ahead kdm={kdm} hybmap={hybmap}")


def test_finite_scan(gpu):
    """check_finite: a NaN or Inf in an input the call reads stops the call
    with HYB_ENONFINITE naming the first bad point, before anything is
    written; without it the call runs; a clean domain gives the same bits
    with and without the scan."""
    import torch
    P, w = _case("finite-scan", idm=60, jdm=40, kdm=32, ntracr=2, trcflg=[0, 0])
    d0 = abi.HostDomain(P).synth(abi.make_synth(P, w, seed=3))
    want = _oracle_out(d0)
    P.check_finite = 1
    dd = abi.DeviceDomain(P).upload(d0)
    dd.hybgen(m=1, n=2)
    torch.cuda.synchronize()
    _assert_same(want, dd.download(), "clean domain with the scan")
    nb = P.nbdy
    ip = d0.a["ip"]
    sea = np.argwhere(ip[nb:nb + P.ii, nb:nb + P.jj] > 0)
    i0, j0 = (int(x) for x in sea[len(sea) // 2])
    for field, extra in (("temp", ()), ("tracer", (1,))):
        bad = d0.copy()
        idx = (i0 + nb, j0 + nb, 5, 1) + extra
        bad.a[field][idx] = np.nan if field == "temp" else np.inf
        dd = abi.DeviceDomain(P).upload(bad)
        with pytest.raises(abi.HybError) as e:
            dd.hybgen(m=1, n=2)
        assert e.value.code == abi.HYB_ENONFINITE
        assert f"i={i0 + 1} j={j0 + 1}" in str(e.value), str(e.value)
        torch.cuda.synchronize()
        assert not ol.diff_domains(bad, dd.download()), "the scan must stop the call"


def _colint(h, f, tr=None):
    P = h.P
    nb = P.nbdy
    x = h.a[f][..., 1] if tr is None else h.a[f][..., 1, tr]
    return (h.a["dp"][..., 1] * x).sum(axis=2)[nb:nb + P.ii, nb:nb + P.jj]


@pytest.mark.parametrize("hybmap", [1, 2, 3])
def test_mixed_on_the_cfg1_shape_with_extreme_inputs(gpu, hybmap):
    """HYB_PREC_MIXED on the exact configs[3] shape, 150x100x32 with 2
    tracers, extreme inputs (1e-3 m layers, 40 C / 40 psu fronts), against
    the FP64 oracle: per field max and RMS error at most 1e-5 of the field's
    scale, interfaces bitwise, and at every sea column the column integrals
    sum_k dp*x of T, S and each tracer within 1e-12 (relative) of the
    input's and of the oracle's."""
    P, w = _case("mixed-cfg1-extreme", hybmap=hybmap)
    d0 = abi.HostDomain(P).synth(abi.make_synth(P, w, seed=100, extreme=1))
    want = _oracle_out(d0)
    P.precision = abi.PREC_MIXED
    dd = abi.DeviceDomain(P).upload(d0)
    dd.hybgen(m=1, n=2)
    got = dd.download()
    for f in ("dp", "p", "dpu", "dpv", "pu", "pv"):
        assert np.array_equal(want.a[f], got.a[f]), f"{f}: the regrid must stay FP64"
    for f in ("temp", "saln", "th3d", "tracer", "u", "v"):
        x, y = want.a[f], got.a[f]
        scale = np.max(np.abs(x))
        mx = np.max(np.abs(x - y)) / scale
        rms = np.sqrt(np.mean((x - y) ** 2)) / scale
        print(f"[mixed] {f}: max {mx:.2e} rms {rms:.2e}", flush=True)
        assert mx <= 1e-5 and rms <= 1e-5, f"{f}: max {mx:.2e} rms {rms:.2e}"
    nb = P.nbdy
    sea = d0.a["ip"][nb:nb + P.ii, nb:nb + P.jj] > 0
    for f, tr in [("temp", None), ("saln", None)] + [("tracer", q) for q in range(P.ntracr)]:
        i0 = _colint(d0, f, tr)
        den = np.maximum(np.abs(i0), 1e-300)
        r_in = (np.abs(_colint(got, f, tr) - i0) / den)[sea].max()
        r_64 = (np.abs(_colint(got, f, tr) - _colint(want, f, tr)) / den)[sea].max()
        print(f"[mixed] {f}{'' if tr is None else tr + 1}: column residual {r_in:.2e}, "
              f"vs oracle {r_64:.2e}", flush=True)
        assert r_in <= 1e-12 and r_64 <= 1e-12, f"{f}{tr}: {r_in:.2e} {r_64:.2e}"


@pytest.mark.parametrize("case", [
    dict(),
    dict(idm=97, jdm=61, kdm=41, ntracr=3, trcflg=[2, 0, 0], hybmap=3, dx00=30.0, dx00f=1.1,
         dx00x=300.0, extreme=1),
    dict(idm=97, jdm=61, hybthk=0.5, hybthn=5.0, nsigma=0, hybrlx=1.0),
    dict(idm=90, jdm=57, ldenfix=1),
    dict(idm=90, jdm=57, kdm=41, nhybrd=35, hybiso=0.01),
], ids=["cfg1", "extreme-dx0k", "hybthk", "denfix", "nhybrd<kdm"])
def test_decisions_match_oracle(gpu, kernel_path, case):
    """hycom_hybgen_record_decisions: each sea p-column's fixlay (the fixed
    layers, hybgen.F90:479-573, HYBGEN_DENFIX included) and kp (the deepest
    inflated layer after the overturn, 590-686) from the GPU equal the
    oracle's, column by column."""
    import torch
    case = dict(case)
    extreme = case.pop("extreme", 0)
    P, w = _case("decisions", **case)
    _select(kernel_path, P)
    d0 = abi.HostDomain(P).synth(abi.make_synth(P, w, seed=44, extreme=extreme))
    want = d0.copy()
    fx_o, kp_o = ol.run_oracle_decisions(want, threads=8)
    lib = abi.load()
    fx = torch.zeros(P.idm * P.jdm, dtype=torch.int32, device="cuda")
    kp = torch.zeros_like(fx)
    dd = abi.DeviceDomain(P).upload(d0)
    abi.check(lib.hycom_hybgen_record_decisions(ctypes.c_void_p(fx.data_ptr()),
                                                ctypes.c_void_p(kp.data_ptr())))
    try:
        dd.hybgen(m=1, n=2)
        torch.cuda.synchronize()
    finally:
        abi.check(lib.hycom_hybgen_record_decisions(None, None))
        abi.check(lib.hycom_hybgen_set_kernel_path(0))
    fx_g = fx.cpu().numpy().reshape(P.idm, P.jdm, order="F")
    kp_g = kp.cpu().numpy().reshape(P.idm, P.jdm, order="F")
    nb = P.nbdy
    sea = d0.a["ip"][nb:nb + P.ii, nb:nb + P.jj] > 0
    assert sea.sum() > 100
    assert np.array_equal(fx_g[sea], fx_o[sea]), f"fixlay differs at {int((fx_g != fx_o)[sea].sum())} columns"
    assert np.array_equal(kp_g[sea], kp_o[sea]), f"kp differs at {int((kp_g != kp_o)[sea].sum())} columns"
    # the decisions vary over the domain, so equality is not vacuous
    assert len(np.unique(fx_o[sea])) > 1 and len(np.unique(kp_o[sea])) > 1
    _assert_same(want, dd.download(), "outputs beside the decisions")


@pytest.mark.parametrize("hybmap,raflag,hybthk", [(0, 0, 0.0), (2, 0, 0.0), (3, 1, 0.0),
                                                  (1, 0, 0.9)])
def test_zlevel_start_matches_oracle(gpu, kernel_path, hybmap, raflag, hybthk):
    """A start from a z-level climatology (oracle_lib.zlevel_start): whole
    layers taken into too-dense massless ones, blocking layers, and
    deepest layers of negative thickness."""
    P, w = _case("zlevel-start", idm=40, jdm=30, kdm=16, hybmap=hybmap, hybthk=hybthk,
                 hybrlx=8.0, ntracr=2, trcflg=[0, 2])
    _select(kernel_path, P)
    for seed, (m, n) in ((3, (1, 2)), (4, (2, 1))):
        print(f"[seed] {seed} m,n={m},{n} raflag={raflag}", flush=True)
        d0 = abi.HostDomain(P).synth(abi.make_synth(P, w, seed=seed), m=m, n=n)
        ol.zlevel_start(d0, m=m, n=n, seed=seed)
        if raflag:
            d0.a["dpo"][...] = d0.a["dp"]
        want = _oracle_out(d0, m=m, n=n, raflag=raflag)
        dd = abi.DeviceDomain(P).upload(d0)
        dd.hybgen(m=m, n=n, raflag=raflag)
        import torch
        torch.cuda.synchronize()
        _assert_same(want, dd.download(), f"zlevel-start hybmap={hybmap} seed={seed}")


@pytest.mark.parametrize("ntracr,itr,raflag", [(0, (1, 2, 3), 0), (2, (3, 5, 4), 1)])
def test_diagnostics_after_remapped_tracers(gpu, kernel_path, ntracr, itr, raflag):
    """itracr(801..803) above ntracr, where the model keeps its mtracr
    diagnostic tracers: the device path and copy mode against the oracle."""
    P, w = _case("diag-after-tracers", idm=48, jdm=36, kdm=22, hybmap=2, ntracr=ntracr,
                 trcflg=[0, 2][:ntracr])
    P.itracr801, P.itracr802, P.itracr803 = itr
    _select(kernel_path, P)
    d0 = abi.HostDomain(P).synth(abi.make_synth(P, w, seed=50))
    rng = np.random.default_rng(51)
    d0.a["tracer"][...] = rng.standard_normal(d0.a["tracer"].shape)
    if raflag:
        d0.a["dpo"][...] = np.abs(d0.a["dp"] * (1.0 + 0.01 * rng.standard_normal(d0.a["dp"].shape)))
    want = _oracle_out(d0, raflag=raflag)
    dd = abi.DeviceDomain(P).upload(d0)
    dd.hybgen(m=1, n=2, raflag=raflag)
    import torch
    torch.cuda.synchronize()
    _assert_same(want, dd.download(), f"device, ntracr={ntracr} itracr={itr}")
    got = d0.copy().hybgen(m=1, n=2, raflag=raflag)
    _assert_same(want, got, f"copy mode, ntracr={ntracr} itracr={itr}")
