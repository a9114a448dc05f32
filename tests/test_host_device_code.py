"""The device column code of the library, compiled for the host
(tests/hostcheck), against the oracle on seeded random parameter draws:
the kernels' own arithmetic, bitwise, without a GPU.  The streaming
kernels cover any kdm; the lane-array ones every configuration."""
import numpy as np
import pytest

import hybgen_abi as abi
import hostcheck_lib as hc
import oracle_lib as ol


def _draw(rng, kdms):
    kdm = int(rng.choice(kdms))
    ntr = int(rng.integers(0, 10))
    dx = bool(rng.integers(0, 2))
    return dict(kdm=kdm, idm=int(rng.integers(20, 70)), jdm=int(rng.integers(10, 40)),
                hybmap=int(rng.integers(0, 4)), nsigma=int(rng.choice([0, min(8, kdm)])),
                hybiso=[0.0, 0.01, 3.0][int(rng.integers(0, 3))],
                hybrlx=[1.0, 8.0][int(rng.integers(0, 2))],
                hybthn=[1.0, 5.0][int(rng.integers(0, 2))],
                hybthk=[0.0, 0.5, 0.9][int(rng.integers(0, 3))],
                nreg=int(rng.integers(0, 5)), ntracr=ntr,
                trcflg=[int(x) for x in rng.choice([0, 2], size=ntr)],
                dx00=30.0 if dx else 9999.0, dx00f=1.1 if dx else 1.0,
                dx00x=300.0 if dx else 9999.0)


@pytest.mark.parametrize("family,kdms", [
    ("streaming", [4, 5, 9, 17, 22, 29, 32, 41, 50, 63, 64]),
    ("streaming-recompute", [5, 17, 22, 41, 64]),
    ("lane-arrays", [5, 22, 41, 64]),
])
@pytest.mark.parametrize("draw", range(10))
def test_device_code_on_host_matches_oracle(family, kdms, draw):
    rng = np.random.default_rng(9000 + 97 * draw + len(family))
    over = _draw(rng, kdms)
    P, w = abi.workload_params("cfg1", **over)
    raflag = int(rng.integers(0, 2))
    m, n = (1, 2) if rng.integers(0, 2) else (2, 1)
    if raflag and P.ntracr >= 3:
        P.itracr801, P.itracr802 = 1, 3
    d = abi.HostDomain(P).synth(abi.make_synth(P, w, seed=draw, extreme=int(rng.integers(0, 2))),
                                m=m, n=n)
    if raflag:
        d.a["dpo"][...] = np.abs(d.a["dp"] * (1.0 + 0.01 * rng.standard_normal(d.a["dp"].shape)))
    a, b = d.copy(), d.copy()
    ol.run_oracle(a, m=m, n=n, raflag=raflag, threads=4)
    hc.run(family, b, m=m, n=n, raflag=raflag)
    df = ol.diff_domains(a, b)
    assert not df, f"{family} {over} raflag={raflag} m={m} n={n}: {df}"


def test_streaming_fallback_columns_are_exercised():
    """The in-place check of the streaming remap sends some columns of an
    extreme domain with a maximum layer thickness (dx0k) to the copy-first
    fallback; results stay bitwise."""
    P, w = abi.workload_params("cfg1", idm=50, jdm=36, kdm=17, hybmap=2, ntracr=2,
                               trcflg=[2, 0], dx00=30.0, dx00f=1.1, dx00x=300.0)
    d = abi.HostDomain(P).synth(abi.make_synth(P, w, seed=1, extreme=1))
    a, b = d.copy(), d.copy()
    ol.run_oracle(a, threads=4)
    hc.run("streaming", b)
    assert hc.unsafe_columns() > 0
    assert not ol.diff_domains(a, b)


def test_streaming_uv_fallback_points_are_exercised():
    """u/v points whose in-place walk is not safe go to point_c; a domain
    with massless upper layers at u/v points (old dpu = 0) has them."""
    found = 0
    for seed in range(6):
        P, w = abi.workload_params("cfg1", idm=60, jdm=40, hybmap=int(seed % 4), hybrlx=1.0,
                                   hybthn=5.0, nsigma=0, nreg=3)
        d = abi.HostDomain(P).synth(abi.make_synth(P, w, seed=1000 + seed))
        a, b = d.copy(), d.copy()
        ol.run_oracle(a, threads=4)
        hc.run("streaming", b)
        found += hc.uv_fallback_points()
        assert not ol.diff_domains(a, b)
    assert found > 0


@pytest.mark.parametrize("ns", [1, 2, 15, 39, 40, 64])
@pytest.mark.parametrize("hybmap", [2, 3])
def test_split_interface_staging(ns, hybmap):
    """The remap with only ns interface levels staged (the batch kernels'
    LDS split, SplitInterfaces: the deeper levels read from p as the walk
    reaches them, starting again for each walk over the column), against
    the oracle on a kdm-41 domain: from one staged level to all of them."""
    P, w = abi.workload_params("cfg1", idm=40, jdm=30, kdm=41, hybmap=hybmap, ntracr=4,
                               trcflg=[2, 0, 0, 0])
    d = abi.HostDomain(P).synth(abi.make_synth(P, w, seed=ns + 10 * hybmap))
    a, b = d.copy(), d.copy()
    ol.run_oracle(a, threads=4)
    lib = hc.lib()
    lib.hostcheck_set_split_levels(ns)
    try:
        hc.run("streaming-recompute", b)
    finally:
        lib.hostcheck_set_split_levels(7)
    df = ol.diff_domains(a, b)
    assert not df, f"ns={ns} hybmap={hybmap}: {df}"


@pytest.mark.parametrize("family", ["streaming", "streaming-recompute", "lane-arrays"])
@pytest.mark.parametrize("hybmap,raflag,hybthk", [(0, 0, 0.0), (2, 0, 0.0), (3, 0, 0.5),
                                                  (2, 1, 0.0), (1, 1, 0.9)])
def test_device_code_on_host_zlevel_start(family, hybmap, raflag, hybthk):
    """The start from a z-level climatology (oracle_lib.zlevel_start): a
    too-dense massless layer takes a whole layer from above and the
    blocking-layer branch reads the previous level's p_hat0; a deepest
    layer of negative thickness takes the thick-thin branches at kk-1 and
    kk, which move the bottom interface."""
    P, w = abi.workload_params("cfg1", idm=40, jdm=30, kdm=16, hybmap=hybmap, hybthk=hybthk,
                               hybrlx=8.0, ntracr=2, trcflg=[0, 2])
    for seed, (m, n) in ((3, (1, 2)), (4, (2, 1))):
        d = abi.HostDomain(P).synth(abi.make_synth(P, w, seed=seed), m=m, n=n)
        ol.zlevel_start(d, m=m, n=n, seed=seed)
        if raflag:
            d.a["dpo"][...] = d.a["dp"]
        a, b = d.copy(), d.copy()
        ol.run_oracle(a, m=m, n=n, raflag=raflag, threads=4)
        hc.run(family, b, m=m, n=n, raflag=raflag)
        df = ol.diff_domains(a, b)
        assert not df, f"{family} hybmap={hybmap} raflag={raflag} seed={seed}: {df}"


@pytest.mark.parametrize("family", ["streaming", "streaming-recompute", "lane-arrays"])
@pytest.mark.parametrize("ntracr,itr,raflag", [(0, (1, 2, 3), 0), (2, (3, 5, 4), 1)])
def test_device_code_on_host_diagnostics_after_remapped_tracers(family, ntracr, itr, raflag):
    """itracr(801..803) above ntracr (the model's mtracr diagnostics):
    written, not remapped."""
    P, w = abi.workload_params("cfg1", idm=48, jdm=36, kdm=22, hybmap=2, ntracr=ntracr,
                               trcflg=[0, 2][:ntracr])
    P.itracr801, P.itracr802, P.itracr803 = itr
    d = abi.HostDomain(P).synth(abi.make_synth(P, w, seed=50))
    rng = np.random.default_rng(51)
    d.a["tracer"][...] = rng.standard_normal(d.a["tracer"].shape)
    if raflag:
        d.a["dpo"][...] = np.abs(d.a["dp"] * (1.0 + 0.01 * rng.standard_normal(d.a["dp"].shape)))
    a, b = d.copy(), d.copy()
    ol.run_oracle(a, raflag=raflag, threads=4)
    hc.run(family, b, raflag=raflag)
    df = ol.diff_domains(a, b)
    assert not df, f"{family} ntracr={ntracr} itracr={itr}: {df}"
