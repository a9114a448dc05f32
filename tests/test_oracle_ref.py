"""The oracle (C restatement, oracle/hybgen_oracle.c) against the live flang
build of the reference hybgen (oracle/_ref), on seeded random parameter
draws beyond the golden set, and the four remappers column by column.
Skipped where the reference build is absent (it needs /root/reference)."""
import ctypes

import numpy as np
import pytest

import hybgen_abi as abi
import oracle_lib as ol

pytestmark = pytest.mark.skipif(not (ol.ref_available() and ol.ref_available(True)),
                                reason="oracle/_ref not built (make -C oracle/ref)")


def _draw(rng):
    kdm = int(rng.choice([5, 12, 22, 32, 41, 50, 64]))
    ntr = int(rng.integers(0, 6))
    dx = bool(rng.integers(0, 2))
    return dict(kdm=kdm, idm=int(rng.integers(20, 60)), jdm=int(rng.integers(10, 40)),
                hybmap=int(rng.integers(0, 4)), nsigma=int(rng.choice([0, min(8, kdm)])),
                hybiso=[0.0, 0.01, 3.0][int(rng.integers(0, 3))],
                hybrlx=[1.0, 3.0, 8.0][int(rng.integers(0, 3))],
                hybthn=[1.0, 5.0][int(rng.integers(0, 2))],
                hybthk=[0.0, 0.5, 0.9][int(rng.integers(0, 3))],
                ldenfix=int(rng.integers(0, 2)), nreg=int(rng.integers(0, 5)),
                ntracr=ntr, trcflg=[int(x) for x in rng.choice([0, 2], size=ntr)],
                dx00=30.0 if dx else 9999.0, dx00f=1.1 if dx else 1.0,
                dx00x=300.0 if dx else 9999.0)


@pytest.mark.parametrize("draw", range(24))
def test_oracle_matches_reference_random(draw):
    rng = np.random.default_rng(5000 + draw)
    over = _draw(rng)
    P, w = abi.workload_params("cfg1", **over)
    raflag = int(rng.integers(0, 2))
    m, n = (1, 2) if rng.integers(0, 2) else (2, 1)
    if raflag and P.ntracr >= 2:
        P.itracr801, P.itracr803 = 1, 2
    d = abi.HostDomain(P).synth(abi.make_synth(P, w, seed=draw, extreme=int(rng.integers(0, 2))),
                                m=m, n=n)
    if raflag:
        d.a["dpo"][...] = np.abs(d.a["dp"] * (1.0 + 0.01 * rng.standard_normal(d.a["dp"].shape)))
    a, b = d.copy(), d.copy()
    ol.run_oracle(a, m=m, n=n, raflag=raflag, threads=4)
    ol.run_ref(b, m=m, n=n, raflag=raflag)
    df = ol.diff_domains(a, b)
    assert not df, f"{over} raflag={raflag} m={m} n={n}: {df}"


def _remap_pair(method, kk, si, pi, po, lcm, thin):
    D = ctypes.POINTER(ctypes.c_double)
    I = ctypes.POINTER(ctypes.c_int)
    so_a = np.zeros(kk)
    so_b = np.zeros(kk)
    args = (si.ctypes.data_as(D), pi.ctypes.data_as(D), po.ctypes.data_as(D),
            lcm.ctypes.data_as(I), ctypes.c_double(thin))
    ol.oracle().oracle_remap_column(method, kk, *args, so_a.ctypes.data_as(D))
    ol.ref().href_remap(method, kk, *args, so_b.ctypes.data_as(D))
    return so_a, so_b


@pytest.mark.parametrize("method", [0, 1, 2, 3])
def test_remappers_match_reference_columns(method):
    rng = np.random.default_rng(77 + method)
    thin = 0.001 * 9.806 * 0.001 * 1000.0 * 0.001
    for trial in range(300):
        kk = int(rng.integers(3, 40))
        dp = rng.exponential(1.0, kk) * (rng.random(kk) > 0.2)
        pi = np.concatenate([[0.0], np.cumsum(dp)])
        dq = rng.exponential(1.0, kk) * (rng.random(kk) > 0.3)
        po = np.concatenate([[0.0], np.cumsum(dq)])
        po *= pi[-1] / po[-1] if po[-1] > 0 else 0.0
        si = np.cumsum(rng.standard_normal(kk))
        lcm = (rng.random(kk) > 0.8).astype(np.int32)
        a, b = _remap_pair(method, kk, si, pi, po, lcm, thin)
        assert np.array_equal(a, b), f"method {method} trial {trial} kk={kk}"


def test_eos_matches_reference_on_grid():
    """sig() (stmt_fns.h:505-511, EOS_17T sigma-2) bitwise on a T,S grid
    that spans the EOS_17T range and beyond (negative salinity takes the
    max(0,s) branch of the square root)."""
    o, r = ol.oracle(), ol.ref()
    for t in np.linspace(-3.0, 42.0, 61):
        for s in np.concatenate([np.linspace(-1.0, 42.0, 53), [0.0, 1e-300]]):
            a, b = o.oracle_sig(t, s), r.href_sig(t, s)
            assert a == b or (np.isnan(a) and np.isnan(b)), (t, s, a, b)


@pytest.mark.parametrize("seed,hybmap", [(1, 2), (2, 3), (3, 0)])
def test_oracle_fixed_layers_agree_with_reference_output(seed, hybmap):
    """The oracle's exported fixlay (hybgen.F90:479-573) against the
    reference's own output: where qdep = 1 (nsigma = 0 here) and no maximum
    thickness applies, the reference keeps interface k+1 at
    min(dp0cum(k+1), p(kk+1)) for every k <= fixlay (1092-1097), dp0cum
    being the running sum of dp0k (468).  An oracle that counts one fixed
    layer too many fails at that level.  The oracle's outputs beside its
    decisions are the reference's bits."""
    P, w = abi.workload_params("cfg1", idm=60, jdm=40, nsigma=0, hybmap=hybmap)
    d0 = abi.HostDomain(P).synth(abi.make_synth(P, w, seed=seed))
    ref = d0.copy()
    ol.run_ref(ref)
    orc = d0.copy()
    fx, _ = ol.run_oracle_decisions(orc, threads=4)
    assert not ol.diff_domains(ref, orc)
    K, nb = P.kdm, P.nbdy
    dp0cum = np.concatenate([[0.0], np.cumsum(np.array(P.dp0k[:K]))])  # dp0cum(1..K+1)
    p = ref.a["p"][nb:nb + P.ii, nb:nb + P.jj, :]
    sea = d0.a["ip"][nb:nb + P.ii, nb:nb + P.jj] > 0
    checked = 0
    for i, j in np.argwhere(sea):
        for k in range(1, min(int(fx[i, j]), K - 1) + 1):
            assert p[i, j, k] == min(dp0cum[k], p[i, j, K]), (i, j, k, int(fx[i, j]))
            checked += 1
    assert checked > 10 * sea.sum()


def test_multi_step_oracle_matches_reference():
    """The oracle against the reference hybgen over 8 model steps, each
    state evolving on its own through the stand-in dynamics of the shim
    A/B (leapfrog time levels, Robert-Asselin branch on), bit for bit after
    every call."""
    import test_fortran_shim as T
    d = T._domain(hybmap=3, ntracr=3, trcflg=[2, 0, 0])
    a, b = d.copy(), d.copy()
    for step in range(8):
        m, n = (1, 2) if step % 2 == 0 else (2, 1)
        T._dynamics(a, m, n, step)
        T._dynamics(b, m, n, step)
        ol.run_ref(a, m=m, n=n, raflag=1)
        ol.run_oracle(b, m=m, n=n, raflag=1)
        df = ol.diff_domains(a, b)
        assert not df, f"step {step}: {df}"


@pytest.mark.parametrize("kdm", [16, 41])
@pytest.mark.parametrize("hybmap", [0, 1, 2, 3])
@pytest.mark.parametrize("raflag,ldenfix,hybthk", [(0, 0, 0.0), (1, 0, 0.0), (0, 1, 0.0),
                                                   (0, 0, 0.5), (1, 0, 0.9)])
def test_oracle_matches_reference_zlevel_start(kdm, hybmap, raflag, ldenfix, hybthk):
    """A start from a z-level climatology (oracle_lib.zlevel_start): the
    first call takes whole layers into too-dense massless ones and reaches
    the blocking-layer branch right after (hybgen.F90:1259-1330), whose
    p_hat0 is the previous level's value from 1479; a deepest layer of
    negative thickness takes the thick-thin branches (1131-1208) at kk-1
    and kk."""
    rng = np.random.default_rng(700 + kdm + 10 * hybmap + raflag + 3 * ldenfix + int(10 * hybthk))
    P, w = abi.workload_params("cfg1", idm=40, jdm=30, kdm=kdm, hybmap=hybmap, hybthk=hybthk,
                               hybrlx=8.0, ldenfix=ldenfix, ntracr=2, trcflg=[0, 2])
    for mn in ((1, 2), (2, 1)):
        d = abi.HostDomain(P).synth(abi.make_synth(P, w, seed=int(rng.integers(1000))),
                                    m=mn[0], n=mn[1])
        ol.zlevel_start(d, m=mn[0], n=mn[1], seed=int(rng.integers(1000)))
        if raflag:
            d.a["dpo"][...] = d.a["dp"]
        a, b = d.copy(), d.copy()
        ol.run_oracle(a, m=mn[0], n=mn[1], raflag=raflag, threads=4)
        ol.run_ref(b, m=mn[0], n=mn[1], raflag=raflag)
        df = ol.diff_domains(a, b)
        assert not df, f"kdm={kdm} hybmap={hybmap} raflag={raflag} m,n={mn}: {df}"


@pytest.mark.parametrize("ntracr,itr,raflag", [(0, (1, 2, 3), 0), (2, (3, 5, 4), 1),
                                               (1, (0, 2, 0), 0), (3, (4, 0, 6), 1)])
def test_diagnostic_tracers_after_the_remapped_ones(ntracr, itr, raflag):
    """itracr(801..803) above ntracr, where the model keeps its mtracr
    diagnostic tracers (blkdat.F90:1307-1330): written before and after
    hybgenaj (hybgen.F90:44-164) and not remapped, since hybgenaj's tracer
    loops stop at ntracr."""
    P, w = abi.workload_params("cfg1", idm=48, jdm=36, kdm=22, hybmap=2, ntracr=ntracr,
                               trcflg=[0, 2, 0][:ntracr])
    P.itracr801, P.itracr802, P.itracr803 = itr
    for mn in ((1, 2), (2, 1)):
        d = abi.HostDomain(P).synth(abi.make_synth(P, w, seed=40 + ntracr), m=mn[0], n=mn[1])
        rng = np.random.default_rng(41)
        d.a["tracer"][...] = rng.standard_normal(d.a["tracer"].shape)
        for name in ("tracer",):
            assert d.a[name].shape[-1] == max(ntracr, *itr)
        if raflag:
            d.a["dpo"][...] = np.abs(d.a["dp"] * (1.0 + 0.01 * rng.standard_normal(d.a["dp"].shape)))
        a, b = d.copy(), d.copy()
        ol.run_oracle(a, m=mn[0], n=mn[1], raflag=raflag, threads=4)
        ol.run_ref(b, m=mn[0], n=mn[1], raflag=raflag)
        df = ol.diff_domains(a, b)
        assert not df, f"ntracr={ntracr} itracr={itr} raflag={raflag} m,n={mn}: {df}"
