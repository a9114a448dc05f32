"""Batched ragged domains (hycom_hybgen_batch) on the GPU.

Every domain of a batch must come out bitwise equal to the same domain
stepped alone by hycom_hybgen_device, which the parity tests pin to the
oracle; the FP64 domains are also checked against the oracle directly.
The first calls of a new domain set run the autotuner's candidate widths,
so comparing every call also checks that no width changes a bit.

Every case prints its name and domains before touching the GPU."""
import numpy as np
import pytest

import hybgen_abi as abi
import oracle_lib as ol

pytestmark = pytest.mark.gpu

# (idm, jdm, kdm, hybmap, precision, nreg, ntracr, extreme, seed[, ldenfix])
RAGGED = [
    (100, 100, 41, 2, 0, 1, 4, 0, 1),
    (37, 211, 22, 0, 0, 0, 1, 0, 2),
    (97, 61, 32, 3, 1, 0, 2, 1, 3),
    (160, 45, 41, 1, 0, 0, 3, 0, 4),
    (64, 64, 22, 2, 1, 1, 0, 0, 5),
    (131, 77, 32, 2, 0, 3, 2, 1, 6),
    (70, 120, 41, 3, 0, 0, 4, 1, 7),
    (45, 33, 32, 1, 1, 0, 1, 0, 8),
    (210, 52, 22, 3, 0, 1, 2, 0, 9),
    (90, 70, 41, 2, 0, 0, 2, 0, 10, 1),  # HYBGEN_DENFIX
    (80, 90, 32, 3, 1, 1, 1, 0, 11, 1),  # HYBGEN_DENFIX, mixed
]


def _domain(spec, n=2):
    idm, jdm, kdm, hybmap, prec, nreg, ntr, extreme, seed = spec[:9]
    denfix = spec[9] if len(spec) > 9 else 0
    dx = dict(dx00=30.0, dx00f=1.1, dx00x=300.0) if extreme else {}
    P, w = abi.workload_params("cfg1", idm=idm, jdm=jdm, kdm=kdm, hybmap=hybmap,
                               nsigma=8, nreg=nreg, ntracr=ntr, ldenfix=denfix,
                               trcflg=[2, 0, 0, 0][:ntr], **dx)
    d0 = abi.HostDomain(P).synth(abi.make_synth(P, w, seed=seed, extreme=extreme),
                                 m=3 - n, n=n)
    P.precision = prec
    return P, d0


def _pairs(specs, n=2):
    print(f"\n[case] batch of {len(specs)} domains: {specs}", flush=True)
    out = []
    for spec in specs:
        P, d0 = _domain(spec, n)
        a = abi.DeviceDomain(P).upload(d0)
        b = abi.DeviceDomain(P).upload(d0)
        out.append((spec, d0, a, b))
    return out


def _same_on_device(a, b):
    import torch
    bad = [f for f in abi.OUTPUTS if f in a.t and not torch.equal(a.t[f], b.t[f])]
    return bad


def test_batch_matches_single_domains(gpu):
    import torch
    abi.check(abi.load().hycom_hybgen_batch_autotune(1))
    pairs = _pairs(RAGGED)
    batch = abi.Batch([a for _, _, a, _ in pairs])
    # every candidate shape twice (remap columns x staged levels, then the
    # regrid widths), then the tuned one
    for call in range(40):
        if call > 0 and batch.config()[2]:
            break
        batch.hybgen(m=1, n=2)
        for _, _, _, b in pairs:
            b.hybgen(m=1, n=2)
        torch.cuda.synchronize()
        for spec, _, a, b in pairs:
            bad = _same_on_device(a, b)
            assert not bad, f"call {call} domain {spec}: {bad} differ, widths {batch.config()}"
        if call == 0:
            # the FP64 domains against the oracle, after one step
            for spec, d0, a, _ in pairs:
                if spec[4] != 0:
                    continue
                want = d0.copy()
                ol.run_oracle(want, m=1, n=2, threads=8)
                df = ol.diff_domains(want, a.download())
                assert not df, f"domain {spec} differs from the oracle: {df}"
    rc, gc, tuned = batch.config()
    ns, kmax = batch.staged_levels()
    print(f"[tuned] remap {rc} regrid {gc}, {ns} of {kmax - 1} levels staged, tuned={tuned} "
          f"after {call} calls", flush=True)
    assert tuned and rc in (64, 128, 256) and gc in (64, 128, 256)
    assert ns in (kmax - 1, 24, 12) and 8 * ns * rc <= 65536


def test_batch_both_time_levels_and_set_change(gpu):
    """n=1 and n=2 replay their own graphs; a different domain set builds a
    new plan, and going back rebuilds the first one."""
    import torch
    abi.check(abi.load().hycom_hybgen_batch_autotune(0))
    try:
        for n in (1, 2):
            pairs = _pairs(RAGGED[:5], n=n)
            full = abi.Batch([a for _, _, a, _ in pairs])
            part = abi.Batch([a for _, _, a, _ in pairs[1:4]])
            for batch, members in ((full, pairs), (part, pairs[1:4]), (full, pairs)):
                batch.hybgen(m=3 - n, n=n)
                for _, _, _, b in members:
                    b.hybgen(m=3 - n, n=n)
                torch.cuda.synchronize()
                for spec, _, a, b in members:
                    bad = _same_on_device(a, b)
                    assert not bad, f"n={n} domain {spec}: {bad} differ"
    finally:
        abi.check(abi.load().hycom_hybgen_batch_autotune(1))


def test_batch_of_one_equals_device_call(gpu):
    import torch
    pairs = _pairs([(90, 57, 41, 2, 0, 1, 4, 1, 11)])
    _, _, a, b = pairs[0]
    abi.Batch([a]).hybgen(m=1, n=2)
    b.hybgen(m=1, n=2)
    torch.cuda.synchronize()
    assert not _same_on_device(a, b)


def test_batch_of_config3_shape_matches_oracle(gpu):
    """The configs[2] batch shape: 64 regional domains of random shape
    (capped at 300x200 so the oracle finishes in seconds), kdm drawn from
    {22, 32, 41}, shelf fraction 0-60% (vanished layers), all four
    remappers, FP64, in one hycom_hybgen_batch call; every domain is
    compared with the oracle bitwise.  The domains are generated on the
    device and again on the host, one at a time, for the check (the two
    generators give the same bits)."""
    import torch
    rng = np.random.default_rng(2024)
    specs = []
    for d in range(64):
        specs.append(dict(idm=int(rng.integers(100, 301)), jdm=int(rng.integers(100, 201)),
                          kdm=int(rng.choice([22, 32, 41])), shelf=float(rng.uniform(0.0, 0.6)),
                          hybmap=int(rng.integers(0, 4)), nreg=int(rng.choice([0, 1])),
                          seed=300000 + 100 * d))
    print(f"\n[case] config-3 shaped batch: {len(specs)} domains "
          f"{[(s['idm'], s['jdm'], s['kdm']) for s in specs]}", flush=True)
    doms, synths = [], []
    for s in specs:
        P, w = abi.workload_params("cfg1", idm=s["idm"], jdm=s["jdm"], kdm=s["kdm"],
                                   hybmap=s["hybmap"], nreg=s["nreg"], shelf=s["shelf"],
                                   nsigma=8)
        S = abi.make_synth(P, w, seed=s["seed"])
        doms.append(abi.DeviceDomain(P, skip=("dpo",)).synth(S))
        synths.append(S)
    torch.cuda.synchronize()
    abi.check(abi.load().hycom_hybgen_batch_autotune(0))
    try:
        abi.Batch(doms).hybgen(m=1, n=2)
        torch.cuda.synchronize()
    finally:
        abi.check(abi.load().hycom_hybgen_batch_autotune(1))
    shelf_cols = 0
    for s, dd, S in zip(specs, doms, synths):
        want = abi.HostDomain(dd.P).synth(S)
        nb = dd.P.nbdy
        ip = want.a["ip"][nb:nb + dd.P.ii, nb:nb + dd.P.jj] > 0
        shelf_cols += int((want.a["depths"][nb:nb + dd.P.ii, nb:nb + dd.P.jj][ip] < 200.0).sum())
        ol.run_oracle(want, m=1, n=2, threads=8)
        got = dd.download([n for n in abi.OUTPUTS])
        got = abi.HostDomain(dd.P, {**{k: v for k, v in want.a.items()}, **got})
        df = ol.diff_domains(want, got)
        assert not df, f"domain {s} differs from the oracle: {df}"
    assert shelf_cols > 10000  # shelf-heavy columns (shallow, vanished layers) are in the set
