"""The request loop (hycom_hybgen_server_*) on the GPU.

A request's result must not depend on what else ran beside it: every
domain stepped inside a mix of concurrent requests (continuous batching,
FP64 and mixed requests together, evictions to host memory) is compared
bitwise with the same domain stepped alone by hycom_hybgen_device, the
path the parity tests pin to the oracle.

Every case prints its name and domains before touching the GPU."""
import numpy as np
import pytest

import hybgen_abi as abi
import oracle_lib as ol

pytestmark = pytest.mark.gpu

# (idm, jdm, kdm, hybmap, nreg, ntracr, seed)
POOL = [
    (150, 100, 32, 2, 0, 2, 21),
    (97, 61, 41, 3, 1, 4, 22),
    (64, 130, 22, 0, 0, 1, 23),
    (120, 80, 41, 1, 0, 3, 24),
    (45, 37, 32, 2, 3, 0, 25),
    (200, 50, 22, 2, 1, 2, 26),
]


def _params(spec):
    idm, jdm, kdm, hybmap, nreg, ntr, seed = spec
    P, w = abi.workload_params("cfg1", idm=idm, jdm=jdm, kdm=kdm, hybmap=hybmap, nreg=nreg,
                               ntracr=ntr, trcflg=[2, 0, 0, 0][:ntr])
    return P, abi.make_synth(P, w, seed=seed, extreme=1 if seed % 2 else 0)


def _alone(spec, steps, precision):
    """The domain stepped alone by hycom_hybgen_device."""
    import torch
    P, S = _params(spec)
    P.precision = precision
    dd = abi.DeviceDomain(P).synth(S)
    for _ in range(steps):
        dd.hybgen(m=1, n=2)
    torch.cuda.synchronize()
    return dd.download()


def _wait(srv, count, timeout_s=60.0):
    import time
    got, t0 = [], time.time()
    while len(got) < count:
        got += srv.poll(wait_ms=50)
        assert time.time() - t0 < timeout_s, f"only {len(got)} of {count} requests completed"
    return got


def _same(want, got, what):
    df = ol.diff_domains(want, got)
    assert not df, f"{what}: differs bitwise from the domain stepped alone: {df}"


def test_server_alone_and_in_the_mix(gpu):
    print(f"\n[case] server mix over {POOL}", flush=True)
    x = POOL[0]
    with abi.Server(max_active=8) as srv:
        P, S = _params(x)
        dx = srv.add_domain(P, S)
        srv.submit(dx, 3, abi.PREC_FP64, tag=1)
        (c,) = _wait(srv, 1)
        assert c.status == 0 and c.tag == 1
        alone = srv.read_domain(dx)
    _same(_alone(x, 3, abi.PREC_FP64), alone, "server, alone")

    with abi.Server(max_active=8) as srv:
        ids = []
        for spec in POOL:
            P, S = _params(spec)
            ids.append(srv.add_domain(P, S))
        P, S = _params(x)
        xm = srv.add_domain(P, S)  # the same domain again, for a mixed request
        # a stream of requests over the other domains, FP64 and mixed;
        # x (3 steps FP64) and its twin (3 steps mixed) arrive while they run
        plan = [(ids[1], 5, 0), (ids[2], 2, 1), (ids[3], 4, 0), (ids[4], 6, 1), (ids[5], 3, 0),
                (ids[1], 2, 0), (ids[4], 1, 1)]
        for q, (d, n, prec) in enumerate(plan[:4]):
            srv.submit(d, n, prec, tag=10 + q)
        srv.submit(ids[0], 3, abi.PREC_FP64, tag=1)
        srv.submit(xm, 3, abi.PREC_MIXED, tag=2)
        for q, (d, n, prec) in enumerate(plan[4:]):
            srv.submit(d, n, prec, tag=20 + q)
        done = _wait(srv, len(plan) + 2)
        assert all(c.status == 0 for c in done), [(c.tag, c.status) for c in done]
        st = srv.stats()
        print(f"[stats] iterations {st.iterations} max active {st.max_active_seen}", flush=True)
        assert st.max_active_seen > 1, "the requests never ran together"
        by_tag = {c.tag: c for c in done}
        others = [by_tag[t] for t in by_tag if t >= 10]
        assert any(o.t_start < by_tag[1].t_done and by_tag[1].t_start < o.t_done
                   for o in others), "x never overlapped another request"
        mix = {d: srv.read_domain(d) for d in ids + [xm]}
    _same(_alone(x, 3, abi.PREC_FP64), mix[ids[0]], "server, FP64 request in the mix")
    _same(_alone(x, 3, abi.PREC_MIXED), mix[xm], "server, mixed request in the mix")
    # the other domains: consecutive requests of one domain run in order
    # (steps of different precisions compose, so compare against the same
    # sequence stepped alone)
    for k, spec in enumerate(POOL[1:], start=1):
        seq = [(n, prec) for d, n, prec in plan if d == ids[k]]
        import torch
        P, S = _params(spec)
        dd = abi.DeviceDomain(P).synth(S)
        for n, prec in seq:
            dd.P.precision = prec
            for _ in range(n):
                dd.hybgen(m=1, n=2)
        torch.cuda.synchronize()
        _same(dd.download(), mix[ids[k]], f"server, domain {spec} steps {seq}")


def _footprint(P):
    """Device bytes of one resident domain (the server's block layout)."""
    al = lambda b: (b + 255) // 256 * 256  # noqa: E731
    n2 = (P.idm + 2 * P.nbdy) * (P.jdm + 2 * P.nbdy)
    K = P.kdm
    n4 = 2 * n2 * K
    sizes = [8 * n4] * 8 + [8 * n4 * P.ntracr] + [8 * n2 * (K + 1)] * 3 + [8 * n2 * K]
    sizes += [8 * n2] * 4 + [24 * n2] * 2 + [4 * n2] * 3
    sizes += [8 * n2 * (K + 1), 8 * n2, 4 * n2, 8 * n2]
    return sum(al(b) for b in sizes)


def test_server_evicts_and_restores(gpu):
    """A budget that holds two of four domains: requests alternate between
    them, so domains go to host memory and come back; every result equals
    the same number of steps run alone."""
    specs = [(80, 60, 32, 2, 0, 2, 31), (70, 50, 32, 3, 0, 1, 32),
             (90, 40, 22, 2, 1, 2, 33), (60, 60, 41, 1, 0, 0, 34)]
    print(f"\n[case] server eviction over {specs}", flush=True)
    Ps = [_params(s) for s in specs]
    budget = int(2.5 * max(_footprint(P) for P, _ in Ps))
    with abi.Server(hbm_budget=budget, max_active=4) as srv:
        ids = [srv.add_domain(P, S) for P, S in Ps]
        for rnd in range(2):
            for d in ids:
                srv.submit(d, 2, abi.PREC_FP64, tag=100 * rnd + d)
        done = _wait(srv, 8)
        assert all(c.status == 0 for c in done), [(c.tag, c.status) for c in done]
        st = srv.stats()
        print(f"[stats] opens {st.opens} evictions {st.evictions} budget {budget}", flush=True)
        assert st.evictions > 0 and st.resident_bytes <= budget
        got = {d: srv.read_domain(d) for d in ids}
    for spec, d in zip(specs, ids):
        _same(_alone(spec, 4, abi.PREC_FP64), got[d], f"server eviction, domain {spec}")


def test_server_rejects(gpu):
    P, S = _params(POOL[0])
    with abi.Server(hbm_budget=_footprint(P) // 2, max_active=2) as srv:
        d = srv.add_domain(P, S)
        with pytest.raises(abi.HybError):
            srv.submit(d, 0)
        with pytest.raises(abi.HybError):
            srv.submit(d + 1, 1)
        with pytest.raises(abi.HybError):
            srv.read_domain(d)  # never stepped
        srv.submit(d, 1, tag=7)
        (c,) = _wait(srv, 1)
        assert c.tag == 7 and c.status == abi.HYB_ENOMEM
