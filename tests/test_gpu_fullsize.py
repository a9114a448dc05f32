"""The GLBb step at its full size (4500x3298x41, 4 tracers, PPM and WENO),
checked through properties that do not depend on the domain's size: the
oracle runs on windows of the same field (test_gpu_parity.py), and here the
whole field must show, after one call,
  * every sea column's integrals of dp*T, dp*S and dp*tracer at time level
    n unchanged to roundoff (the remap conserves them; hybgen.F90:1583-1589);
  * interfaces that start at 0, never decrease, and keep the bottom that dp
    gave on entry (p(kk+1) as the running sum of dp, hybgen.F90:413-416);
  * a trcflg=2 tracer still bitwise equal to T (the reference's
    PIPE_TRACER check);
  * the same bits from the same input twice (no race between columns).
Every case prints its name before touching the GPU."""
import pytest

import hybgen_abi as abi

pytestmark = pytest.mark.gpu


def _views(dd, P):
    NI, NJ, K, T = P.idm + 2 * P.nbdy, P.jdm + 2 * P.nbdy, P.kdm, P.ntracr
    lev = lambda n: dd.t[n].view(2, K, NJ, NI)[1]  # noqa: E731  level n = 2
    trc = lambda q: dd.t["tracer"].view(T, 2, K, NJ, NI)[q, 1]  # noqa: E731
    return NI, NJ, K, T, lev, trc


@pytest.mark.parametrize("hybmap", [2, 3], ids=["ppm", "weno"])
def test_full_glbb_step_properties(gpu, hybmap):
    import torch
    print(f"\n[case] full GLBb 4500x3298x41 step, hybmap={hybmap}, trcflg 2 on tracer 1",
          flush=True)
    P, w = abi.workload_params("glbb", hybmap=hybmap, trcflg=[2, 0, 0, 0])
    dd = abi.DeviceDomain(P, skip=("dpo",))
    S = abi.make_synth(P, w, seed=200)
    dd.synth(S)
    torch.cuda.synchronize()
    NI, NJ, K, T, lev, trc = _views(dd, P)
    nb = P.nbdy
    sea = torch.zeros((NJ, NI), dtype=torch.bool, device="cuda")
    sea[nb:nb + P.jj, nb:nb + P.ii] = dd.t["ip"].view(NJ, NI)[nb:nb + P.jj, nb:nb + P.ii] > 0
    nsea = int(sea.sum().item())
    assert nsea > 9_000_000

    def colints():
        dp = lev("dp")
        out = {"temp": (dp * lev("temp")).sum(0), "saln": (dp * lev("saln")).sum(0)}
        for q in range(T):
            out[f"tracer{q + 1}"] = (dp * trc(q)).sum(0)
        return {k: v[sea] for k, v in out.items()}

    before = colints()
    scale = {k: (lev("dp") * (lev("temp") if k == "temp" else lev("saln") if k == "saln"
                              else trc(int(k[6:]) - 1)).abs()).sum(0)[sea] for k in before}
    # the bottom the call starts from: p(kk+1) as dp's running sum, in the
    # reference's order of additions
    bottom = torch.zeros((NJ, NI), dtype=torch.float64, device="cuda")
    dp0 = lev("dp")
    for k in range(K):
        bottom = bottom + dp0[k]
    bottom = bottom[sea].clone()
    del dp0

    dd.hybgen(m=1, n=2)
    torch.cuda.synchronize()
    after = colints()
    p = dd.t["p"].view(K + 1, NJ, NI)
    assert bool((p[0][sea] == 0.0).all()), "p(1) is not 0 at a sea point"
    for k in range(K):
        assert bool((p[k + 1][sea] >= p[k][sea]).all()), f"p({k + 2}) < p({k + 1}) somewhere"
    assert bool((lev("dp")[:, sea] >= 0.0).all()), "a negative thickness"
    # the bottom stays, except where the layer-1 floor (1051-1063) lifts a
    # column shallower than its first layer's minimum thickness; such a
    # column gains mass, so the conservation check leaves it out
    bot = p[K][sea]
    moved = bot != bottom
    nmoved = int(moved.sum().item())
    print(f"[bottom] {nmoved} of {nsea} sea columns lifted by the layer-1 floor", flush=True)
    assert bool((bot >= bottom).all()), "a bottom p(kk+1) went up the column"
    assert nmoved <= nsea // 1000
    if nmoved:
        assert float(bottom[moved].max().item()) < 20.0 * abi.ONEM
    keep = ~moved
    worst = {}
    for k in before:
        r = ((after[k] - before[k]).abs() / scale[k].clamp_min(1e-300))[keep]
        worst[k] = float(r.max().item())
    print(f"[conservation] max relative column residual {worst}", flush=True)
    for k, r in worst.items():
        assert r <= 1e-12, f"{k}: a column's integral changed by {r:.3e} of its scale"
    # trcflg = 2: tracer 1 is a bit copy of T
    assert torch.equal(trc(0)[:, sea].view(torch.int64), lev("temp")[:, sea].view(torch.int64))

    # the same input again gives the same bits: integer checksums of every
    # output's words (sums mod 2^64, so the reduction order does not matter)
    outs = ("temp", "saln", "th3d", "dp", "u", "v", "dpu", "dpv", "tracer", "p", "pu", "pv")

    def sums():
        out = {}
        for n in outs:
            x = dd.t[n].view(torch.int64)
            acc = [0, 0, 0]
            for c in x.split(1 << 27):  # 1 GiB at a time
                acc[0] += int(c.sum().item())
                acc[1] += int((c * 0x3C6EF372FE94F82B).sum().item())
                acc[2] += int((c ^ (c >> 29)).sum().item())
            out[n] = tuple(a % (1 << 64) for a in acc)
        return out

    first = sums()
    dd.synth(S)
    dd.hybgen(m=1, n=2)
    torch.cuda.synchronize()
    again = sums()
    for n in outs:
        assert again[n] == first[n], f"{n} differs between two runs of the same input"
