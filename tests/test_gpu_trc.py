"""GPU parity of the tracer column remaps (hycom_trc_remap*, trcupd.F90:
808-1147) against the oracle, bitwise.  Every case prints its name and
shape before touching the GPU."""
import ctypes

import numpy as np
import pytest

import hybgen_abi as abi
import oracle_lib as ol
from test_trc import _fields, _grid

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("method", [0, 1])
@pytest.mark.parametrize("ki,ko,ks", [(2, 9, 1), (41, 41, 4), (17, 30, 6), (64, 12, 2)])
def test_trc_host_call_matches_oracle(gpu, method, ki, ko, ks):
    print(f"\n[case] trc host call method={method} ki={ki} ko={ko} ks={ks}", flush=True)
    rng = np.random.default_rng(11 * ki + ko + method)
    P, a = _grid(rng, method, ki, ko, ks)
    b = {k: v.copy(order="F") for k, v in a.items()}
    assert ol.oracle().oracle_trc_remap(ctypes.byref(P), ctypes.byref(_fields(a))) == 0
    abi.check(abi.load().hycom_trc_remap(ctypes.byref(P), ctypes.byref(_fields(b))))
    assert np.array_equal(a["so"], b["so"])


def test_trc_device_call_matches_oracle(gpu):
    import torch
    print("\n[case] trc device call method=1 ki=33 ko=41 ks=3", flush=True)
    rng = np.random.default_rng(5)
    P, a = _grid(rng, 1, 33, 41, 3)
    want = {k: v.copy(order="F") for k, v in a.items()}
    assert ol.oracle().oracle_trc_remap(ctypes.byref(P), ctypes.byref(_fields(want))) == 0
    dev = {k: torch.from_numpy(np.ascontiguousarray(v.ravel(order="F"))).cuda()
           for k, v in a.items()}
    F = abi.HybTrcFields()
    for k, t in dev.items():
        setattr(F, k, t.data_ptr())
    abi.check(abi.load().hycom_trc_remap_device(ctypes.byref(P), ctypes.byref(F), None))
    torch.cuda.synchronize()
    got = dev["so"].cpu().numpy().reshape(a["so"].shape, order="F")
    assert np.array_equal(want["so"], got)


def test_trc_out_of_order_interfaces_are_reported(gpu):
    print("\n[case] trc host call, new interfaces out of order at one column", flush=True)
    rng = np.random.default_rng(3)
    P, a = _grid(rng, 1, 12, 12, 2, idm=8, jdm=5)
    a["mask"][...] = 0
    a["mask"][4, 4] = 1
    po = a["po"][4, 4]
    po[6] = po[3] - 1.0
    po[7:] = np.maximum(po[7:], po[5])
    b = {k: v.copy(order="F") for k, v in a.items()}
    assert ol.oracle().oracle_trc_remap(ctypes.byref(P), ctypes.byref(_fields(a))) == abi.HYB_EINVAL
    rc = abi.load().hycom_trc_remap(ctypes.byref(P), ctypes.byref(_fields(b)))
    assert rc == abi.HYB_EINVAL
    assert b"bad lf" in abi.load().hycom_hybgen_last_error()
    assert np.array_equal(a["so"], b["so"], equal_nan=True)


def test_trc_glbb_window_matches_oracle(gpu):
    """plmtrc as thermf calls it, on the inputs the bench's trc leg builds,
    at window size: a 600x400 window of the GLBb field after one hybgen
    step, 4 tracers of seeded climatology on interfaces half a layer off the
    model's p, remapped onto p; against the oracle, bitwise."""
    import torch
    print("\n[case] trc GLBb window 600x400x41 after a hybgen step, plmtrc, 4 tracers",
          flush=True)
    P, w = abi.workload_params("glbb", idm=600, jdm=400)
    S = abi.make_synth(P, w, seed=200, itdm=4500, jtdm=3298, i0=1800, j0=1200)
    dd = abi.DeviceDomain(P, skip=("dpo",)).synth(S)
    dd.hybgen(m=1, n=2)
    torch.cuda.synchronize()
    NI, NJ = P.idm + 2 * P.nbdy, P.jdm + 2 * P.nbdy
    nij, K, T = NI * NJ, P.kdm, 4
    p = dd.t["p"].view(K + 1, nij)
    pi = torch.empty_like(p)
    pi[0] = 0.0
    pi[1:K] = 0.5 * (p[1:K] + p[2:K + 1])
    pi[K] = p[K]
    g = torch.Generator(device="cuda").manual_seed(7)
    si = torch.cumsum(torch.randn(T, K, nij, device="cuda", generator=g, dtype=torch.float64),
                      dim=1) + 5.0
    so = torch.zeros(T, K, nij, device="cuda", dtype=torch.float64)
    TP = abi.HybTrcParams(idm=P.idm, jdm=P.jdm, ii=P.ii, jj=P.jj, nbdy=P.nbdy, ki=K, ko=K,
                          ks=T, method=1, reserved=0)
    F = abi.HybTrcFields()
    F.si, F.pi, F.po = si.data_ptr(), pi.data_ptr(), p.data_ptr()
    F.mask, F.so = dd.t["ip"].data_ptr(), so.data_ptr()
    abi.check(abi.load().hycom_trc_remap_device(ctypes.byref(TP), ctypes.byref(F), None))
    torch.cuda.synchronize()
    got = so.cpu().numpy()
    h = {"si": si.cpu().numpy(), "pi": pi.cpu().numpy(), "po": p.contiguous().cpu().numpy(),
         "mask": dd.t["ip"].cpu().numpy(), "so": np.zeros((T, K, nij))}
    H = abi.HybTrcFields()
    for k, v in h.items():
        setattr(H, k, v.ctypes.data)
    assert ol.oracle().oracle_trc_remap(ctypes.byref(TP), ctypes.byref(H)) == 0
    sea = h["mask"].reshape(nij) != 0
    assert sea.sum() > 100000
    assert np.array_equal(h["so"][:, :, sea], got[:, :, sea])
