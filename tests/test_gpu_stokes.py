"""GPU parity of the Stokes drift remap (hycom_stokes_vertical*,
mod_stokes.F90:541-546, 650-705) against the oracle, bitwise.  Every case
prints its name and shape before touching the GPU."""
import ctypes

import numpy as np
import pytest

import hybgen_abi as abi
import oracle_lib as ol
from test_stokes import CASES, _domain

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kdm,nsdzi,depth_max", CASES)
def test_stokes_host_call_matches_oracle(gpu, kdm, nsdzi, depth_max):
    print(f"\n[case] stokes host call kdm={kdm} nsdzi={nsdzi} depth_max={depth_max}", flush=True)
    sd = _domain(kdm, nsdzi, seed=kdm * 7 + nsdzi, depth_max=depth_max)
    a, b = sd.copy(), sd.copy()
    ol.run_oracle_stokes(a)
    b.run()
    assert np.array_equal(a.usdp, b.usdp)
    assert np.array_equal(a.vsdp, b.vsdp)


@pytest.mark.parametrize("n", [1, 2])
def test_stokes_device_call_matches_oracle(gpu, n):
    """Device pointers, a short model column (sum of dp(n) above the
    bottom) and the snapshots' deepest levels outside the rest."""
    import torch
    print(f"\n[case] stokes device call kdm=41 nsdzi=30 n={n}", flush=True)
    sd = _domain(41, 30, seed=20 + n, depth_max=600.0, n=n)
    rng = np.random.default_rng(n)
    dp = sd.dom.a["dp"]
    dp[:, :, :, n - 1] *= rng.uniform(0.6, 1.0, dp.shape[:2])[:, :, None]
    sd.usdz[:, :, :, -2:] = rng.choice([-2.0, 2.0], sd.usdz[:, :, :, -2:].shape)
    a = sd.copy()
    ol.run_oracle_stokes(a)
    dev = {}
    for name, arr in (("usdz", sd.usdz), ("vsdz", sd.vsdz), ("dp", dp), ("p", sd.dom.a["p"]),
                      ("ip", sd.dom.a["ip"]), ("usdp", sd.usdp), ("vsdp", sd.vsdp)):
        dev[name] = torch.from_numpy(np.ascontiguousarray(arr.ravel(order="F"))).cuda()
    F = abi.HybStokesFields()
    for name, t in dev.items():
        setattr(F, name, t.data_ptr())
    F.sdzi = sd.sdzi.ctypes.data
    abi.check(abi.load().hycom_stokes_vertical_device(ctypes.byref(sd.P), ctypes.byref(F), n,
                                                      None))
    torch.cuda.synchronize()
    shape = sd.usdp.shape
    got_u = dev["usdp"].cpu().numpy().reshape(shape, order="F")
    got_v = dev["vsdp"].cpu().numpy().reshape(shape, order="F")
    assert np.array_equal(a.usdp, got_u)
    assert np.array_equal(a.vsdp, got_v)
