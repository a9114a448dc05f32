"""The .a reader into device memory (hycom_za_read_device) against the host
reader, bitwise: 32-bit records staged and widened on the device."""
import ctypes

import numpy as np
import pytest

import hybgen_abi as abi
from test_za import IDM, JDM, NB, VOID, halo, lib_read, write_a

pytestmark = pytest.mark.gpu


def test_za_read_device_matches_host(gpu, tmp_path):
    import torch
    print("\n[case] za read device, 5 records 37x23", flush=True)
    rng = np.random.default_rng(9)
    recs = [rng.standard_normal((IDM, JDM)).astype(np.float32) for _ in range(5)]
    recs[1][3, 3] = VOID
    path = str(tmp_path / "d.a")
    write_a(path, recs)
    rc, want, wmn, wmx = lib_read(path, 5)
    assert rc == 0
    init = halo(IDM, JDM, 5)
    dev = torch.from_numpy(np.ascontiguousarray(init.ravel(order="F"))).cuda()
    mn, mx = np.zeros(5), np.zeros(5)
    D = ctypes.POINTER(ctypes.c_double)
    abi.check(abi.load().hycom_za_read_device(path.encode(), IDM, JDM, NB, 1, 5, None,
                                              ctypes.c_void_p(dev.data_ptr()),
                                              mn.ctypes.data_as(D), mx.ctypes.data_as(D), None))
    got = dev.cpu().numpy().reshape(init.shape, order="F")
    assert np.array_equal(got, want)
    assert np.array_equal(mn, wmn) and np.array_equal(mx, wmx)
