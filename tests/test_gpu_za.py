"""The .a reader into device memory (hycom_za_read_device), pinned three
ways, bitwise: against the float32 records the test wrote, widened to
FP64; against the reference's own zaiord3 reading the same file
(href_za, mod_za_sm.h:678-862); and on records written by the
reference's zaiowr3.  The per-record ranges match the reference's.
32-bit records are staged in pinned memory and widened on the device."""
import ctypes

import numpy as np
import pytest

import hybgen_abi as abi
import oracle_lib as ol
from test_za import IDM, JDM, NB, VOID, _records, halo, ref_za, write_a

pytestmark = pytest.mark.gpu
needs_ref = pytest.mark.skipif(not ol.ref_available(), reason="reference build absent")
D = ctypes.POINTER(ctypes.c_double)
I = ctypes.POINTER(ctypes.c_int32)


def device_read(path, nrec, mask=None, rec0=1):
    import torch
    init = halo(IDM, JDM, nrec)
    dev = torch.from_numpy(np.ascontiguousarray(init.ravel(order="F"))).cuda()
    mn, mx = np.zeros(nrec), np.zeros(nrec)
    rc = abi.load().hycom_za_read_device(path.encode(), IDM, JDM, NB, rec0, nrec,
                                         None if mask is None else mask.ctypes.data_as(I),
                                         ctypes.c_void_p(dev.data_ptr()),
                                         mn.ctypes.data_as(D), mx.ctypes.data_as(D), None)
    torch.cuda.synchronize()
    got = dev.cpu().numpy().reshape(init.shape, order="F")
    return rc, got, mn, mx


def test_device_read_is_the_written_records_widened(gpu, tmp_path):
    print("\n[case] za read device vs written records, 5 records 37x23", flush=True)
    rng = np.random.default_rng(9)
    recs = [(rng.standard_normal((IDM, JDM)) * 10.0 ** q).astype(np.float32)
            for q in (-3, 0, 2, 5, 1)]
    recs[1][3, 3] = VOID
    path = str(tmp_path / "d.a")
    write_a(path, recs)
    rc, got, mn, mx = device_read(path, 5)
    assert rc == 0
    for r, x in enumerate(recs):
        assert np.array_equal(got[NB:NB + IDM, NB:NB + JDM, r], x.astype(np.float64)), r
        ok = x != VOID
        assert mn[r] == np.float64(x[ok].min()) and mx[r] == np.float64(x[ok].max()), r
    assert np.all(got[:NB] == -9.0) and np.all(got[:, :NB] == -9.0)  # halo untouched
    assert np.all(got[NB + IDM:] == -9.0) and np.all(got[:, NB + JDM:] == -9.0)
    rc, got2, _, _ = device_read(path, 2, rec0=4)  # records 4..5 alone
    assert rc == 0 and np.array_equal(got2, got[:, :, 3:5])


@needs_ref
def test_device_read_matches_reference_zaiord3(gpu, tmp_path):
    print("\n[case] za read device vs href_za read, 4 records 37x23", flush=True)
    rng = np.random.default_rng(2)
    recs = [rng.standard_normal((IDM, JDM)).astype(np.float32) for _ in range(4)]
    recs[2][5, 6] = VOID
    path = str(tmp_path / "t.a")
    write_a(path, recs)
    h = halo(IDM, JDM, 4)
    ref_mn, ref_mx = ref_za(0, path, h)
    rc, got, mn, mx = device_read(path, 4)
    assert rc == 0
    assert np.array_equal(got[NB:NB + IDM, NB:NB + JDM], h[NB:NB + IDM, NB:NB + JDM])
    assert np.array_equal(mn, ref_mn) and np.array_equal(mx, ref_mx)


@needs_ref
def test_device_read_of_reference_written_records(gpu, tmp_path):
    print("\n[case] za read device of href_za-written records, 5 records 37x23", flush=True)
    rng = np.random.default_rng(1)
    h = _records(rng, 5)
    path = str(tmp_path / "ref.a")
    ref_mn, ref_mx = ref_za(1, path, h.copy(order="F"))
    rc, got, mn, mx = device_read(path, 5)
    assert rc == 0
    want = h[NB:NB + IDM, NB:NB + JDM, :].astype(np.float32).astype(np.float64)
    assert np.array_equal(got[NB:NB + IDM, NB:NB + JDM, :], want)
    assert np.array_equal(mn, ref_mn) and np.array_equal(mx, ref_mx)


def test_device_read_refuses_a_mask(gpu, tmp_path):
    """The device reader writes whole records; a masked read is the host
    reader's (hycom_za_read)."""
    print("\n[case] za read device with a mask: refused", flush=True)
    rng = np.random.default_rng(3)
    path = str(tmp_path / "m.a")
    write_a(path, [rng.standard_normal((IDM, JDM)).astype(np.float32)])
    mask = np.asfortranarray(np.ones((IDM + 2 * NB, JDM + 2 * NB), dtype=np.int32))
    rc, got, _, _ = device_read(path, 1, mask=mask)
    assert rc == abi.HYB_EUNSUPPORTED
    assert np.all(got == -9.0)
