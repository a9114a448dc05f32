"""hycom_hybgen_device_noalloc (workspace in a fixed device array) against
hycom_hybgen_device, bitwise, through the C ABI: two calls in a row on the
same GPU, the second on a smaller domain, so the second call runs over
the workspace the first one left."""
import ctypes

import numpy as np
import pytest

import hybgen_abi as abi
import oracle_lib as ol

pytestmark = pytest.mark.gpu


def _pair(P, w, seed, raflag=0):
    """The same input through both entry points; returns both results."""
    import torch
    h0 = abi.HostDomain(P).synth(abi.make_synth(P, w, seed=seed))
    if raflag:  # an old thickness near the new one, as the RA filter expects
        rng = np.random.default_rng(seed)
        h0.a["dpo"][...] = np.abs(h0.a["dp"] * (1.0 + 0.01 * rng.standard_normal(h0.a["dp"].shape)))
    a = abi.DeviceDomain(P).upload(h0)
    b = abi.DeviceDomain(P).upload(h0)
    a.hybgen(m=1, n=2, raflag=raflag)
    torch.cuda.synchronize()
    b.hybgen(m=1, n=2, raflag=raflag, noalloc=True)
    torch.cuda.synchronize()
    return a.download(), b.download()


CASES = [
    # (name, workload, overrides, seed): the first is the larger domain
    ("glbb-window-ppm", "glbb", dict(idm=300, jdm=200, hybmap=2), 21),
    ("cfg1-weno-smaller", "cfg1", dict(idm=90, jdm=57, kdm=22, hybmap=3), 22),
]


def test_noalloc_matches_device_call_twice(gpu):
    lib = abi.load()
    cap = lib.hycom_hybgen_noalloc_capacity()
    assert cap == 256 << 20
    for name, wl, over, seed in CASES:
        P, w = abi.workload_params(wl, **over)
        need = lib.hycom_hybgen_noalloc_need(ctypes.byref(P), 0)
        print(f"\n[case] noalloc {name} {over} seed={seed} workspace={need} B", flush=True)
        assert 0 < need <= cap
        want, got = _pair(P, w, seed)
        df = ol.diff_domains(want, got)
        assert not df, f"{name}: noalloc differs from hycom_hybgen_device: {df}"


def test_noalloc_raflag_and_lane_family(gpu):
    """The RA branch and HYBGEN_DENFIX (per-lane-array kernels, which take
    no workspace) through the same entry point."""
    for name, over, raflag in [("cfg1-raflag", dict(idm=64, jdm=48, hybmap=2), 1),
                               ("cfg1-denfix", dict(idm=64, jdm=48, hybmap=2, ldenfix=1), 0)]:
        P, w = abi.workload_params("cfg1", **over)
        print(f"\n[case] noalloc {name} {over} raflag={raflag}", flush=True)
        want, got = _pair(P, w, 23, raflag=raflag)
        df = ol.diff_domains(want, got)
        assert not df, f"{name}: noalloc differs from hycom_hybgen_device: {df}"
