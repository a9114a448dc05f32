"""Stokes drift onto the model layers: stokes_vertical_j (mod_stokes.F90:
650-705) for every row, the loop of stokes_forfun (541-546).

Pinning, bottom up:
* the column remap (hybgen_plm_coefs on dpi, then hybgen_plm_remap from ki
  fixed layers onto ko model layers) of the oracle against the reference's
  own two subroutines, built by oracle/ref (oracle/_ref), bitwise;
* the oracle's whole call (blend of the snapshots, interfaces, remap per
  sea point) against the same reference column fed by numpy's IEEE
  elementwise blend and interfaces.  mod_stokes.F90 keeps sdzi and usdz
  private, so the reference's stokes_vertical_j itself cannot be driven
  from outside; its six lines around the two calls are these numpy lines;
* the library's column code (walk_core with ki != ko, compiled for the
  host) against the oracle, bitwise, on domains whose fixed layers end
  above, at and below the bottom.
The GPU path is checked against the oracle in test_gpu_stokes.py."""
import ctypes

import numpy as np
import pytest

import hybgen_abi as abi
import hostcheck_lib as hc
import oracle_lib as ol

D = ctypes.POINTER(ctypes.c_double)
ONEM = abi.ONEM
THIN = abi.DPTHIN


def _column(rng, ki, ko):
    pbot = rng.uniform(2.0, 400.0) * ONEM
    z = np.sort(rng.uniform(0.0, 300.0, ki)) * ONEM
    z[rng.random(ki) < 0.15] = 0.0          # repeated interfaces: empty layers
    sdzi = np.sort(z)
    sdzi[0] = THIN
    pi = np.concatenate([[0.0], np.minimum(sdzi, pbot)])
    dp = rng.exponential(pbot / ko, ko) * (rng.random(ko) > 0.2)
    dp *= pbot * rng.choice([0.97, 1.0, 1.0, 1.02]) / max(dp.sum(), 1.0)
    po = np.zeros(ko + 1)
    for k in range(ko):
        po[k + 1] = min(po[k] + dp[k], pbot)
    si = 0.2 * rng.standard_normal((2, ki)) + np.linspace(0.3, 0.0, ki)
    return si, pi, po


def _oracle_col(ki, ko, si, pi, po):
    so = np.zeros((2, ko))
    ol.oracle().oracle_plm_dpi_column(ki, ko, 2, np.ascontiguousarray(si).ctypes.data_as(D),
                                      pi.ctypes.data_as(D), po.ctypes.data_as(D), THIN,
                                      so.ctypes.data_as(D))
    return so


def _ref_col(ki, ko, si, pi, po):
    so = np.zeros((2, ko))
    ol.ref().href_plm_dpi_column(ki, ko, 2, np.ascontiguousarray(si).ctypes.data_as(D),
                                 np.ascontiguousarray(pi).ctypes.data_as(D),
                                 np.ascontiguousarray(po).ctypes.data_as(D), THIN,
                                 so.ctypes.data_as(D))
    return so


needs_ref = pytest.mark.skipif(not ol.ref_available(), reason="reference build absent")


@needs_ref
@pytest.mark.parametrize("ki,ko", [(2, 5), (8, 41), (21, 32), (41, 41), (64, 22), (30, 64),
                                   (12, 3)])
def test_oracle_column_matches_reference(ki, ko):
    rng = np.random.default_rng(ki * 131 + ko)
    for _ in range(200):
        si, pi, po = _column(rng, ki, ko)
        assert np.array_equal(_oracle_col(ki, ko, si, pi, po), _ref_col(ki, ko, si, pi, po))


def _domain(kdm, nsdzi, seed, depth_max=5500.0, n=2, ws0=0.3):
    P, w = abi.workload_params("cfg1", idm=37, jdm=23, kdm=kdm, ntracr=0, nsigma=min(8, kdm))
    S = abi.make_synth(P, w, seed=seed)
    S.depth_max = depth_max
    S.depth_min = min(S.depth_min, 0.5 * depth_max)
    dom = abi.HostDomain(P).synth(S, m=3 - n, n=n)
    return abi.StokesDomain(dom, nsdzi, seed=seed, ws0=ws0, n=n)


CASES = [(32, 12, 5500.0), (41, 41, 5500.0), (22, 64, 40.0), (64, 8, 5500.0),
         (5, 30, 200.0), (41, 2, 5500.0), (17, 24, 60.0)]


@needs_ref
def test_oracle_call_matches_reference_columns():
    sd = _domain(32, 16, seed=3, depth_max=150.0)
    ol.run_oracle_stokes(sd)
    a, P = sd.dom.a, sd.P
    nb, kk, nz = P.nbdy, P.kdm, P.nsdzi
    checked = 0
    for j in range(1, P.jj + 1):
        for i in range(1, P.ii + 1):
            I, J = i - 1 + nb, j - 1 + nb
            if a["ip"][I, J] == 0:
                continue
            pbot = a["p"][I, J, kk]
            u = P.ws0 * sd.usdz[I, J, 0, :] + P.ws1 * sd.usdz[I, J, 1, :]
            v = P.ws0 * sd.vsdz[I, J, 0, :] + P.ws1 * sd.vsdz[I, J, 1, :]
            pi = np.concatenate([[0.0], np.minimum(sd.sdzi, pbot)])
            po = np.zeros(kk + 1)
            for k in range(kk):
                po[k + 1] = min(po[k] + a["dp"][I, J, k, sd.n - 1], pbot)
            want = _ref_col(nz, kk, np.stack([u, v]), pi, po)
            assert np.array_equal(sd.usdp[I, J, :], want[0]), (i, j)
            assert np.array_equal(sd.vsdp[I, J, :], want[1]), (i, j)
            checked += 1
    assert checked > 100


@pytest.mark.parametrize("kdm,nsdzi,depth_max", CASES)
@pytest.mark.parametrize("n", [1, 2])
def test_library_column_code_matches_oracle(kdm, nsdzi, depth_max, n):
    sd = _domain(kdm, nsdzi, seed=kdm + nsdzi + n, depth_max=depth_max, n=n)
    a, b = sd.copy(), sd.copy()
    ol.run_oracle_stokes(a)
    F = b.fields()
    assert hc.lib().hostcheck_stokes(ctypes.byref(b.P), ctypes.byref(F), b.n) == 0
    assert np.array_equal(a.usdp, b.usdp)
    assert np.array_equal(a.vsdp, b.vsdp)
    land = sd.dom.a["ip"] == 0
    assert np.all(b.usdp[land] == -7.0) and np.all(b.vsdp[land] == 7.0)


def test_library_column_code_short_model_column():
    """Model layers that end above the bottom (sum of dp(n) < p(kk+1)): the
    last output ends before the walk reaches the deepest fixed levels, and
    those still count in the column range that clamps the outputs
    (hybgen_plm_remap's si_min/si_max run over all ki levels)."""
    sd = _domain(22, 40, seed=11, depth_max=300.0)
    rng = np.random.default_rng(5)
    dp = sd.dom.a["dp"]
    dp[:, :, :, sd.n - 1] *= rng.uniform(0.5, 0.95, dp.shape[:2])[:, :, None]
    sd.usdz[:, :, :, -3:] = rng.choice([-3.0, 3.0], sd.usdz[:, :, :, -3:].shape)
    sd.vsdz[:, :, :, -3:] = rng.choice([-3.0, 3.0], sd.vsdz[:, :, :, -3:].shape)
    a, b = sd.copy(), sd.copy()
    ol.run_oracle_stokes(a)
    F = b.fields()
    assert hc.lib().hostcheck_stokes(ctypes.byref(b.P), ctypes.byref(F), b.n) == 0
    assert np.array_equal(a.usdp, b.usdp)
    assert np.array_equal(a.vsdp, b.vsdp)


def test_abi_rejects_bad_arguments():
    """Argument errors come back before any device work (no GPU needed)."""
    lib = abi.load()
    sd = _domain(22, 8, seed=1)
    F = sd.fields()

    def call(P, n=2):
        return lib.hycom_stokes_vertical(ctypes.byref(P), ctypes.byref(F), n)

    for field, value in (("nsdzi", 1), ("nsdzi", 65), ("kdm", 0), ("reserved", 1)):
        P = abi.HybStokesParams.from_buffer_copy(sd.P)
        setattr(P, field, value)
        assert call(P) == abi.HYB_EINVAL, field
    assert call(sd.P, n=3) == abi.HYB_EINVAL
    sd.sdzi[3] = sd.sdzi[2] - 1.0  # the reference stops here (mod_stokes.F90:346-356)
    assert call(sd.P) == abi.HYB_EINVAL
    assert b"non-decreasing" in lib.hycom_hybgen_last_error()
