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
* the oracle's whole call against a closed form that restates none of
  those lines: with snapshots that zigzag in depth every PLM slope is zero
  and each model layer is the exact mean of the blended profile over it;
  the test shows that an interface off by one, swapped weights and the
  other time level's dp each fail the same check;
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


def _overlap_means(s, zi, zo):
    """Thickness-weighted mean of the piecewise-constant profile s (value
    s[l] on [zi[l], zi[l+1]]) over each layer [zo[k], zo[k+1]]; NaN for
    empty layers."""
    lo = np.maximum(zo[:-1, None], zi[None, :-1])
    hi = np.minimum(zo[1:, None], zi[None, 1:])
    w = np.maximum(hi - lo, 0.0)
    with np.errstate(invalid="ignore", divide="ignore"):
        return (w @ s) / (zo[1:] - zo[:-1])


@pytest.mark.parametrize("ws0,n", [(0.3, 2), (1.0, 1), (0.0, 2), (0.75, 1)])
def test_oracle_call_is_the_layer_mean_of_a_zigzag_profile(ws0, n):
    """A closed form of stokes_vertical_j (mod_stokes.F90:676-697) that needs
    neither the reference nor its code lines.  When both snapshots zigzag in
    depth (every fixed layer a strict extremum of its neighbours, after the
    blend too) the PLM limiter (hybgen.F90:2423-2436) gives every layer a
    zero slope, so each model layer's value is the exact mean, over the
    layer, of the time-interpolated profile: fixed layer l spans
    [sdzi(l-1), sdzi(l)] (sdzi(0) = 0), model layer k spans the running sum
    of dp(k, n), both cut at the sea floor p(kk+1).  Swapped blend weights
    or snapshots, u taken for v, an interface off by one or dp of the
    other time level each fail it."""
    sd = _domain(32, 16, seed=int(ws0 * 8) + n, depth_max=150.0, ws0=ws0, n=n)
    P, a, nb, kk = sd.P, sd.dom.a, sd.P.nbdy, sd.P.kdm
    rng = np.random.default_rng(17 + n)
    NI, NJ, _, nz = sd.usdz.shape
    sign = (-1.0) ** np.arange(nz)
    for arr in (sd.usdz, sd.vsdz):
        base = rng.uniform(-0.5, 0.5, (NI, NJ, 2, 1))
        amp = rng.uniform(0.01, 0.2, (NI, NJ, 2, 1)) * rng.uniform(0.5, 1.0, (NI, NJ, 2, nz))
        arr[...] = base + amp * sign
    # the other time level holds the same thicknesses upside down
    a["dp"][:, :, :, 2 - n] = a["dp"][:, :, ::-1, n - 1]
    ol.run_oracle_stokes(sd)
    assert sd.sdzi[-1] > a["p"][:, :, kk].max()       # the fixed layers reach every floor
    checked, caught = 0, [0, 0, 0]
    for j in range(1, P.jj + 1):
        for i in range(1, P.ii + 1):
            I, J = i - 1 + nb, j - 1 + nb
            if a["ip"][I, J] == 0:
                continue
            pbot = a["p"][I, J, kk]
            zi = np.concatenate([[0.0], np.minimum(sd.sdzi, pbot)])
            zo = np.minimum(np.concatenate([[0.0], np.cumsum(a["dp"][I, J, :, n - 1])]), pbot)
            thick = np.diff(zo) > 1.0e-3 * ONEM
            # the same closed form with a planted mistake, to show it discriminates
            zo_lvl = np.minimum(np.concatenate([[0.0], np.cumsum(a["dp"][I, J, :, 2 - n])]), pbot)
            zi_off = np.concatenate([[0.0], np.minimum(np.append(sd.sdzi[1:], sd.sdzi[-1]), pbot)])
            for snap, out in ((sd.usdz, sd.usdp), (sd.vsdz, sd.vsdp)):
                s = P.ws0 * snap[I, J, 0, :] + P.ws1 * snap[I, J, 1, :]
                want = _overlap_means(s, zi, zo)
                got = out[I, J, :]
                scale = np.abs(s).max()
                assert np.all(np.abs(got[thick] - want[thick]) <= 1e-12 * scale), (i, j)
                # thin layers lie within the range of the profile
                assert np.all(got >= s.min() - 1e-12 * scale) and np.all(got <= s.max() + 1e-12 * scale)
                wrong = [_overlap_means(s, zi_off, zo)]
                if ws0 != 0.5:
                    wrong.append(_overlap_means(P.ws1 * snap[I, J, 0, :] + P.ws0 * snap[I, J, 1, :],
                                                zi, zo))
                for q, w in enumerate(wrong):
                    caught[q] += bool(np.any(np.abs(got[thick] - w[thick]) > 1e-9 * scale))
                t2 = np.diff(zo_lvl) > 1.0e-3 * ONEM
                w = _overlap_means(s, zi, zo_lvl)
                caught[2] += bool(np.any(np.abs(got[t2] - w[t2]) > 1e-9 * scale)
                                  or not np.array_equal(t2, thick))
            checked += 1
    assert checked > 100
    # every planted mistake is seen in most columns (both fields counted)
    assert caught[0] > checked and caught[2] > checked
    assert ws0 == 0.5 or caught[1] > checked


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
