"""Model start from a restart (hycom_grid_masks, hycom_state_start,
hycom_restart_read_field; geopar.F90:229-256, bigrid.F90:1-330,
dpthuv.F90:48-123, mod_hycom.F90:1684-1722, dpudpv.F90:85-100).

Pinning:
* a hand-worked closed basin: every mask value, depthu/depthv and the
  halo rule written out in the test;
* the reference's own bigrid, dpthuv, xctilr and dpudpv (oracle/_ref,
  href_start) on seeded closed, periodic and global (ARCTIC build) domains:
  every output bitwise, halos included;
* the cases where bigrid stops, which the library reports as errors.
"""
import ctypes

import numpy as np
import pytest

import hybgen_abi as abi
import oracle_lib as ol

NB = 6
VOID = 2.0 ** 100


def _params(idm, jdm, kdm, nreg=0):
    P, _ = abi.workload_params("cfg1", idm=idm, jdm=jdm, kdm=kdm, ntracr=0, nsigma=min(8, kdm))
    P.nreg = nreg
    return P


def plane(idm, jdm, fill, dtype=np.float64, *extra):
    return np.asfortranarray(np.full((idm + 2 * NB, jdm + 2 * NB) + extra, fill, dtype=dtype))


def lib_masks(depth_interior):
    idm, jdm = depth_interior.shape
    dep = plane(idm, jdm, -7.0)
    dep[NB:NB + idm, NB:NB + jdm] = depth_interior
    ip, iu, iv = (plane(idm, jdm, -1, np.int32) for _ in range(3))
    nreg = ctypes.c_int32(-5)
    rc = abi.load().hycom_grid_masks(idm, jdm, NB, dep.ctypes.data, ip.ctypes.data,
                                     iu.ctypes.data, iv.ctypes.data, ctypes.byref(nreg))
    return rc, dep, ip, iu, iv, nreg.value


def lib_start(P, pbot, dp, ip, iu, iv):
    idm, jdm, kdm = P.idm, P.jdm, P.kdm
    out = {"depthu": plane(idm, jdm, -9.0), "depthv": plane(idm, jdm, -9.0),
           "p": plane(idm, jdm, -9.0, np.float64, kdm + 1),
           "dpu": plane(idm, jdm, -9.0, np.float64, kdm, 2),
           "dpv": plane(idm, jdm, -9.0, np.float64, kdm, 2)}
    F = abi.HybStartFields(pbot.ctypes.data, dp.ctypes.data, ip.ctypes.data, iu.ctypes.data,
                           iv.ctypes.data, *[out[n].ctypes.data for n in
                                             ("depthu", "depthv", "p", "dpu", "dpv")])
    rc = abi.load().hycom_state_start(ctypes.byref(P), ctypes.byref(F))
    return rc, out


def test_hand_worked_closed_basin():
    # depth(i,j), i = 1..5 across, j = 1..4; (3,3) is a data void
    d = np.zeros((5, 4))
    d[2, 0] = 50.0
    d[1:4, 1] = [40.0, 60.0, 70.0]
    d[1, 2], d[2, 2], d[3, 2] = 30.0, VOID, 80.0
    rc, dep, ip, iu, iv, nreg = lib_masks(d)
    assert rc == 0 and nreg == 0
    I = (slice(NB, NB + 5), slice(NB, NB + 4))
    want_ip = np.zeros((5, 4), np.int32)
    for i, j in [(3, 1), (2, 2), (3, 2), (4, 2), (2, 3), (4, 3)]:
        want_ip[i - 1, j - 1] = 1
    want_iu = np.zeros((5, 4), np.int32)
    want_iu[2, 1] = want_iu[3, 1] = 1                 # u(3,2), u(4,2)
    want_iv = np.zeros((5, 4), np.int32)
    for i, j in [(3, 2), (2, 3), (4, 3)]:
        want_iv[i - 1, j - 1] = 1
    assert np.array_equal(ip[I], want_ip)
    assert np.array_equal(iu[I], want_iu)
    assert np.array_equal(iv[I], want_iv)
    # closed basin: the whole halo is land, and so is the void
    for a in (ip, iu, iv):
        h = a.copy()
        h[I] = 0
        assert not h.any()
    assert dep[NB + 2, NB + 2] == 0.0 and not dep[:NB].any() and not dep[:, NB + 4:].any()

    P = _params(5, 4, 5)
    P.vland = 1.0
    pbot = plane(5, 4, -3.0)
    pbot[I] = dep[I] * 9806.0
    dp = plane(5, 4, 0.0, np.float64, 5, 2)
    dp[I] = np.where(ip[I][:, :, None, None] > 0, 1000.0, 0.0)
    rc, out = lib_start(P, pbot, dp, ip, iu, iv)
    assert rc == 0
    pb = lambda i, j: dep[NB + i - 1, NB + j - 1] * 9806.0  # noqa: E731
    du, dv = out["depthu"][I], out["depthv"][I]
    want_du = np.zeros((5, 4))
    want_du[2, 1], want_du[3, 1] = min(pb(3, 2), pb(2, 2)), min(pb(4, 2), pb(3, 2))
    want_dv = np.zeros((5, 4))
    want_dv[2, 1] = min(pb(3, 2), pb(3, 1))
    want_dv[1, 2] = min(pb(2, 3), pb(2, 2))
    want_dv[3, 2] = min(pb(4, 3), pb(4, 2))
    assert np.array_equal(du, want_du) and np.array_equal(dv, want_dv)
    # beyond a closed boundary xctilr writes vland
    assert out["depthu"][NB - 1, NB + 1] == 1.0 and out["depthv"][NB + 2, NB + 4] == 1.0
    # p at sea: the running sum; dpu at u(3,2): each 1000-Pa layer under
    # depthu (60 m and 40 m of water, so min(depthu, ...) never binds)
    assert np.array_equal(out["p"][NB + 2, NB + 1], np.arange(6) * 1000.0)
    assert np.array_equal(out["dpu"][NB + 2, NB + 1], np.full((5, 2), 1000.0))
    assert out["dpu"][NB + 1, NB + 1].sum() == 0.0  # u(2,2) is land


def test_where_bigrid_stops():
    d = np.zeros((5, 4))
    d[2, 1] = 100.0  # one-point sea
    rc, *_ = lib_masks(d)
    assert rc == abi.HYB_EINVAL and b"4 land neighbours" in abi.load().hycom_hybgen_last_error()
    d = np.zeros((5, 4))
    d[1:4, 1:4] = 100.0  # sea in the last row, not periodic
    rc, *_ = lib_masks(d)
    assert rc == abi.HYB_EINVAL and b"non-periodic" in abi.load().hycom_hybgen_last_error()


def bathymetry(rng, idm, jdm, kind):
    """a seeded basin with land: closed, periodic in i (sea in the last
    column) or global (sea in the last column and the last row); sea points
    with four land neighbours are removed, since bigrid stops on them"""
    x = np.linspace(0, 1, idm)[:, None]
    y = np.linspace(0, 1, jdm)[None, :]
    d = 3000 * np.sin(3 * x + 2 * y + rng.uniform(0, 6)) + 1500 * rng.standard_normal((idm, jdm))
    d = np.where(d > 200, d, 0.0)
    if kind == "closed":
        d[-1, :] = 0.0
        d[:, -1] = 0.0
    elif kind == "periodic":
        d[:, -1] = 0.0
        d[-1, :jdm // 2] = 3000.0
    else:
        d[-1, :jdm // 2] = 3000.0
        d[:idm // 2, -1] = 3000.0
    for _ in range(20):
        pad = np.pad(d, 1)
        if kind != "closed":
            pad[0, 1:-1], pad[-1, 1:-1] = d[-1], d[0]
        nz = ((pad[:-2, 1:-1] <= 0).astype(int) + (pad[2:, 1:-1] <= 0) + (pad[1:-1, :-2] <= 0)
              + (pad[1:-1, 2:] <= 0))
        bad = (d > 0) & (nz == 4)
        if not bad.any():
            break
        d[bad] = 0.0
    return d


KINDS = [("closed", 0), ("periodic", 1), ("arctic", 2)]


@pytest.mark.parametrize("kind,nreg", KINDS)
@pytest.mark.parametrize("seed", [1, 2])
def test_start_matches_reference(kind, nreg, seed):
    arctic = kind == "arctic"
    if not ol.start_ref_available(arctic):
        pytest.skip("reference build absent")
    rng = np.random.default_rng(seed)
    idm, jdm, kdm = 23 + seed, 17, 5
    d = bathymetry(rng, idm, jdm, kind)
    if d[2, 3] == 0.0:
        d[2, 3] = VOID  # a void on land reads as land
    rc, dep, ip, iu, iv, got_nreg = lib_masks(d)
    assert rc == 0 and got_nreg == nreg
    P = _params(idm, jdm, kdm, got_nreg)
    P.vland = 1.0
    I = (slice(NB, NB + idm), slice(NB, NB + jdm))
    pbot = plane(idm, jdm, -3.0)
    pbot[I] = dep[I] * 9806.0
    dp = plane(idm, jdm, -5.0, np.float64, kdm, 2)
    dp[I] = rng.uniform(0, 1e5, (idm, jdm, kdm, 2))
    pb_in, dp_in = pbot.copy(order="F"), dp.copy(order="F")
    rc, out = lib_start(P, pbot, dp, ip, iu, iv)
    assert rc == 0

    R = ol.start_ref(arctic)
    rdep = plane(idm, jdm, -7.0)
    rdep[I] = d
    rip, riu, riv = (plane(idm, jdm, 0, np.int32) for _ in range(3))
    rn = ctypes.c_int32()
    ro = {n: np.asfortranarray(np.full(v.shape, -9.0)) for n, v in out.items()}
    rc = R.href_start(ctypes.byref(P), rdep.ctypes.data, pb_in.ctypes.data, dp_in.ctypes.data,
                      rip.ctypes.data, riu.ctypes.data, riv.ctypes.data, ctypes.byref(rn),
                      *[ro[n].ctypes.data for n in ("depthu", "depthv", "p", "dpu", "dpv")])
    assert rc == 0 and rn.value == nreg
    assert np.array_equal(dep, rdep)
    assert np.array_equal(ip, rip) and np.array_equal(iu, riu) and np.array_equal(iv, riv)
    assert np.array_equal(pbot, pb_in) and np.array_equal(dp, dp_in)  # halos as xctilr fills them
    for n in ("depthu", "depthv", "dpu", "dpv"):
        assert np.array_equal(out[n], ro[n]), n
    sea = (ip != 0)[:, :, None]
    assert np.array_equal(np.where(sea, out["p"], 0.0), np.where(sea, ro["p"], 0.0))
    assert np.array_equal(out["p"][:, :, 0], ro["p"][:, :, 0])
    # the halos carry sea where the domain wraps or folds
    halo_i = np.concatenate([ip[:NB, NB:NB + jdm], ip[NB + idm:, NB:NB + jdm]])
    assert halo_i.any() == (nreg > 0)
    assert ip[NB:NB + idm, NB + jdm:].any() == (nreg == 2)
    assert iu[NB:NB + idm, NB + jdm:].any() == (nreg == 2)
    assert out["dpu"].any() and out["dpv"].any()


def test_restart_read_field(tmp_path):
    from test_za import IDM, JDM, write_restart
    P, _ = abi.workload_params("cfg1", idm=IDM, jdm=JDM, kdm=9, ntracr=0, nsigma=8)
    rng = np.random.default_rng(8)
    NI, NJ = IDM + 2 * NB, JDM + 2 * NB
    pb = np.asfortranarray(rng.uniform(1e6, 5e7, (NI, NJ, 1)))
    pbavg = np.asfortranarray(rng.uniform(0, 1, (NI, NJ, 3)))
    base = str(tmp_path / "r")
    write_restart(base, P, [pbavg, pb], ["pbavg", "pbot"])
    out = plane(IDM, JDM, -1.0)
    lib = abi.load()
    rc = lib.hycom_restart_read_field((base + ".a").encode(), (base + ".b").encode(),
                                      ctypes.byref(P), b"pbot", 1, out.ctypes.data)
    assert rc == 0, lib.hycom_hybgen_last_error()
    I = (slice(NB, NB + IDM), slice(NB, NB + JDM))
    assert np.array_equal(out[I], pb[I][:, :, 0].astype(np.float32).astype(np.float64))
    assert np.all(out[:NB] == -1.0)
    rc = lib.hycom_restart_read_field((base + ".a").encode(), (base + ".b").encode(),
                                      ctypes.byref(P), b"pbot", 2, out.ctypes.data)
    assert rc == abi.HYB_EINVAL


def closed_depths(d):
    """a closed basin from a bathymetry: last column and row land, then no
    sea point with four land neighbours (bigrid stops there)"""
    d = d.copy()
    d[-1, :] = 0.0
    d[:, -1] = 0.0
    for _ in range(50):
        pad = np.pad(d, 1)
        nz = ((pad[:-2, 1:-1] <= 0).astype(int) + (pad[2:, 1:-1] <= 0) + (pad[1:-1, :-2] <= 0)
              + (pad[1:-1, 2:] <= 0))
        bad = (d > 0) & (nz == 4)
        if not bad.any():
            return d
        d[bad] = 0.0
    raise AssertionError("bathymetry did not settle")


def load_restart_state(tmp_path, P, d0):
    """Writes the bathymetry of the synthetic domain d0 as an .a file and
    its state as a restart pair (u, v, dp, temp, saln, ubavg, vbavg, pbavg,
    pbot, tracer: 32-bit words, as restart_out writes them), then builds a
    hybgen state the way a model start does: .a read, grid masks, restart
    read, pbot, state start.  theta and topiso, which come from the
    experiment's set-up and not from these files, are taken from d0."""
    from test_za import write_a, write_restart
    lib = abi.load()
    idm, jdm, K = P.idm, P.jdm, P.kdm
    I = (slice(NB, NB + idm), slice(NB, NB + jdm))
    dep_a = str(tmp_path / "depth.a")
    write_a(dep_a, [closed_depths(d0.a["depths"][I])])
    s = abi.HostDomain(P)
    abi.check(lib.hycom_za_read(dep_a.encode(), idm, jdm, NB, 1, 1, None,
                                s.a["depths"].ctypes.data, None, None))
    nreg = ctypes.c_int32()
    abi.check(lib.hycom_grid_masks(idm, jdm, NB, s.a["depths"].ctypes.data,
                                   s.a["ip"].ctypes.data, s.a["iu"].ctypes.data,
                                   s.a["iv"].ctypes.data, ctypes.byref(nreg)))
    P.nreg = nreg.value
    NI, NJ = idm + 2 * NB, jdm + 2 * NB

    def lv(name, n):
        return d0.a[name].reshape(NI, NJ, n, order="F")

    pbot = np.asfortranarray((d0.a["depths"] * 9806.0)[:, :, None])
    pbavg = np.asfortranarray(np.zeros((NI, NJ, 3)))
    ntr = P.ntracr
    names = ["u", "v", "dp", "temp", "saln", "ubavg", "vbavg", "pbavg", "pbot"]
    fields = [lv("u", 2 * K), lv("v", 2 * K), lv("dp", 2 * K), lv("temp", 2 * K),
              lv("saln", 2 * K), lv("ubavg", 3), lv("vbavg", 3), pbavg, pbot]
    if ntr:
        names.append("tracer")
        fields.append(lv("tracer", 2 * K * ntr))
    base = str(tmp_path / "restart_in")
    write_restart(base, P, fields, names)
    F = s.fields()
    abi.check(lib.hycom_restart_read((base + ".a").encode(), (base + ".b").encode(),
                                     ctypes.byref(P), ctypes.byref(F), None, None))
    pb = plane(idm, jdm, 0.0)
    abi.check(lib.hycom_restart_read_field((base + ".a").encode(), (base + ".b").encode(),
                                           ctypes.byref(P), b"pbot", 1, pb.ctypes.data))
    SF = abi.HybStartFields(pb.ctypes.data, *[s.a[n].ctypes.data for n in
                                              ("dp", "ip", "iu", "iv", "depthu", "depthv",
                                               "p", "dpu", "dpv")])
    abi.check(lib.hycom_state_start(ctypes.byref(P), ctypes.byref(SF)))
    s.a["theta"][...] = d0.a["theta"]
    s.a["topiso"][...] = d0.a["topiso"]
    return s


def test_restart_state_through_reference_hybgen(tmp_path):
    """The state of a model start from files, stepped by the oracle and by
    the reference's hybgen: bitwise the same (the GPU test steps the same
    state on the device)."""
    if not ol.ref_available():
        pytest.skip("reference build absent")
    P, w = abi.workload_params("cfg1", idm=60, jdm=44)
    P.ii, P.jj = P.idm, P.jdm
    d0 = abi.HostDomain(P).synth(abi.make_synth(P, w, seed=300))
    s = load_restart_state(tmp_path, P, d0)
    I = (slice(NB, NB + P.idm), slice(NB, NB + P.jdm))
    sea = s.a["ip"][I] != 0
    assert sea.sum() > 0.5 * sea.size
    # the loaded thicknesses are the synthetic ones through 32-bit words
    got = s.a["dp"][I][sea]
    assert np.array_equal(got, d0.a["dp"][I][sea].astype(np.float32).astype(np.float64))
    a, b = s.copy(), s.copy()
    ol.run_oracle(a, threads=8)
    ol.run_ref(b)
    df = ol.diff_domains(a, b)
    assert not df, df
    assert not np.array_equal(a.a["dp"], s.a["dp"])  # hybgen moved interfaces
