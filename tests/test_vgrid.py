"""hycom_hybgen_derive_vgrid against a brute-force evaluation of the
reference's own set-up statements, bitwise, over a parameter sweep:

  * blkdat.F90:296-301 -- a hybrid run with nsigma <= 1 uses nsigma = 1;
  * geopar.F90:449-482 -- dx0k(1) = onem*dx00, then dx0kf = dx0kf*dx00f
    and dx0k(k) = min(dx00*dx0kf, dx00x) (or the given dx0k times onem);
  * geopar.F90:520-558 -- dp0k(1) = onem*dp00, dp0kf = dp0kf*dp00f,
    dp0k(k) = min(dp00*dp0kf, dp00x) for dp00f >= 1 (else max) up to
    nhybrd, 0 below;
  * geopar.F90:567-600 -- ds0k the same way on ds00, ds00f, ds00x up to
    nsigma;
  * geopar.F90:610-630 -- dpns and dsns the sums of dp0k and ds0k over
    k = 1..nsigma, times qonem; ds0k = 0 below nsigma;
  * blkdat.F90:954, 986 and mod_hycom.F90:1447 -- qhybrlx = 1/hybrlx,
    isopcm = hybiso > 0, topiso = onem*isotop.

Each statement is evaluated as written, one IEEE double operation at a time
(Python floats), so any change of order, any dropped term or any
off-by-one in the k ranges shows up as a changed bit."""
import ctypes

import numpy as np
import pytest

import hybgen_abi as abi
import oracle_lib as ol

G, RHOREF = 9.806, 1000.0


def onem():
    svref = 1.0 / RHOREF        # mod_cb_arrays.F90:843-846
    return G / svref


def brute(kdm, nhybrd, nsigma, dp00, dp00f, dp00x, ds00, ds00f, ds00x, dp00i,
          dx00, dx00f, dx00x, dx0k_in, hybrlx, hybiso, isotop):
    om = onem()
    qonem = 1.0 / om
    if nsigma <= 1:                               # blkdat.F90:297-298
        nsigma = 1
    dx0k = [0.0] * kdm
    if dx00 > 0.0:                                # geopar.F90:450-473
        x0, xx = om * dx00, om * dx00x
        dx0k[0] = x0
        f = 1.0
        for k in range(2, kdm + 1):
            f = f * dx00f
            dx0k[k - 1] = min(x0 * f, xx)
    else:                                         # geopar.F90:474-477
        for k in range(1, kdm + 1):
            dx0k[k - 1] = om * dx0k_in[k - 1]
    dp0k = [0.0] * kdm
    d0, dx = om * dp00, om * dp00x                # geopar.F90:520-522
    dp0k[0] = d0
    f = 1.0
    for k in range(2, kdm + 1):
        f = f * dp00f
        if k <= nhybrd:
            dp0k[k - 1] = min(d0 * f, dx) if dp00f >= 1.0 else max(d0 * f, dx)
        else:
            dp0k[k - 1] = 0.0
    ds0k = [0.0] * kdm
    s0, sx = om * ds00, om * ds00x                # geopar.F90:567-569
    ds0k[0] = s0
    f = 1.0
    for k in range(2, nsigma + 1):
        f = f * ds00f
        ds0k[k - 1] = min(s0 * f, sx) if ds00f >= 1.0 else max(s0 * f, sx)
    dpns = dsns = 0.0                             # geopar.F90:617-627
    for k in range(1, nsigma + 1):
        dpns = dpns + dp0k[k - 1]
        dsns = dsns + ds0k[k - 1]
    for k in range(nsigma + 1, kdm + 1):
        ds0k[k - 1] = 0.0
    return dict(nsigma=nsigma, dx0k=dx0k, dp0k=dp0k, ds0k=ds0k, dpns=dpns * qonem,
                dsns=dsns * qonem, dp00i=om * dp00i, qhybrlx=1.0 / hybrlx,
                isopcm=int(hybiso > 0.0), topiso=om * isotop)


def _library(kdm, nhybrd, nsigma, dp00, dp00f, dp00x, ds00, ds00f, ds00x, dp00i,
             dx00, dx00f, dx00x, dx0k_in, hybrlx, hybiso, isotop):
    P = abi.HybParams()
    P.kdm, P.nhybrd, P.nsigma, P.hybiso = kdm, nhybrd, nsigma, hybiso
    for k in range(kdm):
        P.dx0k[k] = dx0k_in[k]
    top = ctypes.c_double()
    abi.check(abi.load().hycom_hybgen_derive_vgrid(
        ctypes.byref(P), dp00, dp00f, dp00x, ds00, ds00f, ds00x, dp00i, dx00, dx00f, dx00x,
        hybrlx, isotop, ctypes.byref(top)))
    return dict(nsigma=P.nsigma, dx0k=list(P.dx0k[:kdm]), dp0k=list(P.dp0k[:kdm]),
                ds0k=list(P.ds0k[:kdm]), dpns=P.dpns, dsns=P.dsns, dp00i=P.dp00i,
                qhybrlx=P.qhybrlx, isopcm=P.isopcm, topiso=top.value)


def _sweep():
    rng = np.random.default_rng(606)
    for q in range(240):
        kdm = int(rng.choice([4, 5, 17, 22, 32, 41, 50, 64]))
        nhybrd = kdm if rng.random() < 0.6 else int(rng.integers(1, kdm + 1))
        nsigma = int(rng.choice([0, 1, 2, min(8, nhybrd), min(14, nhybrd), nhybrd]))
        dpf = float(rng.choice([1.0, 1.125, 1.18, 1.3, 0.92]))
        dsf = float(rng.choice([1.0, 1.125, 1.18, 0.85]))
        dxm = int(rng.integers(0, 3))
        dx00, dx00f, dx00x = [(9999.0, 1.0, 9999.0), (30.0, 1.1, 300.0), (-1.0, 1.0, 0.0)][dxm]
        dx0k_in = list(rng.uniform(20.0, 400.0, kdm)) if dxm == 2 else [0.0] * kdm
        yield dict(kdm=kdm, nhybrd=nhybrd, nsigma=nsigma,
                   dp00=float(rng.choice([1.0, 3.0, 5.0, 0.7])), dp00f=dpf,
                   dp00x=float(rng.choice([2.0, 12.0, 450.0])),
                   ds00=float(rng.choice([0.5, 1.0, 3.0])), ds00f=dsf,
                   ds00x=float(rng.choice([0.2, 12.0, 450.0])),
                   dp00i=float(rng.choice([0.1, 1.0, 3.0])), dx00=dx00, dx00f=dx00f,
                   dx00x=dx00x, dx0k_in=dx0k_in, hybrlx=float(rng.choice([1.0, 3.0, 8.0])),
                   hybiso=float(rng.choice([0.0, 0.01, 3.0])),
                   isotop=float(rng.choice([6.0, 25.0, -1.0])))


@pytest.mark.parametrize("case", list(_sweep()), ids=lambda c: f"k{c['kdm']}h{c['nhybrd']}s{c['nsigma']}")
def test_derive_vgrid_is_the_reference_setup(case):
    want = brute(**case)
    got = _library(**case)
    for key, v in want.items():
        assert got[key] == v, f"{key}: {got[key]} != {v}"


def test_onem_is_the_reference_constant():
    """The brute force's onem is the reference's own (href_consts)."""
    if not ol.ref_available():
        pytest.skip("reference build absent")
    c = (ctypes.c_double * 16)()  # href_consts writes 16 values
    ol.ref().href_consts(c)
    assert c[0] == onem()


def test_sweep_reaches_every_branch():
    cases = list(_sweep())
    assert any(c["dp00f"] < 1.0 for c in cases) and any(c["ds00f"] < 1.0 for c in cases)
    assert any(c["nhybrd"] < c["kdm"] for c in cases)
    assert any(c["nsigma"] <= 1 for c in cases) and any(c["nsigma"] > 1 for c in cases)
    assert {c["dx00"] > 0 for c in cases} == {True, False}
    # the caps bind somewhere (min and max both change a value)
    assert any(min(brute(**c)["dp0k"][1:c["nhybrd"]] or [0]) > 0 and
               brute(**c)["dp0k"][c["nhybrd"] - 1] == onem() * c["dp00x"]
               for c in cases if c["nhybrd"] > 3)
