"""A small closed-basin case for the stand-alone HYCOM model: the input
files a run reads from its working directory.

  blkdat.input          the ~200 run parameters, in the order blkdat.F90
                        reads them (blkdat.F90:40-2548, one `value 'name'`
                        line each after four title lines);
  limits                start and end day (mod_hycom.F90:1208-1209); a
                        negative start day is a new start;
  regional.grid.[ab]    18 records plon..cori (geopar.F90:24-180);
  regional.depth.[ab]   one record, land as the 2^100 void
                        (geopar.F90:205-247);
  relax.{temp,saln,intf}.[ab] (+ relax.trcr with tracers)
                        a z-level climatology, 12 months x kdm records
                        (forfun.F90:2037-2098, rdmonthck 2680-2770), which
                        the model starts from with iniflg=2
                        (inicon.F90:70-92, 311-320);
  relax.weird           accept a climatology whose second and third
                        interfaces are not at dp0k (forfun.F90:3770-3790).

`.a` records are big-endian float32, i fastest, each padded to a multiple
of 4096 words with the void (mod_za_sm.h:497-498, machi_c.c zaio_endian);
each `.b` line's min and max are those of its record to 1e-4
(geopar.F90:170-179).

The start is a z-level state: layer k spans fixed depths, T and S follow
smooth profiles, so the first steps regrid every column hard toward the
isopycnal targets, and the dynamics keep the regridding busy afterwards.
No forcing files are needed (wind, thermal and relaxation forcing off)."""
import math
import os

import numpy as np

ONEM = 9.806 * 1000.0    # one metre in pressure units (mod_cb_arrays.F90:842-846)
VOID = 2.0 ** 100        # data void of .a files (mod_za_sm.h:60-62)
ZI = np.array([0, 5, 15, 30, 50, 80, 120, 170, 240, 330, 450, 600, 800, 1100, 1500,
               2000, 2600, 3200, 3800, 4400], dtype=np.float64)


def _recs(fa, fields, idm, jdm, mask=None):
    n = idm * jdm
    nrec = ((n + 4095) // 4096) * 4096
    out = []
    for a in fields:
        w = np.full(nrec, VOID, dtype=">f4")
        a = np.asarray(a, dtype=np.float64)
        if mask is not None:
            a = np.where(mask, a, VOID)
        w[:n] = a.astype(np.float32).reshape(-1)
        fa.write(w.tobytes())
        v = w[:n][w[:n] < VOID / 2]
        out.append((float(v.min()), float(v.max())))
    return out


def write_case(path, *, idm=40, jdm=30, kdm=16, hybmap=2, hybraf=False, ntracr=0,
               trcflg=3, days=1.0, periodic=False, nsigma=0, hybthk=0.0, shelf=20.0,
               diag=(), hybiso=0.01, hybrlx=8.0, dx00=None, res=0.5, baclin=900.0,
               batrop=30.0, tnudge=0):
    """Writes the case into directory `path` (created if missing).  The run
    starts at day 1 (step 96 with baclin 900 s) and ends `days` later.
    periodic: a channel periodic in i (sea in the last column makes the
    model take the domain as periodic, bigrid.F90:21-33), land rows at both
    j edges; otherwise a closed basin.  shelf: the shallowest depth, m
    (under the nsigma layers' deep limit dpns the columns are terrain
    following, geopar.F90:449-630).  diag: diagnostic tracer types (801
    hybgen's change of dp, 802 of T, 803 of S), kept after the ntracr
    tracers (mtracr, blkdat.F90:1307-1330).  hybiso 0 turns the PCM
    remap of near-isopycnal layers off (isopcm); dx00=(dx00, dx00x, dx00f)
    sets maximum layer thicknesses (the dx00 form of blkdat.F90:234-250).
    res: grid spacing in degrees (from 15N, 60W).  tnudge: the T climatology
    moved by that many float32 ulps everywhere (a perturbation at the
    precision of the input files, to measure how far the model's own
    sensitivity carries a tiny difference)."""
    assert 4 <= kdm <= 64 and idm >= 8 and jdm >= 8
    os.makedirs(path, exist_ok=True)
    P = lambda f: os.path.join(path, f)  # noqa: E731

    # grid: res degrees, 15N.., 60W..
    assert 15.0 + res * jdm < 80.0
    ii = np.arange(idm)[None, :] * np.ones((jdm, 1))
    jj = np.arange(jdm)[:, None] * np.ones((1, idm))
    plon, plat = -60.0 + res * ii, 15.0 + res * jj
    rad, R, h = math.pi / 180.0, 6371001.0, 0.5 * res
    scx = lambda lat: R * rad * res * np.cos(lat * rad)  # noqa: E731
    scy = lambda lat: R * rad * res * np.ones_like(lat)  # noqa: E731
    qlat, ulat, vlat = plat - h, plat, plat - h
    grid = [("plon", plon), ("plat", plat), ("qlon", plon - h), ("qlat", qlat),
            ("ulon", plon - h), ("ulat", ulat), ("vlon", plon), ("vlat", vlat),
            ("pang", np.zeros_like(plon)),
            ("pscx", scx(plat)), ("pscy", scy(plat)), ("qscx", scx(qlat)), ("qscy", scy(qlat)),
            ("uscx", scx(ulat)), ("uscy", scy(ulat)), ("vscx", scx(vlat)), ("vscy", scy(vlat)),
            ("cori", 2.0 * 7.2921e-5 * np.sin(qlat * rad))]
    with open(P("regional.grid.a"), "wb") as fa:
        rng = _recs(fa, [a for _, a in grid], idm, jdm)
    with open(P("regional.grid.b"), "w") as fb:
        fb.write("%5d    'idm   ' = longitudinal array size\n" % idm)
        fb.write("%5d    'jdm   ' = latitudinal  array size\n" % jdm)
        fb.write("    0    'mapflg' = map flag (0=mercator)\n")
        for (name, _), (lo, hi) in zip(grid, rng):
            fb.write("%s: min,max = %16.8e %16.8e\n" % (name, lo, hi))

    # a bowl of 4000 m with shelves as shallow as `shelf`, one land ring; or a
    # channel of the same section, periodic in i
    xi, yj = (ii + 0.5) / idm, (jj + 0.5) / jdm
    if periodic:
        r = np.abs(2 * yj - 1)
        bump = np.sin(2.0 * math.pi * 2.0 * xi) * np.cos(5.0 * yj)
    else:
        r = np.maximum(np.abs(2 * xi - 1), np.abs(2 * yj - 1))
        bump = np.sin(7.0 * xi) * np.cos(5.0 * yj)
    depth = shelf + (4000.0 - shelf) * np.clip((1.0 - r) / 0.35, 0.0, 1.0) ** 1.5
    depth += 150.0 * bump * (depth > 400.0)
    depth[0, :] = depth[-1, :] = 0.0
    if not periodic:
        depth[:, 0] = depth[:, -1] = 0.0
    sea = depth > 0.0
    with open(P("regional.depth.a"), "wb") as fa:
        (lo, hi), = _recs(fa, [depth], idm, jdm, mask=sea)
    with open(P("regional.depth.b"), "w") as fb:
        fb.write("closed basin\n\n\n\n\n")
        fb.write("depth: min,max = %16.8e %16.8e\n" % (lo, hi))

    with open(P("limits"), "w") as f:
        f.write("  -1.0  %.4f\n" % (1.0 + days))

    # layer target densities (sigma-2) and the z-level climatology
    sigma = np.round(28.10 + (36.70 - 28.10) * np.linspace(0.0, 1.0, kdm) ** 0.8, 4)
    # z-levels: ZI, or for more layers thicknesses growing 13% a layer
    # from 5 m (40 layers reach 5000 m)
    zi = ZI[:kdm] if kdm <= len(ZI) else np.concatenate(
        [[0.0], np.cumsum(5.0 * 1.13 ** np.arange(kdm - 1))])
    zmid = [0.5 * (zi[k] + min(zi[k + 1] if k + 1 < kdm else 1e5, zi[k] + 400.0))
            for k in range(kdm)]
    sst = 27.0 - 0.45 * (plat - 15.0) + 0.8 * np.sin(0.3 * plon)
    clim = {
        "temp": [4.0 + (sst - 4.0) * np.exp(-z / 500.0) - 2.0 * z / 5000.0 for z in zmid],
        "saln": [34.9 + 0.6 * np.exp(-z / 250.0)
                 - 0.1 * (plat - 22.0) / 8.0 * np.exp(-z / 800.0) for z in zmid],
        "intf": [np.minimum(zi[k], depth) * ONEM for k in range(kdm)],
    }
    if ntracr:
        # trcflg's decimal digits, least significant first, the last one
        # repeated (blkdat.F90:1288-1300); a temperature tracer (2) starts
        # from T's climatology, the others from their own profiles
        flags, t = [], trcflg
        for q in range(ntracr):
            flags.append(t % 10)
            t = t // 10 if t >= 10 else t
        clim["trcr"] = [clim["temp"][k] if flags[q] == 2 else
                        (10.0 + 3.0 * q) * np.exp(-zmid[k] / (150.0 + 200.0 * q))
                        * (1.0 + 0.3 * np.cos(0.2 * ii + q)) + 0.01 * (100.0 * yj + q)
                        for q in range(ntracr) for k in range(kdm)]
    if tnudge:
        clim["temp"] = [np.nextafter(np.asarray(t, np.float32),
                                     np.float32(np.inf) if tnudge > 0 else np.float32(-np.inf))
                        .astype(np.float64) for t in clim["temp"]]
    for name, fl in clim.items():
        with open(P(f"relax.{name}.a"), "wb") as fa, open(P(f"relax.{name}.b"), "w") as fb:
            fb.write(f"z-level {name} climatology\n\n\n\n\n")
            for mon in range(1, 13):
                for r_, (lo, hi) in enumerate(_recs(fa, fl, idm, jdm, mask=sea)):
                    k = r_ % kdm
                    fb.write(" %s: month,layer,dens,range = %2d %3d %7.3f %16.8e %16.8e\n"
                             % (name, mon, k + 1, sigma[k], lo, hi))
    open(P("relax.weird"), "w").close()

    L = []
    I = lambda n, v: L.append("%11d    '%-6s'" % (int(v), n))  # noqa: E731,E741
    F = lambda n, v: L.append("%11s    '%-6s'" % (repr(float(v)), n))  # noqa: E731
    I("iversn", 23); I("iexpt", 10); I("idm", idm); I("jdm", jdm)
    I("itest", -1); I("jtest", -1); I("kdm", kdm); I("nhybrd", kdm); I("nsigma", nsigma)
    # minimum thicknesses dp0k = dp00*dp00f**(k-1) up to dp00x; with many
    # layers a smaller start and cap keep their sum a few hundred metres
    dp00, dp00x = (3.0, 450.0) if kdm <= len(ZI) else (1.0, 12.0)
    if dx00:
        F("dx00", dx00[0]); F("dx00x", dx00[1]); F("dx00f", dx00[2])
    F("dp00", dp00); F("dp00x", dp00x); F("dp00f", 1.18)
    F("ds00", 1.0); F("ds00x", 15.0); F("ds00f", 1.18); F("dp00i", 1.0); F("isotop", 6.0)
    F("oneta0", 0.1); F("saln0", 35.0); I("locsig", 0); I("kapref", 0); I("thflag", 2)
    F("thbase", 34.0); I("vsigma", 0)
    for s in sigma:
        F("sigma", s)
    I("iniflg", 2); I("jerlv0", 1); I("yrflag", 3); I("sshflg", 0)
    F("dsurfq", 1.0); F("diagfq", days); F("proffq", 0.0); F("tilefq", 0.0)
    F("meanfq", 0.0); F("rstrfq", 365.0); F("bnstfq", 0.0); F("nestfq", 0.0)
    F("cplifq", -1.0); F("baclin", baclin); F("batrop", batrop)
    I("incflg", 0); I("incstp", 0); I("incupf", 0)
    I("stfflg", 0); F("stfrdt", 0.0); F("stfrds", 0.0); F("stfrdv", 0.0)
    F("ra2fac", 0.125); F("wbaro", 0.125); I("btrlfr", 1); I("btrmas", 0)
    I("hybraf", hybraf); F("hybrlx", hybrlx); F("hybiso", hybiso); F("hybthn", 1.0)
    F("hybthk", hybthk); I("hybmap", hybmap); I("hybflg", 0); I("advflg", 0)
    I("advtyp", 2); I("momtyp", 2); F("slip", 0.0); F("visco2", 0.05); F("visco4", 0.2)
    F("facdf4", 0.0); F("veldf2", 0.02); F("veldf4", 0.01); F("thkdf2", 0.0)
    F("thkdf4", 0.005); F("temdf2", 0.015); F("temdfc", 1.0); F("vertmx", 0.0)
    F("cbar", 0.05); F("cb", 2.2e-3); I("cbtidc", 0); F("drglim", 0.0)
    F("thkbot", 10.0); F("sigjmp", 0.02); F("tmljmp", 0.2); F("sefold", 15.0)
    F("thkmls", 0.0); F("thkmlt", 0.0); F("thkriv", 0.0); F("thkcdw", 3.0)
    F("thkfrz", 0.0); I("iceflg", 0); F("tfrz_0", 0.0); F("tfrz_s", -0.054)
    F("frzifq", -1.0); F("ticegr", 0.0); F("hicemn", 0.5); F("hicemx", 10.0)
    # a negative ntracr starts the tracers from relax.trcr (blkdat.F90:1272-1274)
    I("ishelf", 0); I("ntracr", -ntracr); I("trcflg", trcflg); I("mtracr", len(diag))
    for d in diag:
        I("dtrflg", d)
    I("mstrcr", 0); I("tsofrq", 0); F("tofset", 0.0); F("sofset", 0.0); I("mlflag", 0)
    I("pensol", 0); F("dtrate", 0.0); F("thkmin", 0.0); I("dypflg", 0); I("mixfrq", 1)
    F("diapyc", 0.0); F("rigr", 0.25); F("ribc", 0.45); F("rinfty", 0.7)
    F("rcnvec", 0.25); F("ricr", 0.45); F("bldmin", 0.0); F("bldmax", 1200.0)
    F("cekman", 0.7); F("cmonob", 1.0); I("bblkpp", 0); I("shinst", 1); I("dbdiff", 1)
    I("nonloc", 1); I("botdiw", 0); I("difout", 0); I("difsmo", 0); F("difm0", 0.05)
    F("difs0", 0.05); F("difmiw", 3.0e-5); F("difsiw", 1.0e-5); F("dsfmax", 0.1)
    F("rrho0", 1.9); F("cs", 98.96); F("cstar", 10.0); F("cv", 0.0); F("c11", 5.0)
    I("hblflg", 2); I("niter", 2); I("langmr", 0)
    I("fltflg", 0); I("nfladv", 4); I("nflsam", 1); I("intpfl", 0); I("iturbv", 0)
    I("ismpfl", 1); F("tbvar", 4.63e-6); F("tdecri", 0.4)
    I("lbflag", 0); I("lbmont", 0); I("tidflg", 0); I("tidein", 0); I("tidcon", 0)
    F("tidsal", 0.06); I("tiddrg", 0); F("thkdrg", 0.0); F("drgscl", 0.0)
    for n in ("tidfm2", "tidfs2", "tidfk1", "tidfo1", "drgscm", "drgscs", "drgsck", "drgsco"):
        F(n, 0.0)
    I("tidgen", 0); F("tidrmp", 0.0); F("tid_t0", 0.0); I("tidnud", 0)
    F("nudrmp", 0.0); F("nud_t0", 0.0)
    I("clmflg", 12); I("wndflg", 0); F("ocnscl", 0.0); I("ustflg", 1); I("flxflg", 0)
    I("empflg", 0); F("emptgt", 0.0); I("empbal", 0); I("dswflg", 0); I("albflg", 0)
    I("sssflg", 0); I("sssbal", 0); I("lwflag", 0); I("sstflg", 0); I("icmflg", 0)
    F("prsbas", 0.0); I("mslprf", 0); I("stroff", 0); I("flxoff", 0); I("flxsmo", 0)
    I("relax", 0); I("trcrlx", 0); I("priver", 0); I("epmass", 0)
    I("stdflg", 0); I("stdsur", 0); I("stdbot", 0); I("stdarc", 0); I("nsdzi", 0)
    with open(P("blkdat.input"), "w") as f:
        f.write("Closed basin, 0.5 degree, 15N-30N, from a z-level climatology\n")
        f.write("no wind, thermal or relaxation forcing\n")
        f.write(f"hybmap {hybmap}, hybraf {int(bool(hybraf))}, {ntracr} tracers\n")
        f.write("-" * 80 + "\n")
        f.write("\n".join(L) + "\n")
    return path
