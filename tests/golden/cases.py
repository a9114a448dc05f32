"""Golden cases of the hybgen path: parameter sets and input construction.

Shared by make_golden.py (which runs the flang build of the reference
hybgen on them) and the CPU tests (which run the oracle and the host build
of the device code on the same inputs).  Inputs come from the library's
seeded host generator; their SHA-256 is part of each fixture, so a change
of the generator shows up as an input mismatch, not as a parity failure."""
import hashlib

import numpy as np

import hybgen_abi as abi

# §8(d) parameter sweep: hybmap x {isopcm, hybrlx, hybthn, hybthk, nsigma,
# dx0k, ldenfix, trcflg incl. 2}, on a 60x40x32 config-1 tile
_SWEEP_SEED = 20240601


def _sweep():
    rng = np.random.default_rng(_SWEEP_SEED)
    out = []
    for hybmap in range(4):
        for c in range(8):
            dx = bool(rng.integers(0, 2))
            v = dict(hybmap=hybmap, idm=60, jdm=40,
                     hybiso=[0.0, 0.01][int(rng.integers(0, 2))],
                     hybrlx=[1.0, 8.0][int(rng.integers(0, 2))],
                     hybthn=[1.0, 5.0][int(rng.integers(0, 2))],
                     hybthk=[0.0, 0.5][int(rng.integers(0, 2))],
                     nsigma=[0, 8][int(rng.integers(0, 2))],
                     ldenfix=int(rng.integers(0, 2)),
                     nreg=int(rng.integers(0, 4)),
                     dx00=30.0 if dx else 9999.0, dx00f=1.1 if dx else 1.0,
                     dx00x=300.0 if dx else 9999.0)
            out.append(dict(name=f"cfg1_sweep_h{hybmap}_{c}", workload="cfg1", over=v,
                            seed=1000 + 10 * hybmap + c))
    return out


def cases():
    c = _sweep()
    c.append(dict(name="cfg1_full_ppm", workload="cfg1", over=dict(hybmap=2), seed=101))
    c.append(dict(name="cfg1_full_weno", workload="cfg1", over=dict(hybmap=3), seed=101))
    for m, n in ((1, 2), (2, 1)):
        c.append(dict(name=f"cfg1_raflag_m{m}n{n}", workload="cfg1",
                      over=dict(idm=48, jdm=30, ntracr=4, trcflg=[2, 0, 0, 0]),
                      itracr=(2, 3, 4), seed=300 + m, m=m, n=n, raflag=1))
    for hybmap in (2, 3):
        c.append(dict(name=f"glbb_tile_h{hybmap}", workload="glbb",
                      over=dict(idm=96, jdm=32, hybmap=hybmap),
                      window=dict(itdm=4500, jtdm=3298, i0=2100, j0=1500), seed=200))
    c.append(dict(name="k22_notracer_pcm", workload="cfg1",
                  over=dict(idm=50, jdm=30, kdm=22, ntracr=0, trcflg=[], hybmap=0), seed=7))
    c.append(dict(name="k41_extreme_ppm", workload="cfg1",
                  over=dict(idm=50, jdm=30, kdm=41, ntracr=2, trcflg=[2, 0], hybmap=2),
                  seed=8, extreme=1))
    c.append(dict(name="k17_extreme_weno", workload="cfg1",
                  over=dict(idm=40, jdm=30, kdm=17, nsigma=6, ntracr=3, trcflg=[0, 2, 0],
                            hybmap=3), seed=9, extreme=1))
    return c


def build(case):
    """(P, HostDomain, m, n, raflag) of a golden case."""
    P, w = abi.workload_params(case["workload"], **case["over"])
    it = case.get("itracr")
    if it:
        P.itracr801, P.itracr802, P.itracr803 = it
    win = case.get("window", {})
    S = abi.make_synth(P, w, seed=case["seed"], extreme=case.get("extreme", 0), **win)
    m, n = case.get("m", 1), case.get("n", 2)
    d = abi.HostDomain(P).synth(S, m=m, n=n)
    raflag = case.get("raflag", 0)
    if raflag:
        rng = np.random.default_rng(case["seed"])
        d.a["dpo"][...] = np.abs(d.a["dp"] * (1.0 + 0.01 * rng.standard_normal(d.a["dp"].shape)))
    return P, d, m, n, raflag


def sha(a):
    return hashlib.sha256(np.asfortranarray(a).tobytes(order="F")).hexdigest()


def input_sha(d):
    h = hashlib.sha256()
    for nme in abi.FIELD_NAMES:
        h.update(np.asfortranarray(d.a[nme]).tobytes(order="F"))
    return h.hexdigest()


def output_sha(d):
    return {nme: sha(d.a[nme]) for nme in abi.OUTPUTS}
