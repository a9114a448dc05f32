"""Writes tests/golden/hybgen_golden.json from the reference hybgen.

The expected outputs come from HYCOM-src's own hybgen.F90 / dpudpv.F90,
compiled by flang from /root/reference (oracle/ref/Makefile ->
oracle/_ref/libhybgen_ref*.so, the HYBGEN_DENFIX variant for ldenfix=1) and
called through oracle/ref/hybgen_ref_driver.F90.  Each case stores the
SHA-256 of every input array and of every output array (bitwise), and the
full output columns of a few sea points as exact hex floats.

    python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "hycom-src_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, HERE)

import cases as gc  # noqa: E402
import oracle_lib as ol  # noqa: E402

SAMPLE_FIELDS = ("dp", "temp", "saln", "th3d", "p", "u", "v")


def samples(P, d, n, k=3):
    nb = P.nbdy
    ip = d.a["ip"][nb:nb + P.ii, nb:nb + P.jj]
    pts = np.argwhere(ip > 0)
    if len(pts) == 0:
        return []
    sel = pts[np.linspace(0, len(pts) - 1, k).astype(int)]
    out = []
    for i0, j0 in sel:
        i, j = int(i0) + nb, int(j0) + nb
        rec = {"i": int(i0) + 1, "j": int(j0) + 1}
        for f in SAMPLE_FIELDS:
            a = d.a[f]
            col = a[i, j, :, n - 1] if a.ndim == 4 else a[i, j, :]
            rec[f] = [float(x).hex() for x in col]
        out.append(rec)
    return out


def main():
    if not ol.ref_available() or not ol.ref_available(True):
        sys.exit("oracle/_ref is not built: make -C oracle/ref")
    out = {"about": "expected outputs of HYCOM-src hybgen (flang build of hybgen.F90, "
                    "dpudpv.F90) on the golden cases of tests/golden/cases.py",
           "cases": []}
    for c in gc.cases():
        P, d, m, n, raflag = gc.build(c)
        rec = dict(c)
        rec["input_sha256"] = gc.input_sha(d)
        ol.run_ref(d, m=m, n=n, raflag=raflag)
        rec["output_sha256"] = gc.output_sha(d)
        rec["samples"] = samples(P, d, n)
        out["cases"].append(rec)
        print(c["name"], flush=True)
    with open(os.path.join(HERE, "hybgen_golden.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
