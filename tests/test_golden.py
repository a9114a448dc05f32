"""The oracle and the host build of the device code against the golden
fixtures written from the flang build of the reference hybgen
(tests/golden/make_golden.py): every output array bitwise, by SHA-256."""
import json
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import cases as gc  # noqa: E402
import hostcheck_lib as hc  # noqa: E402
import oracle_lib as ol  # noqa: E402

with open(os.path.join(HERE, "golden", "hybgen_golden.json")) as _f:
    GOLDEN = json.load(_f)["cases"]
IDS = [c["name"] for c in GOLDEN]


def _inputs(case):
    P, d, m, n, raflag = gc.build(case)
    assert gc.input_sha(d) == case["input_sha256"], \
        f"{case['name']}: generated inputs differ from the fixture's"
    return P, d, m, n, raflag


def _check(case, d, n, who):
    got = gc.output_sha(d)
    bad = [k for k, v in case["output_sha256"].items() if got[k] != v]
    assert not bad, f"{who} on {case['name']}: {bad} differ from the reference"
    for s in case["samples"]:
        i, j = s["i"] - 1 + d.P.nbdy, s["j"] - 1 + d.P.nbdy
        for f in ("dp", "temp", "p", "u"):
            a = d.a[f]
            col = a[i, j, :, n - 1] if a.ndim == 4 else a[i, j, :]
            want = np.array([float.fromhex(x) for x in s[f]])
            assert np.array_equal(col, want), f"{who}: {case['name']} {f} column ({s['i']},{s['j']})"


@pytest.mark.parametrize("case", GOLDEN, ids=IDS)
def test_oracle_matches_reference_fixture(case):
    P, d, m, n, raflag = _inputs(case)
    ol.run_oracle(d, m=m, n=n, raflag=raflag, threads=4)
    _check(case, d, n, "oracle")


@pytest.mark.parametrize("family", list(hc.FAMILIES))
@pytest.mark.parametrize("case", GOLDEN, ids=IDS)
def test_device_code_on_host_matches_reference_fixture(case, family):
    P, d, m, n, raflag = _inputs(case)
    if not hc.covers(family, P):
        pytest.skip(f"{family} kernels do not cover this case")
    hc.run(family, d, m=m, n=n, raflag=raflag)
    _check(case, d, n, f"{family} (host build)")
