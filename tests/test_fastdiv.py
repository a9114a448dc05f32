"""The FP64 quotient helpers of hyb_math.h (rcp64/qdiv/qdiv_pos/qrcp/
qdiv_any/div3), host build, against IEEE division bit for bit.

The device forms its first reciprocal with v_rcp_f64; here the first
reciprocal is supplied, so the two Newton steps and Markstein's correction
are checked from reciprocals of a chosen error: the correctly rounded one,
a few ulp off, a float-precision one and a relative error of 2^-14.  The
operands cover the range the helpers claim (|b| in [2^-8, 2^100], a = 0
or |a| in [2^-918, 2^760)), its ends, and significands of all ones."""
import ctypes

import numpy as np
import pytest

import hostcheck_lib as hc

N = 400_000


def _quot(which, a, b=None, y0=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ones_like(a) if b is None else np.ascontiguousarray(b, dtype=np.float64)
    y0 = np.ones_like(a) if y0 is None else np.ascontiguousarray(y0, dtype=np.float64)
    out = np.empty_like(a)
    f = hc.lib().hostcheck_quot
    f.argtypes = [ctypes.c_int] + [ctypes.c_void_p] * 4 + [ctypes.c_int64]
    f(which, a.ctypes.data, b.ctypes.data, y0.ctypes.data, out.ctypes.data, a.size)
    return out


def _rand(rng, n, emin, emax):
    """Random doubles with exponents uniform in [emin, emax) and random
    significands."""
    m = 1.0 + rng.integers(0, 1 << 52, n, dtype=np.int64).astype(np.float64) * 2.0 ** -52
    return np.ldexp(m, rng.integers(emin, emax, n))


def _operands(rng):
    a = _rand(rng, N, -917, 759) * rng.choice([-1.0, 1.0], N)
    a[:1000] = 0.0
    a[1000:2000] = -0.0
    b = _rand(rng, N, -8, 100) * rng.choice([-1.0, 1.0], N)
    # ends of the range and all-ones significands
    ends = np.array([2.0 ** -918, np.nextafter(2.0 ** 760, 0), 2.0 ** 759, 1.0,
                     np.nextafter(2.0, 0), np.nextafter(1.0, 2)])
    a[2000:2000 + ends.size * 4] = np.repeat(ends, 4)
    b[2000:2000 + ends.size * 4] = np.tile([2.0 ** -8, 2.0 ** 100, np.nextafter(1.0, 2),
                                            np.nextafter(2.0, 0)], ends.size)
    # exact quotients and a == b
    a[3000:4000] = b[3000:4000] * np.round(_rand(rng, 1000, 0, 20))
    a[4000:4100] = b[4000:4100]
    return a, b


def _first_reciprocals(rng, b):
    r = 1.0 / b
    return {
        "rn": r,
        "plus3ulp": np.nextafter(np.nextafter(np.nextafter(r, np.inf), np.inf), np.inf),
        "minus2ulp": np.nextafter(np.nextafter(r, -np.inf), -np.inf),
        "float": (1.0 / b.astype(np.float32).astype(np.float64)).astype(np.float32).astype(np.float64),
        "rel2^-14": r * (1.0 + rng.uniform(-1, 1, b.size) * 2.0 ** -14),
    }


def _same_bits(x, y):
    return np.array_equal(x.view(np.int64), y.view(np.int64))


def _all_ones(b):
    return (b.view(np.int64) & ((1 << 52) - 1)) == (1 << 52) - 1


def test_qdiv_is_ieee_from_the_rounded_reciprocal():
    rng = np.random.default_rng(5)
    a, b = _operands(rng)
    want = a / b
    assert _same_bits(want, _quot(0, a, b, 1.0 / b))
    assert _same_bits(want, _quot(3, a, b, 1.0 / b))


def test_qdiv_is_ieee_from_inexact_first_reciprocals():
    """From a first reciprocal that is not the rounded one, the result is
    the IEEE quotient for every divisor except those whose significand is
    all ones: there 1/b lies next to a rounding tie, the Newton steps can
    stop one ulp short and the correction can miss by one ulp.  The device
    sequence is the compiler's own division without its (inactive) scaling
    steps, from the same v_rcp_f64, so for those divisors it gives what the
    compiler's a/b gives; the GPU test compares the two directly."""
    rng = np.random.default_rng(5)
    a, b = _operands(rng)
    want = a / b
    ones = _all_ones(b)
    assert ones.sum() >= 6  # the edge rows hold some
    for name, y0 in _first_reciprocals(rng, b).items():
        for which in (0, 3):
            got = _quot(which, a, b, y0)
            bad = np.flatnonzero((want.view(np.int64) != got.view(np.int64)) & ~ones)
            assert bad.size == 0, (f"which {which}, y0 {name}: {bad.size} quotients differ, "
                                   f"e.g. {a[bad[0]]!r}/{b[bad[0]]!r}")
            ulps = np.abs(want.view(np.int64)[ones] - got.view(np.int64)[ones])
            assert ulps.max() <= 1


def test_qdiv_pos_and_qrcp():
    rng = np.random.default_rng(6)
    a = _rand(rng, N, -917, 759)
    b = _rand(rng, N, -8, 100)
    b[:4] = [np.nextafter(2.0, 0), np.nextafter(1.0, 0) * 2.0 ** 40, 2.0 ** -8, 2.0 ** 100]
    ones = _all_ones(b)
    for name, y0 in _first_reciprocals(rng, b).items():
        ok = ones if name != "rn" else np.zeros_like(ones)
        for which, want in ((1, a / b), (2, 1.0 / b)):
            got = _quot(which, a, b, y0)
            assert np.array_equal(want.view(np.int64)[~ok], got.view(np.int64)[~ok]), \
                f"which {which}, y0 {name}"


def test_qdiv_any_outside_the_range():
    rng = np.random.default_rng(7)
    b = _rand(rng, 8, -8, 100)
    a = np.array([5e-324, -2.0 ** -1000, 2.0 ** -919, 2.0 ** 800, np.inf, -np.inf, 1e308,
                  2.0 ** -1022])
    got = _quot(3, a, b, 1.0 / b)
    want = a / b
    assert _same_bits(want, got)


def test_div3():
    rng = np.random.default_rng(8)
    x = _rand(rng, N, -917, 759) * rng.choice([-1.0, 1.0], N)
    x[:10] = [0.0, -0.0, 3.0, -3.0, 1.0, 2.0 ** -918, 2.0 ** 759, 6.0, 1.5, 0.75]
    x[10:5000] = 3.0 * np.round(_rand(rng, 4990, 0, 40))  # exact multiples of 3
    assert _same_bits(x / 3.0, _quot(4, x))


def test_the_correction_matters():
    """The check discriminates: without Markstein's correction (q0 = a*y
    alone) a sizeable share of the quotients is off by an ulp."""
    rng = np.random.default_rng(9)
    a, b = _operands(rng)
    q0 = a * (1.0 / b)
    wrong = np.count_nonzero(q0.view(np.int64) != (a / b).view(np.int64))
    assert wrong > N // 20
