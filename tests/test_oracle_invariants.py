"""Checks of the oracle that do not use the reference: closed forms of the
remappers (hybgen.F90:2276-3107) and invariants of hybgen as a whole.
Each targets a plausible mistake (a dropped term, a swapped coefficient, a
wrong index) that the bitwise fixtures would also catch, but these say
what is wrong."""
import ctypes
from fractions import Fraction

import numpy as np
import pytest

import hybgen_abi as abi
import oracle_lib as ol

THIN = 0.001 * (9.806 / 0.001) * 0.001  # dpthin = 0.001*onemm (hybgen.F90:374)
D = ctypes.POINTER(ctypes.c_double)
I = ctypes.POINTER(ctypes.c_int)


def remap(method, si, pi, po, lcm=None):
    kk = len(si)
    si = np.ascontiguousarray(si, dtype=np.float64)
    pi = np.ascontiguousarray(pi, dtype=np.float64)
    po = np.ascontiguousarray(po, dtype=np.float64)
    lc = np.zeros(kk, np.int32) if lcm is None else np.ascontiguousarray(lcm, np.int32)
    so = np.zeros(kk)
    ol.oracle().oracle_remap_column(method, kk, si.ctypes.data_as(D), pi.ctypes.data_as(D),
                                    po.ctypes.data_as(D), lc.ctypes.data_as(I),
                                    ctypes.c_double(THIN), so.ctypes.data_as(D))
    return so


def cell_avg(f, a, b):
    """exact average over [a,b] of the polynomial with coefficients f
    (f[0] + f[1] x + f[2] x^2), in rationals"""
    a, b = Fraction(a), Fraction(b)
    F = lambda x: sum(Fraction(c) * x ** (n + 1) / (n + 1) for n, c in enumerate(f))
    return (F(b) - F(a)) / (b - a)


@pytest.mark.parametrize("method", [0, 1, 2, 3])
def test_constant_profile_is_reproduced_exactly(method):
    rng = np.random.default_rng(method)
    for _ in range(50):
        kk = int(rng.integers(3, 30))
        pi = np.concatenate([[0.0], np.cumsum(rng.exponential(10.0, kk))])
        po = np.concatenate([[0.0], np.sort(rng.uniform(0, pi[-1], kk - 1)), [pi[-1]]])
        c = float(rng.normal(10, 5))
        so = remap(method, np.full(kk, c), pi, po)
        assert (so == c).all()


@pytest.mark.parametrize("method", [0, 1, 2, 3])
def test_identity_grid_returns_the_input(method):
    rng = np.random.default_rng(10 + method)
    kk = 25
    pi = np.concatenate([[0.0], np.cumsum(rng.exponential(10.0, kk))])
    si = np.cumsum(rng.normal(0, 1, kk))
    so = remap(method, si, pi, pi.copy())
    assert np.allclose(so, si, rtol=0, atol=1e-12 * np.abs(si).max())


@pytest.mark.parametrize("method,degree", [(1, 1), (2, 2), (3, 2)])
def test_polynomial_reproduction_on_uniform_grid(method, degree):
    """PLM reproduces a linear profile, PPM and WENO a quadratic one, in the
    cells away from the ends of the column, onto an irregular new grid (the
    edge values of Colella & Woodward 1984 are exact for quadratics on a
    uniform grid; the limiters leave monotone polynomials alone).  PCM and
    PLM miss a quadratic here by 0.8 and 0.015."""
    kk, h = 30, 4.0
    f = [1.0, 0.37, 0.011][: degree + 1]
    pi = h * np.arange(kk + 1, dtype=np.float64)
    si = np.array([float(cell_avg(f, pi[k], pi[k + 1])) for k in range(kk)])
    po = pi + h * np.random.default_rng(0).uniform(0.1, 0.9, kk + 1)
    po[0], po[-1] = 0.0, pi[-1]
    so = remap(method, si, pi, po)
    for k in range(4, kk - 4):
        want = float(cell_avg(f, po[k], po[k + 1]))
        assert abs(so[k] - want) <= 1e-11 * abs(want), (k, so[k], want)


def test_pcm_is_the_overlap_average():
    rng = np.random.default_rng(3)
    for _ in range(40):
        kk = 8
        pi = np.concatenate([[0.0], np.cumsum(rng.uniform(1, 10, kk))])
        po = np.concatenate([[0.0], np.sort(rng.uniform(0, pi[-1], kk - 1)), [pi[-1]]])
        si = rng.normal(0, 1, kk)
        so = remap(0, si, pi, po)
        for k in range(kk):
            a, b = Fraction(po[k]), Fraction(po[k + 1])
            if b - a <= Fraction(THIN):
                continue
            tot = Fraction(0)
            for l in range(kk):
                lo, hi = max(a, Fraction(pi[l])), min(b, Fraction(pi[l + 1]))
                if hi > lo:
                    tot += (hi - lo) * Fraction(si[l])
            want = float(tot / (b - a))
            assert abs(so[k] - want) <= 1e-12 * max(1.0, abs(want))


@pytest.mark.parametrize("method", [0, 1, 2, 3])
def test_column_conservation_and_bounds(method):
    """sum(so*dpo) == sum(si*dpi) without thin outputs and without clamping
    (the lconserve sums, hybgen.F90:1583-1589); every output inside the
    input range (the clamp, e.g. 2836-2837)."""
    rng = np.random.default_rng(20 + method)
    for _ in range(60):
        kk = int(rng.integers(4, 30))
        pi = np.concatenate([[0.0], np.cumsum(rng.uniform(1, 20, kk))])
        po = np.concatenate([[0.0], np.sort(rng.uniform(0, pi[-1], kk - 1)), [pi[-1]]])
        si = np.sin(np.linspace(0, 3, kk)) * 5 + rng.normal(0, 0.01, kk)
        so = remap(method, si, pi, po)
        assert so.min() >= si.min() and so.max() <= si.max()
        mass_in = np.sum(si * np.diff(pi))
        mass_out = np.sum(so * np.diff(po))
        if not ((so == si.min()) | (so == si.max())).any():
            assert abs(mass_out - mass_in) <= 1e-10 * np.sum(np.abs(si) * np.diff(pi))


def _run(over, seed=5, extreme=0):
    P, w = abi.workload_params("cfg1", **over)
    d = abi.HostDomain(P).synth(abi.make_synth(P, w, seed=seed, extreme=extreme))
    a = d.copy()
    ol.run_oracle(a, threads=4)
    return P, d, a


@pytest.mark.parametrize("hybmap", [0, 1, 2, 3])
def test_t_tracer_stays_equal_to_temperature(hybmap):
    """A trcflg=2 tracer initialised to T stays bitwise equal to T through
    hybgen (the reference's PIPE_TRACER check, mod_pipe.F90:1524-1530)."""
    P, w = abi.workload_params("cfg1", idm=50, jdm=30, hybmap=hybmap, ntracr=2,
                               trcflg=[2, 0])
    d = abi.HostDomain(P).synth(abi.make_synth(P, w, seed=4))
    d.a["tracer"][:, :, :, 1, 0] = d.a["temp"][:, :, :, 1]
    ol.run_oracle(d, threads=4)
    assert np.array_equal(d.a["tracer"][:, :, :, 1, 0], d.a["temp"][:, :, :, 1])


@pytest.mark.parametrize("extreme", [0, 1])
def test_interfaces_ordered_and_bottom_kept(extreme):
    P, d, a = _run(dict(idm=60, jdm=40, hybmap=2), extreme=extreme)
    nb = P.nbdy
    sea = d.a["ip"][nb:nb + P.ii, nb:nb + P.jj] > 0
    p0 = np.concatenate([np.zeros(d.a["dp"].shape[:2] + (1,)),
                         np.cumsum(d.a["dp"][:, :, :, 1], axis=2)], axis=2)
    p1 = a.a["p"]
    dp1 = a.a["dp"][:, :, :, 1]
    inner = (slice(nb, nb + P.ii), slice(nb, nb + P.jj))
    assert (dp1[inner][sea] >= 0).all()
    assert (np.diff(p1[inner][sea], axis=1) >= 0).all()
    bot0, bot1 = p0[inner][sea][:, -1], p1[inner][sea][:, -1]
    assert np.allclose(bot1, bot0, rtol=1e-12, atol=0)
    assert (p1[inner][sea][:, 0] == 0).all()
