"""The FP64 quotient helpers on the device (hycom_hybgen_selftest_quotients)
against the compiler's division on the same operands and against IEEE
division on the host, bit for bit: random operands over the range the
helpers claim, its ends, divisors with all-ones significands (next to a
rounding tie) and, for qdiv_any, operands outside the range."""
import ctypes

import numpy as np
import pytest

import hybgen_abi as abi
import test_fastdiv as T

pytestmark = pytest.mark.gpu


def _device(which, a, b):
    import torch
    ta = torch.from_numpy(np.ascontiguousarray(a)).cuda()
    tb = torch.from_numpy(np.ascontiguousarray(b)).cuda()
    fast = torch.empty_like(ta)
    ieee = torch.empty_like(ta)
    abi.check(abi.load().hycom_hybgen_selftest_quotients(
        which, ta.data_ptr(), tb.data_ptr(), fast.data_ptr(), ieee.data_ptr(), ta.numel(), None))
    torch.cuda.synchronize()
    return fast.cpu().numpy(), ieee.cpu().numpy()


def _bits(x):
    return x.view(np.int64)


@pytest.mark.parametrize("which,name", [(0, "qdiv"), (3, "qdiv_any")])
def test_device_quotients(gpu, which, name):
    rng = np.random.default_rng(50 + which)
    a, b = T._operands(rng)
    b[5000:5400] = np.ldexp(np.nextafter(2.0, 0), rng.integers(-8, 99, 400))  # all ones
    print(f"\n[case] {name} {a.size} operands", flush=True)
    fast, ieee = _device(which, a, b)
    assert np.array_equal(_bits(ieee), _bits(a / b)), "compiler division is not IEEE here"
    bad = np.flatnonzero(_bits(fast) != _bits(ieee))
    assert bad.size == 0, f"{name}: {bad.size} differ, e.g. {a[bad[0]]!r}/{b[bad[0]]!r}"


def test_device_qdiv_pos_qrcp(gpu):
    rng = np.random.default_rng(60)
    a = T._rand(rng, T.N, -917, 759)
    b = T._rand(rng, T.N, -8, 100)
    b[:400] = np.ldexp(np.nextafter(2.0, 0), rng.integers(-8, 99, 400))
    for which, want in ((1, a / b), (2, 1.0 / b)):
        print(f"\n[case] which={which}", flush=True)
        fast, ieee = _device(which, a, b)
        assert np.array_equal(_bits(ieee), _bits(want))
        assert np.array_equal(_bits(fast), _bits(ieee))


def test_device_qdiv_any_outside_the_range(gpu):
    rng = np.random.default_rng(61)
    a = np.array([5e-324, -2.0 ** -1000, 2.0 ** -919, 2.0 ** 800, np.inf, -np.inf, 1e308,
                  2.0 ** -1022, 0.0, -0.0])
    b = T._rand(rng, a.size, -8, 100)
    print("\n[case] qdiv_any outside the range", flush=True)
    fast, ieee = _device(3, a, b)
    assert np.array_equal(_bits(fast), _bits(ieee))
    assert np.array_equal(_bits(ieee), _bits(a / b))


def test_device_div3(gpu):
    rng = np.random.default_rng(62)
    x = T._rand(rng, T.N, -917, 759) * rng.choice([-1.0, 1.0], T.N)
    x[:10] = [0.0, -0.0, 3.0, -3.0, 1.0, 2.0 ** -918, 2.0 ** 759, 6.0, 1.5, 0.75]
    print("\n[case] div3", flush=True)
    fast, ieee = _device(4, x, np.ones_like(x))
    assert np.array_equal(_bits(ieee), _bits(x / 3.0))
    assert np.array_equal(_bits(fast), _bits(ieee))
