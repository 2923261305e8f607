"""The step controller's pow (analogsat_amd/csrc/asat_powctrl.h), host build.

The device kernels and this host build compile the same source with the same
operation order (no contraction), so these checks pin what the device
computes.  Bar: x^-0.25 and x^-0.2 rounded to nearest, i.e. equal to a
200-bit mpmath pow rounded once (the reference's std::pow at
AnalogSAT/cpu_sat/CpuRungeKutta.cpp:236-252 is glibc's, correctly rounded in
practice), and equal to the oracle's independent correctly rounded pow
(seeded from the library pow) on a large sample.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def pc(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("powctrl") / "libpowctrl.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared",
                    "-I" + os.path.join(ROOT, "analogsat_amd", "csrc"),
                    os.path.join(ROOT, "tests", "cpp", "powctrl_host.cpp"), "-o", out], check=True)
    lib = ctypes.CDLL(out)
    for f in (lib.pc_pow_many,):
        f.argtypes = [ctypes.c_void_p, ctypes.c_longlong, ctypes.c_int, ctypes.c_void_p]
    lib.pc_ln_many.argtypes = [ctypes.c_void_p, ctypes.c_longlong, ctypes.c_void_p]

    def pow_many(x, q):
        x = np.ascontiguousarray(x, np.float64)
        y = np.empty_like(x)
        lib.pc_pow_many(x.ctypes.data, x.size, q, y.ctypes.data)
        return y

    def ln_many(x):
        x = np.ascontiguousarray(x, np.float64)
        y = np.empty_like(x)
        lib.pc_ln_many(x.ctypes.data, x.size, y.ctypes.data)
        return y

    return pow_many, ln_many


def _samples(seed, n):
    rng = np.random.default_rng(seed)
    return np.concatenate([
        10 ** rng.uniform(np.log10(1.89e-4), 0.0, n),   # accepted steps: ERRCON < errmax <= 1 (q = 5)
        10 ** rng.uniform(0.0, 8.0, n // 2),             # rejected steps: errmax > 1 (q = 4)
        10 ** rng.uniform(-299.0, 299.0, n // 4),        # the rest of the normal range
        [1.0, 1.89e-4, 2.0, 0.5, 1e-299, 1e299, 1.0000000000000002, 0.9999999999999999],
    ])


def test_controller_pow_correctly_rounded(pc):
    mpmath = pytest.importorskip("mpmath")
    mpmath.mp.prec = 200
    pow_many, _ = pc
    xs = _samples(7, 1600)
    for q, y in ((4, "-0.25"), (5, "-0.2")):
        got = pow_many(xs, q)
        yv = mpmath.mpf(float(y))   # the double nearest -0.2, as the reference passes it
        for x, g in zip(xs, got):
            t = mpmath.power(mpmath.mpf(float(x)), yv)
            assert g == mpmath.libmp.to_float(t._mpf_, rnd="n"), (x, q)


def test_controller_pow_equals_oracle_pow(pc):
    pow_many, _ = pc
    xs = _samples(11, 40000)
    for q in (4, 5):
        got = pow_many(xs, q)
        want = np.array([O.pow_cr(float(x), q) for x in xs])
        bad = np.nonzero(got != want)[0]
        assert bad.size == 0, (q, xs[bad[:5]], got[bad[:5]], want[bad[:5]])


def test_controller_pow_special_values(pc):
    pow_many, _ = pc
    libm = ctypes.CDLL("libm.so.6")
    libm.pow.argtypes = [ctypes.c_double, ctypes.c_double]
    libm.pow.restype = ctypes.c_double
    xs = np.array([0.0, np.inf, np.nan, 1e-305, 1e305, -1.0])
    for q, y in ((4, -0.25), (5, -0.2)):
        got = pow_many(xs, q)
        want = np.array([libm.pow(float(x), y) for x in xs])   # outside (1e-300, 1e300): the library pow
        assert np.array_equal(got, want, equal_nan=True), (q, got, want)


def test_ln_series_accuracy(pc):
    _, ln_many = pc
    rng = np.random.default_rng(5)
    xs = np.concatenate([10 ** rng.uniform(-299, 299, 100000), [1.0, np.sqrt(2.0), np.nextafter(np.sqrt(2.0), 2)]])
    err = np.abs(ln_many(xs) - np.log(xs))
    assert err.max() < 2e-9, err.max()
