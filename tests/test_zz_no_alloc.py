"""asat_solve_noalloc (every device buffer in one static per-device array)
against asat_rk_solve on the same problems, through the C ABI: result,
solver steps, last violation count, integrator counters, time, step size and
the final state, bit for bit, on two calls in a row, the second with a
smaller problem (so the array keeps the first call's larger contents).
Also: a problem too large for the path is refused with an error."""
import ctypes
import sys

import numpy as np
import pytest

import analogsat_amd as A
import oracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print("[gpu-noalloc]", msg, flush=True)
    sys.stdout.flush()


def _with_handles(p, prec, pub, batch, max_steps, max_time):
    sat = A.HipSat(prec)
    sat.SetProblem(A.SatProblem(p.literals, p.lengths))
    rk = A.HipRungeKutta()
    rk.SetEpsilon(1e-6)
    rk.Configure(sat)
    assert rk.UsesOneWorkgroup()
    old, new = sat.MakeState(), sat.MakeState()
    old.CopyFrom(pub)
    r = rk.SolveBatches(old, new, batch, max_steps, max_time)
    return r, rk.stats(), old.CopyTo()


def _noalloc(p, prec, pub, batch, max_steps, max_time):
    L = A.lib()
    out = A.SolveResult()
    st = A.RkStats()
    dt = np.float64 if prec == "f64" else np.float32
    pub = np.ascontiguousarray(pub, dt)
    res = np.zeros_like(pub)
    lits = np.ascontiguousarray(p.literals, np.int32)
    lens = np.ascontiguousarray(p.lengths, np.int32)
    rc = L.asat_solve_noalloc(0 if prec == "f64" else 1, lits.ctypes.data, lens.ctypes.data, len(lens),
                              pub.ctypes.data, 1e-6, batch, max_steps, max_time, 0, res.ctypes.data,
                              ctypes.byref(out), ctypes.byref(st))
    return rc, out, st, res


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_noalloc_equals_handle_solve_two_calls(prec):
    dt = np.float64 if prec == "f64" else np.float32
    cases = [(O.fastsat(200, 852, 3, 21), 7, 1, 600), (O.fastsat(50, 213, 3, 4), 8, 1, 10000)]
    for p, seed, batch, max_steps in cases:
        pub = np.concatenate([np.random.default_rng(seed).uniform(-1, 1, p.n), np.ones(p.m)]).astype(dt)
        say("%s: N=%d M=%d batch %d max_steps %d" % (prec, p.n, p.m, batch, max_steps))
        r0, s0, x0 = _with_handles(p, prec, pub, batch, max_steps, 1e8)
        rc, r1, s1, x1 = _noalloc(p, prec, pub, batch, max_steps, 1e8)
        assert rc == 0, A.lib().asat_last_error()
        assert (r1.result, r1.steps, r1.last_violations) == (r0.result, r0.steps, r0.last_violations)
        for f in ("time", "hdid", "hnext", "errmax", "steps", "attempts", "rhscount"):
            assert getattr(s1, f) == getattr(s0, f), f
        assert np.array_equal(x1, x0)


def test_noalloc_refuses_too_large_problem():
    p = O.fastsat(20000, 85000, 3, 1)
    pub = np.concatenate([np.zeros(p.n), np.ones(p.m)])
    rc, _, _, _ = _noalloc(p, "f64", pub, 1, 10, 1e8)
    assert rc == 1
    assert b"too large" in A.lib().asat_last_error()
