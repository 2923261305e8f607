"""Big clause tiles (FP32, 3-literal systems with at least 2048 clauses per
resident workgroup: per-problem tile sizes, run tables, primary words) with
the tanh family (AnalogSAT/cpu_sat/CpuSatTanh.cpp:64-94,192-286) and the
bias term (AnalogSAT/cpu_sat/CpuSat.cpp:232-239), against the oracle.

The CTDS family without bias is bit-identical on this path
(tests/test_gpu_plain.py::test_big_tile_sizes_bitexact).  Here the bar is a
tolerance, for the reasons the small-problem tests of these families give:
the device tanh is not glibc's bit for bit, and the bias mean is a tree sum
in double (the oracle's mean_mode('double')).
  * RHS: |GPU - oracle| <= 16 ulp (2^-23) of the largest per-variable sum of
    |slot term| (tanh) or of the largest |RHS| (bias)
  * 10 integrator steps: the oracle's attempts, states within 1e-3 relative
"""
import sys

import numpy as np
import pytest

import analogsat_amd as A
import oracle as O

pytestmark = pytest.mark.gpu

N, M, Q, B = 310000, 1300000, 1.1, 2.5


def say(msg):
    print("[gpu-big]", msg, flush=True)
    sys.stdout.flush()


def _tanh_scale(p, x, q):
    """max over variables of sum_m |slot term| (bounds the rounding of the sums)"""
    C = np.asarray(p.C, np.int64).reshape(p.m, p.k)
    x = np.asarray(x, np.float64)
    s = np.where(C != 0, x[np.abs(C)], -1.0)
    f = np.where(C < 0, 1.0 + s, 1.0 - s)
    a = x[p.n + 1:p.n + 1 + p.m]
    tot = np.zeros(p.n + 1)
    for qi in range(p.k):
        prod = np.prod(np.delete(f, qi, axis=1), axis=1)
        term = 2.0 ** -p.k * a * np.abs(prod) * np.abs(np.tanh(q * f[:, qi]))
        np.add.at(tot, np.abs(C[:, qi]), np.where(C[:, qi] != 0, term, 0.0))
    return tot[1:].max() + 1e-300


@pytest.mark.parametrize("family", ["tanh", "bias"])
def test_big_tiles_tanh_and_bias_vs_oracle(family):
    say("big tiles, %s family, N=%d M=%d" % (family, N, M))
    p = O.fastsat(N, M, 3, 34)
    n, m = p.n, p.m
    if family == "tanh":
        sat = A.HipSatTanh("f32", q=Q)
        kw = dict(family="tanh", q=Q)
    else:
        sat = A.HipSat("f32")
        kw = dict(b=B)
    sat.SetProblem(A.SatProblem(p.literals, p.lengths))
    if family == "bias":
        sat.Set_B(B)
    lay = sat.Layout()
    say("  layout %s" % lay)
    assert 2048 < lay["tile_clauses"] <= 4608 and lay["run_tables"] == 1 and lay["primary_words"] == 1
    rng = np.random.default_rng(9)
    pub = np.concatenate([rng.uniform(-1, 1, n), rng.uniform(1, 2, m)]).astype(np.float32)
    x, d = sat.MakeState(), sat.MakeState()
    x.CopyFrom(pub)
    sat.GetDerivatives(d, x, 0.0)
    got = d.CopyTo().astype(np.float64)
    xi = np.concatenate([[-1.0], pub]).astype(np.float32)
    with O.mean_mode("double"):
        want = O.rhs(p, xi, mode="mix", **kw)[1:].astype(np.float64)
    if family == "tanh":
        scale = _tanh_scale(p, xi, Q)
    else:
        scale = np.abs(want).max() + np.abs(O.rhs(p, xi, mode="mix")[1:]).max()
    err = np.abs(got - want).max()
    say("  RHS max |diff| %.3e = %.2f ulp of the scale; bit-identical share %.4f"
        % (err, err / (2.0 ** -23 * scale), np.mean(got == want)))
    assert err <= 16 * 2.0 ** -23 * scale
    rk = A.HipRungeKutta()
    rk.SetEpsilon(1e-6)
    rk.Configure(sat)
    old, new = sat.MakeState(), sat.MakeState()
    old.CopyFrom(pub)
    with O.pow_mode("cr"), O.mean_mode("double"):
        orc = O.RK(p, eps=1e-6, mode="mix", **kw)
        orc.set_state(pub)
        rk.Step(old, new, 10)
        orc.step(10)
    g, ref = old.CopyTo().astype(np.float64), orc.get_state().astype(np.float64)
    rel = np.abs(g - ref).max() / np.abs(ref).max()
    st = rk.stats()
    say("  10 steps: rel diff %.3e, attempts %d vs %d" % (rel, st.attempts, orc.attempts))
    assert st.steps == 10 and st.attempts == orc.attempts
    assert rel <= 1e-3
