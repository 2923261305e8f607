"""GPU tanh CTDS family (CpuSatTanh, AnalogSAT/cpu_sat/CpuSatTanh.cpp:64-94,192-286)
against the reference's own outputs and the oracle.

The slot term calls tanh; the device library's tanh is not glibc's bit for
bit, so the bar is a tolerance written here, not bit equality:
  * RHS: |GPU - reference| <= 8 ulp of the largest |term| (FP64: 2^-52,
    FP32: 2^-23) per element, against the committed CpuSatTanh fixtures
    (16 ulp with the bias term, whose mean is a tree sum here); the share of
    bit-identical elements is printed
  * trajectories (50 steps from the C1 start): same accepted step count and
    attempts as the oracle; states within 1e-9 (FP64) / 1e-4 (FP32) relative
"""
import os
import sys

import numpy as np
import pytest

import analogsat_amd as A
import oracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print("[gpu-tanh]", msg, flush=True)
    sys.stdout.flush()


def _rhs(sat, pub):
    x = sat.MakeState()
    x.CopyFrom(pub)
    d = sat.MakeState()
    sat.GetDerivatives(d, x, 0.0)
    return d.CopyTo()


def _term_scale(p, x, q):
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


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("case", ["k3", "k4", "mixed", "k12", "n2000", "bias"])
def test_tanh_rhs_vs_reference(golden_dir, case, precision):
    say("tanh rhs %s %s" % (precision, case))
    r = np.load(os.path.join(golden_dir, "rhs_cases.npz"), allow_pickle=False)
    t = np.load(os.path.join(golden_dir, "tanh_cases.npz"), allow_pickle=False)
    q = float(t[case + "_q"])
    sat = A.HipSatTanh(precision, q=q)
    sat.SetProblem(A.SatProblem(r[case + "_lits"], r[case + "_lens"]))
    sat.Set_B(float(r[case + "_b"]))
    assert sat.Get_Q() == pytest.approx(q)
    dt = np.float64 if precision == "f64" else np.float32
    x = r[case + "_x"]
    got = _rhs(sat, x[1:].astype(dt)).astype(np.float64)
    want = (t[case + "_rhs64"] if precision == "f64" else t[case + "_rhs32"])[1:].astype(np.float64)
    ulp = 2.0 ** -52 if precision == "f64" else 2.0 ** -23
    p = O.Problem(r[case + "_lits"], r[case + "_lens"])
    scale = max(_term_scale(p, x, q), np.abs(want).max())
    tol = 8 if float(r[case + "_b"]) == 0.0 else 16   # the bias mean is a tree sum here
    err = np.abs(got - want).max()
    say("  max |diff| = %.3e (%.2f ulp of scale), bit-identical %.1f%%"
        % (err, err / (ulp * scale), 100.0 * np.mean(got == want)))
    assert err <= tol * ulp * scale


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_tanh_trajectory_vs_oracle(golden_dir, precision):
    say("tanh trajectory %s 50 steps" % precision)
    g = np.load(os.path.join(golden_dir, "c1_traj.npz"), allow_pickle=False)
    p = O.fastsat(50, 213, 3, 0)
    mode = "f64" if precision == "f64" else "mix"
    dt = np.float64 if precision == "f64" else np.float32
    sat = A.HipSatTanh(precision, q=1.1)
    sat.SetProblem(A.SatProblem(p.literals, p.lengths))
    rk = A.HipRungeKutta()
    rk.SetEpsilon(1e-6)
    rk.Configure(sat)
    old, new = sat.MakeState(), sat.MakeState()
    pub = g["pub0"].astype(dt)
    old.CopyFrom(pub)
    with O.pow_mode("cr"):
        orc = O.RK(p, eps=1e-6, mode=mode, family="tanh", q=1.1)
        orc.set_state(pub)
        rk.Step(old, new, 50)
        orc.step(50)
    x, ref = old.CopyTo().astype(np.float64), orc.get_state().astype(np.float64)
    rel = np.abs(x - ref).max() / np.abs(ref).max()
    say("  rel diff %.3e, attempts %d vs %d" % (rel, rk.stats().attempts, orc.attempts))
    assert rk.stats().steps == 50 and rk.stats().attempts == orc.attempts
    assert rel <= (1e-9 if precision == "f64" else 1e-4)
    assert sat.GetClauseViolationCount(old) == orc.violations()
