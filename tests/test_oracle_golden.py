"""Pin the C oracle against the reference's own CPU path (CPU only).

Golden fixtures in tests/golden/ were written by tests/golden/make_golden.py
from the reference sources compiled in place (oracle/_ref).  Every check here is
bit-exact (FP64 equality), which a dropped term, a swapped operand order or a
wrong sign would break.

Pinned by these tests:
  orc_rhs_f64        CpuSat<double>::GetDerivatives, k = 3, 4, mixed lengths
                     with padding, k = 12 (generic CalculateAll order), bias
  orc_rk_*_f64       CpuRungeKutta<double>: per-step hdid/time/attempts and
                     states over 1000 steps (rkck + rkqs + controller + pow)
  orc_solve_f64      SatSolver::DoSolve result, steps, per-batch violations
  orc_violations     per-step violation counts
  orc_fastsat        FastSatMaker literal streams (libstdc++ minstd_rand0)
  orc_uniform_f64    CpuRandom<double> uniform streams
  orc_problem_build  SatProblem::AddClauses (zero literals dropped, N/M/K)
"""
import os

import numpy as np
import pytest

import oracle as O


def _load(golden_dir, name):
    return np.load(os.path.join(golden_dir, name), allow_pickle=False)


def read_cnf(path):
    lits, lens = [], []
    with open(path) as f:
        for line in f:
            t = line.split()
            if not t or t[0] in ("c", "p"):
                continue
            vals = [int(v) for v in t]
            assert vals[-1] == 0
            lits += vals[:-1]
            lens.append(len(vals) - 1)
    return O.Problem(np.array(lits, np.int32), np.array(lens, np.int32))


@pytest.fixture(scope="module")
def c1(golden_dir):
    return read_cnf(os.path.join(golden_dir, "c1_n50_m213_s0.cnf"))


def test_c1_instance_is_fastsat_seed0(c1):
    assert (c1.n, c1.m, c1.k) == (50, 213, 3)
    g = O.fastsat(50, 213, 3, 0)
    assert np.array_equal(g.literals, c1.literals)


def test_rhs_cases_bitexact(golden_dir):
    d = _load(golden_dir, "rhs_cases.npz")
    names = sorted({k.rsplit("_", 1)[0] for k in d.files if k.endswith("_rhs")})
    assert set(names) == {"k3", "k4", "mixed", "k12", "bias", "n2000"}
    for nm in names:
        p = O.Problem(d[nm + "_lits"], d[nm + "_lens"])
        got = O.rhs(p, d[nm + "_x"], b=float(d[nm + "_b"]))
        want = d[nm + "_rhs"]
        assert got.shape == want.shape, nm
        assert np.array_equal(got, want), (nm, np.abs(got - want).max())


def test_problem_build_drops_zeros(golden_dir):
    d = _load(golden_dir, "rhs_cases.npz")
    p = O.Problem(d["mixed_lits"], d["mixed_lens"])
    assert p.m == len(d["mixed_lens"])
    assert p.k == 5
    assert 0 not in set(p.C.reshape(p.m, p.k)[:, 0].tolist())
    # padded matrix holds exactly the non-zero literals in order
    assert np.array_equal(p.C[p.C != 0], d["mixed_lits"][d["mixed_lits"] != 0])


def test_trajectory_bitexact(golden_dir, c1):
    d = _load(golden_dir, "c1_traj.npz")
    rk = O.RK(c1, eps=float(d["eps"]))
    rk.set_state(d["pub0"])
    T = len(d["hdid"])
    for s in range(1, T + 1):
        assert rk.step(1)
        assert rk.hdid == d["hdid"][s - 1], s
        assert rk.time == d["time"][s - 1], s
        assert rk.attempts == d["attempts"][s - 1], s
        assert rk.violations() == d["violations"][s - 1], s
        key = "state_%d" % s
        if key in d.files:
            assert np.array_equal(rk.get_state(), d[key]), s
    assert rk.rhscount == int(d["rhscount"])


def test_solve_bitexact(golden_dir, c1):
    d = _load(golden_dir, "c1_solve.npz")
    t = _load(golden_dir, "c1_traj.npz")
    rk = O.RK(c1, eps=1e-6)
    rk.set_state(t["pub0"])
    res, viol = rk.solve(max_time=1e8, max_steps=10000, batch=1)
    assert res == int(d["result"]) == 0
    assert np.array_equal(viol, d["batch_violations"])
    assert rk.steps == int(d["steps"])
    assert rk.rhscount == int(d["rhscount"])
    assert rk.time == float(d["time"])
    assert np.array_equal(rk.get_state(), d["final"])


def test_generator_streams(golden_dir):
    d = _load(golden_dir, "gen_cases.npz")
    for key in d.files:
        parts = key.split("_")
        if parts[0] == "fastsat":
            n, m, k, sd = map(int, parts[1:])
            assert np.array_equal(O.fastsat(n, m, k, sd).literals, d[key]), key
        else:
            sd = int(parts[1])
            assert np.array_equal(O.uniform_f64(sd, len(d[key])), d[key]), key


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built here")
def test_live_reference_random_instances():
    """Where the reference build exists, compare on fresh seeds as well."""
    R = O.ref()
    for seed in range(5):
        p = O.fastsat(40 + 10 * seed, int(4.26 * (40 + 10 * seed)), 3, 100 + seed)
        pub = O.random_state_f64(seed, p.n, p.m)
        pub[p.n:] = 1.0
        rk = O.RK(p, eps=1e-6)
        rk.set_state(pub)
        h = R.ref_rk_create(p.literals, p.lengths, p.nclauses, 1e-6)
        R.ref_rk_set_state(h, pub)
        for _ in range(20):
            rk.step(25)
            R.ref_rk_step(h, 25)
            xr = np.zeros(p.n + p.m)
            R.ref_rk_get_state(h, xr)
            assert np.array_equal(rk.get_state(), xr)
            assert rk.hdid == R.ref_rk_hdid(h)
        R.ref_rk_destroy(h)


@pytest.mark.parametrize("case", ["k3", "k4", "mixed", "k12", "bias", "n2000"])
def test_tanh_cases_bitexact(golden_dir, case):
    """The tanh family (AnalogSAT/cpu_sat/CpuSatTanh.cpp:64-94,192-286) equals
    the reference's CpuSatTanh<double> and CpuSatTanh<float> bit for bit."""
    r = np.load(os.path.join(golden_dir, "rhs_cases.npz"), allow_pickle=False)
    t = np.load(os.path.join(golden_dir, "tanh_cases.npz"), allow_pickle=False)
    p = O.Problem(r[case + "_lits"], r[case + "_lens"])
    x, b, q = r[case + "_x"], float(r[case + "_b"]), float(t[case + "_q"])
    assert np.array_equal(O.rhs(p, x, b=b, family="tanh", q=q), t[case + "_rhs64"])
    x32 = x.astype(np.float32)
    assert np.array_equal(O.rhs(p, x32, mode="mix", b=b, family="tanh", q=q), t[case + "_rhs32"])


def test_tanh_zero_spin_closed_form():
    """At s = 0 every factor is 1: da_m = a_m 4^-k and
    ds_i = sum_m c_mi 2^-k a_m tanh(q)."""
    p = O.fastsat(40, 170, 3, 4)
    a = np.random.default_rng(0).uniform(1, 2, p.m)
    x = np.concatenate([[-1.0], np.zeros(p.n), a])
    q = 1.7
    got = O.rhs(p, x, family="tanh", q=q)
    want_s = np.zeros(p.n + 1)
    lits = p.literals.reshape(p.m, 3)
    for j in range(p.m):
        for l in lits[j]:
            want_s[abs(l)] += np.sign(l) * 0.125 * a[j] * np.tanh(q)
    np.testing.assert_allclose(got[1:p.n + 1], want_s[1:], rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(got[p.n + 1:], a / 64.0, rtol=1e-15)


def test_long_double_trajectory_bitexact(golden_dir, c1):
    """The extended-precision mode (CpuSat<long double> +
    CpuRungeKutta<long double>, AnalogSAT/cpu_sat/CpuSat.cpp:393,
    AnalogSAT/cpu_sat/CpuRungeKutta.cpp:273) equals the reference's own
    long double run: states after 10, 100 and 1000 steps and all 1000 step
    sizes."""
    d = np.load(os.path.join(golden_dir, "c1_traj_ld.npz"), allow_pickle=False)
    p = c1
    rk = O.RK(p, eps=1e-6, mode="ld")
    rk.set_state(d["pub0"])
    h = []
    for s in range(1, 1001):
        rk.step(1)
        h.append(rk.hdid)
        if s in (10, 100, 1000):
            assert np.array_equal(rk.get_state(), d["x%d" % s]), s
    assert np.array_equal(np.array(h), d["hdid"].astype(np.float64))


def test_precision_study_against_long_double(golden_dir, c1):
    """FP64 (the reference's <double>) and the FP32-state mode measured
    against the extended-precision trajectory on the C1 start.  Before the
    step sequences part, FP64 stays within a few ulp of extended precision and
    FP32 within float rounding; both are printed for the record."""
    d = np.load(os.path.join(golden_dir, "c1_traj_ld.npz"), allow_pickle=False)
    p = c1
    out = {}
    for mode in ("f64", "mix"):
        rk = O.RK(p, eps=1e-6, mode=mode)
        rk.set_state(d["pub0"].astype(rk.dt))
        for s in range(1, 1001):
            rk.step(1)
            if s in (10, 100, 1000):
                ref = d["x%d" % s].astype(np.float64)
                out[(mode, s)] = np.abs(rk.get_state().astype(np.float64) - ref).max() / np.abs(ref).max()
    print("relative distance to long double:",
          {"%s@%d" % k: "%.2e" % v for k, v in out.items()})
    assert out[("f64", 10)] < 1e-14
    assert out[("mix", 10)] < 1e-5
    assert out[("f64", 10)] < out[("mix", 10)]
