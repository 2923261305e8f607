"""Pin the oracle against mathematics that fixes its answers (CPU only).

These checks are independent of the reference build: closed forms of the
CTDS right-hand side (AnalogSAT/cpu_sat/CpuSat.cpp:197-241), the gradient
identity f_s = -dV/ds for V = sum_m a_m K_m^2, an exactly solvable one-clause
system for the Cash-Karp tableau and controller
(AnalogSAT/cpu_sat/CpuRungeKutta.cpp:180-267), brute-force violation counts,
and the SatSolver loop's limit semantics (AnalogSAT/solver/SatSolver.h:317-339).
They also pin the mixed (float state / double control) mode, which the
reference does not have.
"""
import numpy as np
import pytest

import oracle as O


def internal(p, s, a):
    return np.concatenate([[-1.0], np.asarray(s, float), np.asarray(a, float)])


def energy(p, x):
    """V = sum_m a_m K_m^2 with K_m = 2^-len prod (1 - c s), computed directly from the clauses."""
    V = 0.0
    off = 0
    for j, L in enumerate(p.lengths):
        lits = p.literals[off:off + L]
        off += L
        lits = lits[lits != 0]
        K = 2.0 ** (-p.k)
        for lit in lits:
            c = 1.0 if lit > 0 else -1.0
            K *= 1.0 - c * x[abs(lit)]
        K *= 2.0 ** (p.k - len(lits))          # padding literals contribute a factor 2
        V += x[p.n + 1 + j] * K * K
    return V


def test_zero_spins_closed_form():
    p = O.fastsat(30, 120, 3, 5)
    a = 1.0 + np.arange(p.m) / p.m
    x = internal(p, np.zeros(p.n), a)
    d = O.rhs(p, x)
    assert np.array_equal(d[p.n + 1:], a * 4.0 ** -3)
    want = np.zeros(p.n + 1)
    C = p.C.reshape(p.m, p.k)
    for j in range(p.m):
        for lit in C[j]:
            want[abs(lit)] += 2 * a[j] * np.sign(lit) * 4.0 ** -3
    np.testing.assert_allclose(d[1:p.n + 1], want[1:], rtol=1e-14, atol=0)
    assert d[0] == 0.0


def test_satisfied_literal_kills_clause():
    p = O.Problem.from_clauses([[1, -2, 3]])
    for s1, s2, s3 in ([1.0, 0.3, -0.2], [0.4, -1.0, 0.9], [-0.5, 0.5, 1.0]):
        d = O.rhs(p, internal(p, [s1, s2, s3], [1.7]))
        assert np.all(d == 0.0)


def test_padding_factor_two():
    # a 2-literal clause inside a k=3 problem has K = 2^-2 (1-c1 s1)(1-c2 s2)
    p = O.Problem.from_clauses([[1, -2], [1, 2, 3]])
    assert p.k == 3
    s = np.array([0.25, -0.5, 0.125])
    x = internal(p, s, [1.5, 2.0])
    d = O.rhs(p, x)
    K0 = 0.25 * (1 - s[0]) * (1 + s[1])
    assert d[p.n + 1] == pytest.approx(1.5 * K0 * K0, rel=1e-15)


def test_gradient_identity_fd():
    p = O.fastsat(25, 100, 3, 8)
    rng = np.random.default_rng(0)
    x = internal(p, rng.uniform(-1, 1, p.n), rng.uniform(1, 2, p.m))
    d = O.rhs(p, x)
    h = 1e-6
    for i in range(1, p.n + 1):
        xp, xm = x.copy(), x.copy()
        xp[i] += h
        xm[i] -= h
        g = (energy(p, xp) - energy(p, xm)) / (2 * h)
        assert d[i] == pytest.approx(-g, rel=1e-6, abs=1e-10)
    # f_a = a K^2 exactly = V restricted to one clause
    for j in range(p.m):
        y = x.copy()
        y[p.n + 1:] = 0.0
        y[p.n + 1 + j] = x[p.n + 1 + j]
        assert d[p.n + 1 + j] == pytest.approx(energy(p, y), rel=1e-14)


def test_violations_bruteforce_and_planted():
    rng = np.random.default_rng(1)
    p = O.fastsat(12, 60, 3, 2)
    C = p.C.reshape(p.m, p.k)
    for _ in range(20):
        s = rng.uniform(-1, 1, p.n)
        truth = s > 0
        want = 0
        for row in C:
            sat = any((truth[abs(l) - 1] if l > 0 else not truth[abs(l) - 1]) for l in row)
            want += not sat
        assert O.violations(p, np.concatenate([s, np.ones(p.m)])) == want
        assert O.violations(p, np.concatenate([s, np.ones(p.m)]), mode="mix") == want
    # planted: clauses built to satisfy a fixed assignment
    plant = rng.random(20) < 0.5
    clauses = []
    while len(clauses) < 80:
        v = rng.choice(np.arange(1, 21), 3, replace=False)
        lits = [int(x) if rng.random() < 0.5 else -int(x) for x in v]
        if any((plant[abs(l) - 1] if l > 0 else not plant[abs(l) - 1]) for l in lits):
            clauses.append(lits)
    pp = O.Problem.from_clauses(clauses)
    s = np.where(plant, 0.5, -0.5)[: pp.n]
    assert O.violations(pp, np.concatenate([s, np.ones(pp.m)])) == 0


def test_mixed_mode_rhs_matches_f64_to_float_precision():
    p = O.fastsat(200, 850, 3, 3)
    rng = np.random.default_rng(2)
    x = internal(p, rng.uniform(-1, 1, p.n), rng.uniform(1, 2, p.m)).astype(np.float32).astype(np.float64)
    d64 = O.rhs(p, x)
    d32 = O.rhs(p, x, mode="mix").astype(np.float64)
    scale = np.abs(d64).max()
    assert np.abs(d32 - d64).max() <= 2e-6 * scale


def _one_clause_exact(u0, a0, t):
    """One clause (x1), k=1: u = 1-s obeys w=u^2, w' = -C w + w^2/4 with C = a + u^2/4 conserved."""
    C = a0 + u0 * u0 / 4.0
    w = 1.0 / (1.0 / (4 * C) + (1.0 / (u0 * u0) - 1.0 / (4 * C)) * np.exp(C * t))
    return w


@pytest.mark.parametrize("mode", ["f64", "mix"])
def test_tableau_and_controller_on_exact_solution(mode):
    p = O.Problem.from_clauses([[1]])
    s0, a0 = -0.5, 1.25
    eps = 1e-10 if mode == "f64" else 1e-6
    rk = O.RK(p, eps=eps, mode=mode)
    rk.set_state([s0, a0])
    for _ in range(40):
        assert rk.step(5)
        s, a = rk.get_state().astype(np.float64)
        w = _one_clause_exact(1 - s0, a0, rk.time)
        if mode == "f64":
            assert (1 - s) ** 2 == pytest.approx(w, rel=1e-8)
            assert a + (1 - s) ** 2 / 4 == pytest.approx(a0 + (1 - s0) ** 2 / 4, rel=1e-9)
        else:
            # float state: compare s itself (1 - s cancels near the fixed point s = 1)
            assert s == pytest.approx(1 - np.sqrt(w), abs=3e-6)
            assert a + (1 - s) ** 2 / 4 == pytest.approx(a0 + (1 - s0) ** 2 / 4, rel=3e-6)


def test_initial_step_growth_is_five_fold():
    # from h0 = 1e-9 the error is far below ERRCON, so hdid grows 5x per step
    p = O.fastsat(50, 213, 3, 0)
    pub = O.random_state_f64(1, p.n, p.m)
    pub[p.n:] = 1.0
    rk = O.RK(p, eps=1e-6)
    rk.set_state(pub)
    h = 1e-9
    for _ in range(4):
        rk.step(1)
        assert rk.hdid == h
        h = 5.0 * h
    assert rk.attempts == 4 and rk.rhscount == 24


def test_solve_limits_checked_between_batches():
    p = O.fastsat(60, 270, 3, 4)
    pub = O.random_state_f64(2, p.n, p.m)
    pub[p.n:] = 1.0
    rk = O.RK(p, eps=1e-6)
    rk.set_state(pub)
    res, viol = rk.solve(max_time=1e8, max_steps=1, batch=20)
    # maxSteps is compared only after a whole batch
    assert res in (0, 4) and rk.steps == 20 and len(viol) == 1
    rk.set_state(pub)
    res, viol = rk.solve(max_time=1e-12, max_steps=100000, batch=7)
    assert res in (0, 3) and rk.steps == 7


def test_correctly_rounded_controller_pow():
    """orc_pow_cr (the GPU controller's pow) against 200-bit mpmath, nearest rounding."""
    mpmath = pytest.importorskip("mpmath")
    mpmath.mp.prec = 200
    rng = np.random.default_rng(3)
    xs = np.concatenate([10 ** rng.uniform(-4, 0.3, 2500), 10 ** rng.uniform(0, 8, 500), [1.0, 1.89e-4, 2.0]])
    for q, y in ((4, -0.25), (5, -0.2)):
        for x in xs:
            t = mpmath.power(mpmath.mpf(float(x)), mpmath.mpf(y))
            want = mpmath.libmp.to_float(t._mpf_, rnd="n")
            assert O.pow_cr(x, q) == want, (x, q)


def test_pow_mode_only_changes_controller(golden_dir):
    """With the correctly rounded pow the oracle still follows the reference
    trajectory exactly up to the first step whose controller pow glibc misrounds."""
    import os
    g = np.load(os.path.join(golden_dir, "c1_traj.npz"), allow_pickle=False)
    p = O.fastsat(50, 213, 3, 0)
    with O.pow_mode("cr"):
        rk = O.RK(p, eps=1e-6)
        rk.set_state(g["pub0"])
        first = None
        for s in range(1, 1001):
            rk.step(1)
            if rk.hdid != g["hdid"][s - 1] and first is None:
                first = s
                break
    # the first 12 steps never call pow (5x growth); afterwards a difference may
    # only come from a glibc misrounding, which is rare
    assert first is None or first > 12



# Cash & Karp (1990) embedded 4(5) tableau, written out independently of the
# oracle's C code (which follows AnalogSAT/cpu_sat/CpuRungeKutta.cpp:180-215).
_CK_B = [[], [1 / 5], [3 / 40, 9 / 40], [3 / 10, -9 / 10, 6 / 5], [-11 / 54, 5 / 2, -70 / 27, 35 / 27],
         [1631 / 55296, 175 / 512, 575 / 13824, 44275 / 110592, 253 / 4096]]
_CK_C5 = np.array([37 / 378, 0, 250 / 621, 125 / 594, 0, 512 / 1771])
_CK_C4 = np.array([2825 / 27648, 0, 18575 / 48384, 13525 / 55296, 277 / 14336, 1 / 4])


def _ck_errmax(p, x, h, h_scale, eps, floor):
    """max |xerr / xscal| / eps of one Cash-Karp attempt of size h from x, with
    xscal = |x| + |h_scale k1| + floor (h_scale: the step size at step start)."""
    ks = []
    for row in _CK_B:
        y = x + h * sum(b * k for b, k in zip(row, ks)) if row else x
        ks.append(O.rhs(p, y))
    err = h * sum((c5 - c4) * k for c5, c4, k in zip(_CK_C5, _CK_C4, ks))
    scale = np.abs(x) + np.abs(h_scale * ks[0]) + floor
    return np.max(np.abs(err / scale)) / eps


@pytest.mark.parametrize("floor", [1e-20, 1e-6, 1e-3, 0.5])
def test_error_scale_floor(floor):
    """The constant term of xscal (TINY = 1e-20 in
    AnalogSAT/cpu_sat/CpuRungeKutta.cpp:67-76) enters the error norm as
    |x| + |h k1| + floor: the accepted attempt's errmax equals an independent
    Cash-Karp evaluation, and raising the floor can only lower it."""
    p = O.fastsat(60, 255, 3, 5)
    pub = O.random_state_f64(2, p.n, p.m)
    pub[: p.n : 3] = 0.0          # spins at zero: there xscal is |h k1| + floor
    rk = O.RK(p, eps=1e-6, floor=floor)
    rk.set_state(pub)
    checked = 0
    for _ in range(14):
        x0 = np.concatenate([[-1.0], rk.get_state()])
        h_scale, att = rk.hnext, rk.attempts
        rk.step(1)
        want = _ck_errmax(p, x0, rk.hdid, h_scale, 1e-6, floor)
        assert rk.errmax == pytest.approx(want, rel=1e-9, abs=1e-300)
        assert _ck_errmax(p, x0, rk.hdid, h_scale, 1e-6, floor * 100) <= want
        checked += rk.attempts - att
    assert rk.hdid > 1e-4 and checked >= 14


def test_bias_term_and_mean_modes():
    """The bias term adds mean(a) (M/N) b 4^-k (pi/2) sin(pi s_i) to ds_i
    (AnalogSAT/cpu_sat/CpuSat.cpp:232-239): the b > 0 minus b = 0 difference
    equals that expression.  In FP64 the two mean modes coincide; in FP32 the
    double-accumulated mean is the float-rounded exact sum."""
    p = O.fastsat(60, 255, 3, 8)
    pub = O.random_state_f64(5, p.n, p.m)
    x = np.concatenate([[-1.0], pub])
    b = 1.75
    d = O.rhs(p, x, b=b) - O.rhs(p, x, b=0.0)
    a = x[p.n + 1:]
    want = a.mean() * (p.m / p.n) * b / 64.0 * (np.pi / 2) * np.sin(np.pi * x[1:p.n + 1])
    np.testing.assert_allclose(d[1:p.n + 1], want, rtol=1e-10, atol=1e-14)
    assert np.all(d[p.n + 1:] == 0)
    with O.mean_mode("double"):
        d64 = O.rhs(p, x, b=b)
    assert np.array_equal(d64, O.rhs(p, x, b=b))
    x32 = x.astype(np.float32)
    with O.mean_mode("double"):
        r_dbl = O.rhs(p, x32, mode="mix", b=b)
    r0 = O.rhs(p, x32, mode="mix", b=0.0)
    f32 = np.float32
    mean32 = f32(np.sum(x32[p.n + 1:].astype(np.float64))) / f32(p.m)
    pref = mean32 * (f32(p.m) / f32(p.n)) * (f32(b) * f32(0.125) * f32(0.125)) * f32(1.570796326794897)
    want32 = pref * np.sin(f32(np.pi) * x32[1:p.n + 1])
    np.testing.assert_allclose(r_dbl[1:p.n + 1] - r0[1:p.n + 1], want32, rtol=2e-5, atol=2e-6)
