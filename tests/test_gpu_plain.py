"""GPU parity of the plain single-trajectory solver against the CPU oracle.

Bar (all through the C ABI):
  * integer work (violation counts, SAT results, step/attempt counters): exact
  * FP64 RHS: bit-identical to the reference's CpuSat<double> (golden fixtures)
  * FP32 RHS: bit-identical to the oracle's float mode (same operation order)
  * FP64 trajectories: states, h, time and attempts bit-identical to the oracle
    at every one of 1000 steps, with the oracle's controller pow correctly
    rounded (the GPU definition); on the committed fixture that oracle equals
    the reference's own CPU trajectory at every step, so the GPU does too
  * FP32-state trajectories: bit-identical to the oracle's mixed mode; the
    distance to the FP64 oracle (adaptive, its own step sizes) is printed at
    steps 10 and 100 and bounded by 1e-5 at step 1000
"""
import os
import sys

import numpy as np
import pytest

import analogsat_amd as A
import oracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print("[gpu-plain]", msg, flush=True)
    sys.stdout.flush()


def _internal_to_public(p, x):
    return np.asarray(x[1:], dtype=np.float64)


def _sys_with(problem_lits, problem_lens, precision):
    s = A.HipSat(precision)
    s.SetProblem(A.SatProblem(problem_lits, problem_lens))
    return s


@pytest.fixture(scope="module")
def rhs_cases(golden_dir):
    return np.load(os.path.join(golden_dir, "rhs_cases.npz"), allow_pickle=False)


def _gpu_rhs(sat, pub):
    x = sat.MakeState()
    x.CopyFrom(pub)
    d = sat.MakeState()
    sat.GetDerivatives(d, x, 0.0)
    return d.CopyTo()


@pytest.mark.parametrize("case", ["k3", "k4", "mixed", "k12", "n2000"])
def test_rhs_f64_matches_reference_golden(rhs_cases, case):
    say("rhs f64 golden case=%s" % case)
    d = rhs_cases
    sat = _sys_with(d[case + "_lits"], d[case + "_lens"], "f64")
    x = d[case + "_x"]
    got = _gpu_rhs(sat, x[1:])
    want = d[case + "_rhs"][1:]
    assert np.array_equal(got, want), np.abs(got - want).max()


@pytest.mark.parametrize("case", ["k3", "k4", "mixed", "k12", "n2000"])
def test_rhs_f32_matches_oracle_float_mode(rhs_cases, case):
    say("rhs f32 vs oracle mix case=%s" % case)
    d = rhs_cases
    sat = _sys_with(d[case + "_lits"], d[case + "_lens"], "f32")
    p = O.Problem(d[case + "_lits"], d[case + "_lens"])
    x = d[case + "_x"].astype(np.float32)
    got = _gpu_rhs(sat, x[1:])
    want = O.rhs(p, x, mode="mix")[1:]
    assert np.array_equal(got, want), np.abs(got - want).max()


def test_violations_exact_and_random_state_ranges():
    say("violations")
    p = O.fastsat(300, 1278, 3, 21)
    for prec in ("f64", "f32"):
        sat = _sys_with(p.literals, p.lengths, prec)
        rng = np.random.default_rng(5)
        for _ in range(5):
            pub = np.concatenate([rng.uniform(-1, 1, p.n), np.ones(p.m)])
            assert sat.GetClauseViolationCount(pub) == O.violations(p, pub)
        st = sat.MakeState()
        sat.SetRandomState(st, A.HipRandom(3))
        v = st.CopyTo()
        assert np.all(v[:p.n] >= -1) and np.all(v[:p.n] < 1)
        assert np.all(v[p.n:] >= 1) and np.all(v[p.n:] < 2)
        assert abs(v[:p.n].mean()) < 0.15


def _c1(golden_dir):
    g = np.load(os.path.join(golden_dir, "c1_traj.npz"), allow_pickle=False)
    p = O.fastsat(50, 213, 3, 0)
    return p, g


def test_trajectory_f64_vs_reference_golden(golden_dir):
    """FP64 GPU trajectory == oracle (correctly rounded controller pow) bit for
    bit at every step, and == the reference's own CPU trajectory wherever the
    oracle is (all 1000 steps of this fixture)."""
    say("trajectory f64 C1 1000 steps")
    p, g = _c1(golden_dir)
    sat = _sys_with(p.literals, p.lengths, "f64")
    rk = A.HipRungeKutta()
    rk.SetEpsilon(1e-6)
    rk.Configure(sat)
    old, new = sat.MakeState(), sat.MakeState()
    old.CopyFrom(g["pub0"])
    with O.pow_mode("cr"):
        orc = O.RK(p, eps=1e-6)
        orc.set_state(g["pub0"])
        oracle_matches_ref = True
        for s in range(1, 1001):
            assert rk.Step(old, new, 1)
            orc.step(1)
            st = rk.stats()
            assert st.hdid == orc.hdid and st.attempts == orc.attempts and st.time == orc.time, s
            x = old.CopyTo()
            assert np.array_equal(x, orc.get_state()), s
            assert sat.GetClauseViolationCount(old) == orc.violations(), s
            oracle_matches_ref &= orc.hdid == g["hdid"][s - 1]
            key = "state_%d" % s
            if key in g.files and oracle_matches_ref:
                assert np.array_equal(x, g[key]), s
    assert oracle_matches_ref


def test_trajectory_f32_vs_oracle_mixed_and_f64(golden_dir):
    """FP32-state GPU trajectory == oracle mixed mode bit for bit; distance to
    the FP64 oracle reported and bounded."""
    say("trajectory f32 C1 1000 steps")
    p, g = _c1(golden_dir)
    sat = _sys_with(p.literals, p.lengths, "f32")
    rk = A.HipRungeKutta()
    rk.SetEpsilon(1e-6)
    rk.Configure(sat)
    old, new = sat.MakeState(), sat.MakeState()
    pub32 = g["pub0"].astype(np.float32)
    old.CopyFrom(pub32)
    with O.pow_mode("cr"):
        om = O.RK(p, eps=1e-6, mode="mix")
        om.set_state(pub32)
        o64 = O.RK(p, eps=1e-6, mode="f64")
        o64.set_state(pub32.astype(np.float64))
        for s in range(1, 1001):
            rk.Step(old, new, 1)
            om.step(1)
            o64.step(1)
            x = old.CopyTo()
            assert np.array_equal(x, om.get_state()), s
            assert rk.stats().hdid == om.hdid, s
            if s in (10, 100, 1000):
                ref = o64.get_state()
                rel = np.abs(x - ref).max() / np.abs(ref).max()
                say("step %d: max rel err vs FP64 oracle = %.3e (h %s)" % (s, rel, rk.stats().hdid == o64.hdid))
                # The two runs choose their own step sizes, so before they
                # settle step s of one is at another analog time than step s
                # of the other (5.7e-5 apart at step 100 on this start); at
                # step 1000 both sit on the same fixed point.  The matched-
                # step bound is pinned with frozen step sizes in
                # test_fp32_with_fp64_steps_within_1e5_of_fp64_oracle.
                if s == 1000:
                    assert rel <= 1e-5, rel


def test_satsolver_loop_matches_reference(golden_dir):
    say("SatSolver f64 C1 solve")
    p, g = _c1(golden_dir)
    d = np.load(os.path.join(golden_dir, "c1_solve.npz"), allow_pickle=False)
    sat = A.HipSat("f64")
    rk = A.HipRungeKutta()
    rk.SetEpsilon(1e-6)
    solver = A.SatSolver(sat, rk)
    solver.SetProblem(A.SatProblem(p.literals, p.lengths))
    solver.SetMaxTime(1e8)
    solver.SetMaxSteps(10000)
    solver.SetBatchSize(1)
    solver.SetSatState(g["pub0"])
    seen = []
    solver.SetCallback(lambda s: seen.append(s.GetLastClauseViolationCount()) or True)
    res = solver.Solve()
    assert int(res) == int(d["result"])
    assert solver.GetStepCount() == int(d["steps"])
    assert solver.GetRhsCount() == int(d["rhscount"])
    assert np.array_equal(np.array(seen), d["batch_violations"])
    assert np.array_equal(solver.GetSatState().CopyTo(), d["final"])


def test_step_count_semantics_and_reset():
    say("step semantics")
    p = O.fastsat(200, 852, 3, 9)
    sat = _sys_with(p.literals, p.lengths, "f32")
    rk = A.HipRungeKutta()
    rk.SetEpsilon(1e-6)
    rk.Configure(sat)
    old, new = sat.MakeState(), sat.MakeState()
    pub = O.random_state_f64(4, p.n, p.m).astype(np.float32)
    old.CopyFrom(pub)
    for c in (1, 7, 50):
        before = rk.GetStepCount()
        assert rk.Step(old, new, c)
        assert rk.GetStepCount() - before == c
        st = rk.stats()
        assert st.rhscount == st.steps + 5 * st.attempts
    rk.Reset()
    st = rk.stats()
    assert (st.steps, st.attempts, st.rhscount, st.time) == (0, 0, 0, 0.0)
    assert st.hnext == 1e-9


def test_kernel_timer_and_stream_probe():
    """Measurement hooks: the kernel timer times one launch of the chosen
    kind per enqueued attempt and does not change the trajectory; the
    stream probe reports positive rates."""
    say("kernel timer and stream probe")
    p = O.fastsat(3000, 12600, 3, 21)
    pub = O.random_state_f64(5, p.n, p.m).astype(np.float32)
    states = []
    for kind in (-1, 5, 11):
        sat = _sys_with(p.literals, p.lengths, "f32")
        rk = A.HipRungeKutta()
        rk.SetEpsilon(1e-6)
        rk.Configure(sat)
        old, new = sat.MakeState(), sat.MakeState()
        old.CopyFrom(pub)
        rk.set_kernel_timer(kind)
        assert rk.Step(old, new, 30)
        ms, n = rk.kernel_timer()
        if kind < 0:
            assert (ms, n) == (0.0, 0)
        else:
            assert n == rk.GetAttemptCount() and ms > 0.0
        states.append(old.CopyTo())
    assert np.array_equal(states[0], states[1]) and np.array_equal(states[0], states[2])
    with pytest.raises(ValueError):
        A.HipRungeKutta().set_kernel_timer(13)
    bw = A.stream_bench(64 << 20, 2)
    assert bw["copy_GBps"] > 0 and bw["read_GBps"] > 0


def test_step_log_records_accepted_steps():
    """asat_rk_set_step_log: after the k-th accepted step since Reset the
    device log holds its step size, for Step calls of any size (single
    attempts and captured groups of attempts alike)."""
    torch = pytest.importorskip("torch")
    import ctypes
    say("step log")
    p = O.fastsat(200, 852, 3, 9)
    sat = _sys_with(p.literals, p.lengths, "f64")
    rk = A.HipRungeKutta()
    rk.SetEpsilon(1e-6)
    rk.Configure(sat)
    old, new = sat.MakeState(), sat.MakeState()
    old.CopyFrom(O.random_state_f64(8, p.n, p.m))
    log = torch.full((60,), -1.0, dtype=torch.float64, device="cuda")
    assert A.lib().asat_rk_set_step_log(rk._h, ctypes.c_void_p(log.data_ptr()), 50) == 0
    rk.Reset()
    seen = []
    for c in (1, 1, 3, 20, 30):
        assert rk.Step(old, new, c)
        seen.append(rk.stats().hdid)
    h = log.cpu().numpy()
    assert h[0] == seen[0] and h[1] == seen[1] and h[4] == seen[2] and h[24] == seen[3]
    assert np.all(h[:50] > 0) and np.all(h[50:] == -1.0)   # 55 accepted steps, 50 logged
    assert A.lib().asat_rk_set_step_log(rk._h, None, 0) == 0


def test_large_instance_rhs_property():
    """Standalone RHS (spin-gather clause pass) at N = 2e5, M = 8.52e5: GPU
    FP32 equals the oracle's float mode bit for bit.  The benchmark's own
    size (N = 1e6) and its slot-spin integrator path are checked in
    test_gpu_full_size.py."""
    say("rhs f32 N=200000 M=852000")
    n, m = 200000, 852000
    p = O.fastsat(n, m, 3, 1)
    sat = _sys_with(p.literals, p.lengths, "f32")
    u = O.random_state_f64(7, n, m).astype(np.float32)
    got = _gpu_rhs(sat, u)
    x = np.concatenate([[-1.0], u]).astype(np.float32)
    want = O.rhs(p, x, mode="mix")[1:]
    assert np.array_equal(got, want)


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_trajectory_multi_group_bitexact(precision):
    """N = 30000 spans several variable groups (XCD-affine contribution
    regions): 60 steps bit-identical to the oracle."""
    say("trajectory %s N=30000 60 steps" % precision)
    n, m = 30000, 127800
    p = O.fastsat(n, m, 3, 77)
    mode = "f64" if precision == "f64" else "mix"
    dt = np.float64 if precision == "f64" else np.float32
    pub = np.concatenate([np.random.default_rng(3).uniform(-1, 1, n), np.ones(m)]).astype(dt)
    sat = _sys_with(p.literals, p.lengths, precision)
    rk = A.HipRungeKutta()
    rk.SetEpsilon(1e-6)
    rk.Configure(sat)
    old, new = sat.MakeState(), sat.MakeState()
    old.CopyFrom(pub)
    with O.pow_mode("cr"):
        orc = O.RK(p, eps=1e-6, mode=mode)
        orc.set_state(pub)
        for blk in range(6):
            rk.Step(old, new, 10)
            orc.step(10)
            assert np.array_equal(old.CopyTo(), orc.get_state()), blk
            st = rk.stats()
            assert (st.hdid, st.attempts, st.rhscount) == (orc.hdid, orc.attempts, orc.rhscount)
    assert sat.GetClauseViolationCount(old) == orc.violations()


def _mixed_problem(n, m, seed):
    """Clauses of 1, 2 or 3 distinct literals (padding slots in the 3-wide rows)."""
    rng = np.random.default_rng(seed)
    lens = rng.choice([1, 2, 3], size=m, p=[0.1, 0.3, 0.6]).astype(np.int32)
    lits = []
    for L in lens:
        v = rng.choice(np.arange(1, n + 1), size=L, replace=False)
        lits.extend(np.where(rng.random(L) < 0.5, -v, v))
    return O.Problem(np.asarray(lits, np.int32), lens)


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("case", ["k4", "mixed"])
def test_trajectory_k4_and_padding_bitexact(precision, case):
    """The tiled clause pass with slot spins for 4-literal clauses and for
    rows with padding slots, over several tiles and groups: 30 steps
    bit-identical to the oracle, also after a standalone RHS (which reuses
    the contribution array) and after the state is replaced between Step
    calls (the slot spins are rebuilt from the new state)."""
    say("trajectory %s %s 3 x 10 steps" % (precision, case))
    if case == "k4":
        n, m = 3000, 29700
        p = O.fastsat(n, m, 4, 5)
    else:
        p = _mixed_problem(3000, 9000, 6)
    n, m = p.n, p.m
    mode = "f64" if precision == "f64" else "mix"
    dt = np.float64 if precision == "f64" else np.float32
    pub = O.random_state_f64(9, n, m).astype(dt)
    sat = _sys_with(p.literals, p.lengths, precision)
    rk = A.HipRungeKutta()
    rk.SetEpsilon(1e-6)
    rk.Configure(sat)
    old, new = sat.MakeState(), sat.MakeState()
    old.CopyFrom(pub)
    with O.pow_mode("cr"):
        orc = O.RK(p, eps=1e-6, mode=mode)
        orc.set_state(pub)
        for blk in range(3):
            if blk == 1:
                d = sat.MakeState()
                sat.GetDerivatives(d, old, 0.0)
            if blk == 2:
                moved = old.CopyTo()
                moved[:n] = np.clip(moved[:n] * dt(0.5), -1, 1)
                old.CopyFrom(moved)
                orc.set_state(moved)
            rk.Step(old, new, 10)
            orc.step(10)
            assert np.array_equal(old.CopyTo(), orc.get_state()), blk
            st = rk.stats()
            assert (st.hdid, st.attempts) == (orc.hdid, orc.attempts), blk
    assert sat.GetClauseViolationCount(old) == orc.violations()


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_trajectory_with_error_floor_bitexact(golden_dir, precision):
    """A raised xscal floor (SetErrorFloor) changes the error norm the same
    way on the GPU as in the oracle: 300 steps bit-identical."""
    say("trajectory %s C1 floor=1e-3 300 steps" % precision)
    p, g = _c1(golden_dir)
    mode = "f64" if precision == "f64" else "mix"
    dt = np.float64 if precision == "f64" else np.float32
    sat = _sys_with(p.literals, p.lengths, precision)
    rk = A.HipRungeKutta()
    rk.SetEpsilon(1e-6)
    rk.SetErrorFloor(1e-3)
    rk.Configure(sat)
    old, new = sat.MakeState(), sat.MakeState()
    pub = g["pub0"].astype(dt)
    old.CopyFrom(pub)
    with O.pow_mode("cr"):
        orc = O.RK(p, eps=1e-6, mode=mode, floor=1e-3)
        orc.set_state(pub)
        for blk in range(30):
            rk.Step(old, new, 10)
            orc.step(10)
            assert np.array_equal(old.CopyTo(), orc.get_state()), blk
            st = rk.stats()
            assert (st.hdid, st.time, st.attempts) == (orc.hdid, orc.time, orc.attempts), blk


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_high_degree_variable_uses_gather_pass(precision):
    """A variable in every clause has more contributions than one LDS region
    holds: the var pass falls back to gathering from global memory.  RHS and
    20 steps stay bit-identical to the oracle."""
    say("high-degree variable %s" % precision)
    rng = np.random.default_rng(12)
    n, m = 2000, 45000
    lits = np.empty((m, 3), np.int32)
    lits[:, 0] = 1
    lits[:, 1:] = rng.integers(2, n + 1, (m, 2))
    lits *= np.where(rng.random((m, 3)) < 0.5, -1, 1).astype(np.int32)
    p = O.Problem(lits.ravel(), np.full(m, 3, np.int32))
    mode = "f64" if precision == "f64" else "mix"
    dt = np.float64 if precision == "f64" else np.float32
    sat = _sys_with(p.literals, p.lengths, precision)
    pub = O.random_state_f64(3, n, m).astype(dt)
    got = _gpu_rhs(sat, pub)
    want = O.rhs(p, np.concatenate([[-1.0], pub]).astype(dt), mode=mode)[1:]
    assert np.array_equal(got, want)
    rk = A.HipRungeKutta()
    rk.SetEpsilon(1e-6)
    rk.Configure(sat)
    old, new = sat.MakeState(), sat.MakeState()
    old.CopyFrom(pub)
    with O.pow_mode("cr"):
        orc = O.RK(p, eps=1e-6, mode=mode)
        orc.set_state(pub)
        rk.Step(old, new, 20)
        orc.step(20)
    assert np.array_equal(old.CopyTo(), orc.get_state())
    assert rk.stats().attempts == orc.attempts


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_bias_term_rhs_and_trajectory(rhs_cases, precision):
    """b > 0 (AnalogSAT/cpu_sat/CpuSat.cpp:232-239).  The GPU sums a in double
    over a tree and rounds once (the oracle's mean_mode('double')); the
    reference sums serially in the state type.  RHS: within 16 ulp of the
    largest |term| of the reference (FP64 fixture) and of the oracle in both
    mean modes (FP32).  Trajectory, 30 steps: same attempts; the difference in
    the mean and in sin is a persistent rounding-level change of the vector
    field, which the flow integrates (on the CPU, switching only the FP32
    mean's accumulation moves this trajectory by 3.6e-4), so the bound is
    1e-9 (FP64) / 1e-3 (FP32) relative."""
    say("bias b=2.5 %s" % precision)
    d = rhs_cases
    p = O.Problem(d["bias_lits"], d["bias_lens"])
    b = float(d["bias_b"])
    dt = np.float64 if precision == "f64" else np.float32
    mode = "f64" if precision == "f64" else "mix"
    sat = _sys_with(d["bias_lits"], d["bias_lens"], precision)
    sat.Set_B(b)
    x = d["bias_x"]
    got = _gpu_rhs(sat, x[1:].astype(dt)).astype(np.float64)
    ulp = 2.0 ** -52 if precision == "f64" else 2.0 ** -23
    scale = np.abs(d["bias_rhs"]).max() + np.abs(O.rhs(p, x, b=0.0)[1:]).max()
    wants = [d["bias_rhs"][1:]] if precision == "f64" else []
    for mm in ("serial", "double"):
        with O.mean_mode(mm):
            wants.append(O.rhs(p, x.astype(dt), mode=mode, b=b)[1:].astype(np.float64))
    for w in wants:
        assert np.abs(got - w).max() <= 16 * ulp * scale
    rk = A.HipRungeKutta()
    rk.SetEpsilon(1e-6)
    rk.Configure(sat)
    old, new = sat.MakeState(), sat.MakeState()
    old.CopyFrom(x[1:].astype(dt))
    rk.Step(old, new, 30)
    g = old.CopyTo().astype(np.float64)
    for mm in ("serial", "double"):
        with O.pow_mode("cr"), O.mean_mode(mm):
            orc = O.RK(p, eps=1e-6, mode=mode, b=b)
            orc.set_state(x[1:].astype(dt))
            orc.step(30)
        ref = orc.get_state().astype(np.float64)
        rel = np.abs(g - ref).max() / np.abs(ref).max()
        say("  30 steps vs oracle (mean %s): rel diff %.3e, attempts %d vs %d"
            % (mm, rel, rk.stats().attempts, orc.attempts))
        assert rk.stats().attempts == orc.attempts
        assert rel <= (1e-9 if precision == "f64" else 1e-3)


def test_new_problem_invalidates_states_and_integrator():
    """After a second SetProblem, states and an integrator made for the first
    problem are rejected, even when the second problem has the same buffer
    sizes (N = 100, M = 200 and N = 127, M = 250 both take 384 elements)."""
    say("problem generation")
    p1, p2 = O.fastsat(100, 200, 3, 1), O.fastsat(127, 250, 3, 2)
    sat = _sys_with(p1.literals, p1.lengths, "f64")
    rk = A.HipRungeKutta()
    rk.SetEpsilon(1e-6)
    rk.Configure(sat)
    old, new = sat.MakeState(), sat.MakeState()
    old.CopyFrom(O.random_state_f64(1, p1.n, p1.m))
    assert rk.Step(old, new, 3)
    assert old.GetAllocSize() == 384
    sat.SetProblem(A.SatProblem(p2.literals, p2.lengths))
    fresh = sat.MakeState()
    assert fresh.GetAllocSize() == 384
    with pytest.raises(RuntimeError, match="changed since"):
        rk.Step(old, new, 1)
    with pytest.raises(ValueError, match="current problem"):
        sat.GetClauseViolationCount(old)
    rk.Configure(sat)
    a, b = sat.MakeState(), sat.MakeState()
    a.CopyFrom(O.random_state_f64(2, p2.n, p2.m))
    assert rk.Step(a, b, 3)


@pytest.mark.parametrize("n, m", [(310000, 1300000), (520000, 2200000)])
def test_big_tile_sizes_bitexact(n, m):
    """FP32 3-literal systems large enough to give every resident workgroup
    (two per CU) at least 2048 clauses use tiles sized per problem to fill
    whole rounds (up to 4608 clauses, five for some threads; the kernel
    takes a thread's clauses one at a time), rebuild their sorted positions
    from per-tile run tables and carry 16-bit primary words.  On 256 CUs
    these get tiles of 2560 and 4352 clauses, not multiples of the
    workgroup size.  The standalone RHS and 12 integrator steps
    bit-identical to the oracle's mixed mode."""
    say("big tiles N=%d M=%d" % (n, m))
    p = O.fastsat(n, m, 3, 31)
    sat = _sys_with(p.literals, p.lengths, "f32")
    lay = sat.Layout()
    say("  layout %s" % lay)
    assert lay["direct_primary"] == 1 and lay["tile_clauses"] % 64 == 0 and 2048 < lay["tile_clauses"] <= 4608
    assert lay["tiles"] * lay["tile_clauses"] >= m
    assert lay["run_tables"] == 1 and lay["primary_words"] == 1
    assert 4 * (lay["groups"] + 1) <= 2 * lay["tile_clauses"]
    if lay["tile_clauses"] % 1024 == 0:   # the size follows the CU count (256 CUs here)
        pytest.skip("this device sizes the tiles to a multiple of the workgroup")
    _rhs_and_steps_bitexact(sat, p, n, m)


def test_far_apart_primaries_keep_2048_clause_tiles():
    """N = 2.2e6, M = 1.1e6 (FP32, 3 literals): big enough for tiles of 2176
    clauses, but some of those tiles would span up to about 18 000 primary
    variables (half of them about 3 400), more than a 16-bit primary word's
    13 relative-variable bits hold; its 1280 groups would also make a run
    table larger than a quarter of the positions.  The system keeps
    2048-clause tiles with sorted positions, a rank for every slot and
    4-byte primary variables.  The standalone RHS and 12 integrator steps
    bit-identical to the oracle's mixed mode."""
    n, m = 2200000, 1100000
    say("far-apart primaries N=%d M=%d" % (n, m))
    p = O.fastsat(n, m, 3, 31)
    sat = _sys_with(p.literals, p.lengths, "f32")
    lay = sat.Layout()
    say("  layout %s" % lay)
    assert lay["direct_primary"] == 1 and lay["tile_clauses"] == 2048 and lay["groups"] == 1280
    assert lay["primary_words"] == 0 and lay["run_tables"] == 0
    _rhs_and_steps_bitexact(sat, p, n, m)


def test_small_systems_keep_2048_clause_tiles():
    """N = 2000, M = 8500 and N = 140000, M = 6e5 (FP32, 3 literals): too
    few clauses for 2048 per resident workgroup, so 2048-clause tiles (the
    per-problem size would be 1024 and 1216); RHS and 12 steps
    bit-identical to the oracle's mixed mode."""
    for n, m in ((2000, 8500), (140000, 600000)):
        say("small system N=%d M=%d" % (n, m))
        p = O.fastsat(n, m, 3, 33)
        sat = _sys_with(p.literals, p.lengths, "f32")
        lay = sat.Layout()
        assert lay["direct_primary"] == 1 and lay["tile_clauses"] == 2048, lay
        assert lay["primary_words"] == 0 and lay["run_tables"] == 0
        _rhs_and_steps_bitexact(sat, p, n, m)


def _rhs_and_steps_bitexact(sat, p, n, m):
    # N is the largest variable the clauses name (SatProblem), which random
    # instances with M << N can leave below the generator's N
    n, m = p.n, p.m
    assert (sat.Get_N(), sat.Get_M()) == (n, m)
    pub = np.concatenate([np.random.default_rng(8).uniform(-1, 1, n), np.ones(m)]).astype(np.float32)
    x, d = sat.MakeState(), sat.MakeState()
    x.CopyFrom(pub)
    sat.GetDerivatives(d, x, 0.0)
    want = O.rhs(p, np.concatenate([[-1.0], pub]).astype(np.float32), mode="mix")[1:]
    got = d.CopyTo()
    bad = np.flatnonzero(got != want)
    assert bad.size == 0, ("%d spin / %d aux entries differ; first %s; got %s want %s"
                           % ((bad < n).sum(), (bad >= n).sum(), bad[:6], got[bad[:3]], want[bad[:3]]))
    rk = A.HipRungeKutta()
    rk.SetEpsilon(1e-6)
    rk.Configure(sat)
    old, new = sat.MakeState(), sat.MakeState()
    old.CopyFrom(pub)
    with O.pow_mode("cr"):
        orc = O.RK(p, eps=1e-6, mode="mix")
        orc.set_state(pub)
        for blk in range(3):
            rk.Step(old, new, 4)
            orc.step(4)
            assert np.array_equal(old.CopyTo(), orc.get_state()), blk
            st = rk.stats()
            assert (st.hdid, st.attempts, st.rhscount) == (orc.hdid, orc.attempts, orc.rhscount)
    assert sat.GetClauseViolationCount(old) == orc.violations()
