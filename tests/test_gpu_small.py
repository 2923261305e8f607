"""GPU parity of the one-workgroup persistent path of the plain solver.

Problems whose whole integrator fits in one workgroup's LDS run Step and
SatSolver's loop (no callback) in one persistent kernel.  Bar:
  * states, h, time, errmax and every counter bit-identical to the
    multi-kernel path (forced with SetOneWorkgroup(False)) on the same
    problem and start, FP64 and FP32, across several Step calls, for K = 3,
    K = 4, mixed clause lengths with padding, K = 12, the tanh family and a
    problem with one variable in every clause (its term lists stay ragged:
    padding every variable's run to the hub's length would cost more than
    twice the terms; the others take uniform runs padded with -0.0)
  * the device-side SatSolver loop gives the reference's own solve of the C1
    fixture (result, steps, RHS count, final state: AnalogSAT/solver/SatSolver.h:313-355)
    and, on problems that end by SAT, by the step limit and by the time
    limit, exactly what the host loop gives
  * FP64 solves of 20 C1 trajectories equal the FP64 oracle's SatSolver loop
    (result, steps, RHS count, time, final state)
"""
import os
import sys

import numpy as np
import pytest

import analogsat_amd as A
import oracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print("[gpu-small]", msg, flush=True)
    sys.stdout.flush()


def _problem(case):
    if case == "k3":
        return O.fastsat(50, 213, 3, 0)
    if case == "k4":
        return O.fastsat(40, 400, 4, 3)
    if case == "hub":
        rng = np.random.default_rng(13)
        lits = []
        for _ in range(200):
            v = np.concatenate([[1], rng.choice(np.arange(2, 41), size=2, replace=False)])
            lits.extend((v * rng.choice([-1, 1], size=3)).tolist())
        return O.Problem(np.array(lits, np.int32), np.full(200, 3, np.int32))
    rng = np.random.default_rng(11 if case == "mixed" else 12)
    n, m = (60, 250) if case == "mixed" else (40, 90)
    lens = rng.integers(1, 6, m) if case == "mixed" else np.full(m, 12)
    lits = []
    for L in lens:
        v = rng.choice(n, size=int(L), replace=False) + 1
        lits.extend((v * rng.choice([-1, 1], size=int(L))).tolist())
    lits[0] = n if lits[0] > 0 else -n   # make sure variable n appears
    return O.Problem(np.array(lits, np.int32), lens.astype(np.int32))


def _start(p, seed, dtype):
    rng = np.random.default_rng(seed)
    return np.concatenate([rng.uniform(-1.0, 1.0, p.n), np.ones(p.m)]).astype(dtype)


def _pair(p, prec, family):
    out = []
    for one_wg in (True, False):
        sat = A.HipSatTanh(prec) if family == "tanh" else A.HipSat(prec)
        sat.SetProblem(A.SatProblem(p.literals, p.lengths))
        rk = A.HipRungeKutta()
        rk.SetEpsilon(1e-6)
        rk.SetOneWorkgroup(one_wg)
        rk.Configure(sat)
        assert rk.UsesOneWorkgroup() == one_wg
        out.append((sat, rk, sat.MakeState(), sat.MakeState()))
    return out


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("case,family", [("k3", "ctds"), ("k4", "ctds"), ("mixed", "ctds"), ("k12", "ctds"),
                                         ("hub", "ctds"), ("k3", "tanh"), ("mixed", "tanh")])
def test_one_workgroup_equals_multi_kernel(case, family, prec):
    say("one workgroup vs multi-kernel: %s %s %s" % (case, family, prec))
    p = _problem(case)
    dt = np.float64 if prec == "f64" else np.float32
    pub = _start(p, 5, dt)
    (sa, ra, oa, na), (sb, rb, ob, nb) = _pair(p, prec, family)
    oa.CopyFrom(pub)
    ob.CopyFrom(pub)
    for count in (1, 7, 50, 200):
        assert ra.Step(oa, na, count) == rb.Step(ob, nb, count)
        a, b = ra.stats(), rb.stats()
        for f in ("time", "hdid", "hnext", "errmax", "steps", "attempts", "rhscount"):
            assert getattr(a, f) == getattr(b, f), (count, f, getattr(a, f), getattr(b, f))
        assert np.array_equal(oa.CopyTo(), ob.CopyTo()), count
        assert sa.GetClauseViolationCount(oa) == sb.GetClauseViolationCount(ob)


def test_one_workgroup_pauses_and_resumes_bitexact():
    """A Step longer than one launch's attempt budget (8192) resumes where the
    previous launch paused, still bit-identical to the multi-kernel path."""
    p = _problem("k3")
    (sa, ra, oa, na), (sb, rb, ob, nb) = _pair(p, "f32", "ctds")
    pub = _start(p, 6, np.float32)
    oa.CopyFrom(pub)
    ob.CopyFrom(pub)
    say("one workgroup: Step(9000) f32")
    assert ra.Step(oa, na, 9000) and rb.Step(ob, nb, 9000)
    a, b = ra.stats(), rb.stats()
    assert (a.time, a.hdid, a.steps, a.attempts, a.rhscount) == (b.time, b.hdid, b.steps, b.attempts, b.rhscount)
    assert np.array_equal(oa.CopyTo(), ob.CopyTo())


def _solver(p, prec, one_wg, batch, max_steps, max_time, pub, callback=False):
    sat = A.HipSat(prec)
    rk = A.HipRungeKutta()
    rk.SetEpsilon(1e-6)
    rk.SetOneWorkgroup(one_wg)
    s = A.SatSolver(sat, rk)
    s.SetProblem(A.SatProblem(p.literals, p.lengths))
    s.SetMaxTime(max_time)
    s.SetMaxSteps(max_steps)
    s.SetBatchSize(batch)
    s.SetSatState(pub)
    seen = []
    if callback:
        s.SetCallback(lambda x: seen.append(x.GetLastClauseViolationCount()) or True)
    return s, seen


def test_device_solve_matches_reference_fixture(golden_dir):
    say("device SatSolver loop, C1 fixture, f64")
    g = np.load(os.path.join(golden_dir, "c1_traj.npz"), allow_pickle=False)
    d = np.load(os.path.join(golden_dir, "c1_solve.npz"), allow_pickle=False)
    p = O.fastsat(50, 213, 3, 0)
    s, _ = _solver(p, "f64", True, 1, 10000, 1e8, g["pub0"])
    res = s.Solve()
    assert int(res) == int(d["result"])
    assert s.GetStepCount() == int(d["steps"])
    assert s.GetRhsCount() == int(d["rhscount"])
    assert s.GetElapsedTime() == float(d["time"])
    assert s.GetLastClauseViolationCount() == 0
    assert np.array_equal(s.GetSatState().CopyTo(), d["final"])


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("batch,max_steps,max_time", [(1, 10000, 1e8), (7, 10000, 1e8), (5, 33, 1e8),
                                                      (3, 10000, 0.75)])
def test_device_solve_matches_host_loop(prec, batch, max_steps, max_time):
    dt = np.float64 if prec == "f64" else np.float32
    for seed in (0, 1, 2):
        p = O.fastsat(50, 213, 3, seed)
        pub = _start(p, 100 + seed, dt)
        say("device vs host SatSolver loop: %s seed %d batch %d max_steps %d max_time %g"
            % (prec, seed, batch, max_steps, max_time))
        dev, _ = _solver(p, prec, True, batch, max_steps, max_time, pub)
        host, seen = _solver(p, prec, True, batch, max_steps, max_time, pub, callback=True)
        rd, rh = dev.Solve(), host.Solve()
        assert rd == rh
        assert dev.GetStepCount() == host.GetStepCount()
        assert dev.GetRhsCount() == host.GetRhsCount()
        assert dev.GetElapsedTime() == host.GetElapsedTime()
        assert dev.GetLastStepSize() == host.GetLastStepSize()
        assert dev.GetLastClauseViolationCount() == host.GetLastClauseViolationCount()
        assert np.array_equal(dev.GetSatState().CopyTo(), host.GetSatState().CopyTo())
        if max_steps < 10000 and rd == A.SatResult.SAT_MAXITER_REACHED:
            assert dev.GetStepCount() == 35
        if max_time < 1:
            assert rd in (A.SatResult.SAT_MAXTIME_REACHED, A.SatResult.SAT_SOLUTION_FOUND)


def test_device_solve_c1_sample_vs_fp64_oracle():
    """20 C1 trajectories (instance seeds 0..1, start seeds 0..9, the
    benchmark's recipe), SAT checked every step, 1e4-step cap: the FP64
    device loop equals the FP64 oracle's SatSolver loop."""
    with O.pow_mode("cr"):
        for seed in range(2):
            p = O.fastsat(50, 213, 3, seed)
            for ic in range(10):
                s0 = np.random.default_rng(1000 * seed + ic).uniform(-1.0, 1.0, 50)
                pub = np.concatenate([s0, np.ones(213)])
                s, _ = _solver(p, "f64", True, 1, 10000, 1e300, pub)
                res = s.Solve()
                o = O.RK(p, eps=1e-6)
                o.set_state(pub)
                ores, viol = o.solve(max_time=1e300, max_steps=10000, batch=1, max_batches=10001)
                say("C1 seed %d ic %d: %s in %d steps" % (seed, ic, A.SatResult(res).name, s.GetStepCount()))
                assert int(res) == int(ores)
                assert s.GetStepCount() == o.steps and s.GetRhsCount() == o.rhscount
                assert s.GetElapsedTime() == o.time
                assert s.GetLastClauseViolationCount() == int(viol[-1])
                assert np.array_equal(s.GetSatState().CopyTo(), o.get_state())
