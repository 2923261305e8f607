"""GPU parity at the measured workloads' own sizes, against the CPU oracle.

C2 (the benchmark's single large instance): N = 1e6, M = 4.2e6, k = 3,
FastSatMaker seed 0, s(0) uniform from numpy seed 0, a(0) = 1, FP32 state with
FP64 control.  Checked: the layout the benchmark runs (1024 variable groups of
977, tiled clause pass, LDS var pass, slot spins), the standalone RHS, five
integrator steps through the slot-spin path (state, h, time, attempts and
violation count after every step), all bit-identical to the oracle's mixed
mode (AnalogSAT/cpu_sat/CpuSat.cpp:197-241, CpuRungeKutta.cpp:217-267 in
float state) with the correctly rounded controller pow.

C5 (the adjoint workload): N = 1e4, M = 42 500, 64 trajectories, T = 500
accepted steps, a checkpoint every 32.  Checked: the logged step sizes of six
trajectories equal the oracle's adaptive mixed-mode runs, and the gradients of
two trajectories equal the oracle's FP32-forward / FP64-adjoint within 1e-12
of max |grad|.

C1 over many seeds: 100 instances x 10 initial conditions (N = 50, M = 213),
SAT checked after every accepted step, 1e4-step cap, all 1000 trajectories in
one batched run.  Checked against the oracle's SatSolver loop in mixed mode
(AnalogSAT/solver/SatSolver.h:273-355): the set of solved trajectories, the
accepted step at which each is solved, and each minimum violation count.
"""
import concurrent.futures
import sys

import numpy as np
import pytest

import analogsat_amd as A
import oracle as O

pytestmark = pytest.mark.gpu

WORKERS = 16


def say(msg):
    print("[gpu-full]", msg, flush=True)
    sys.stdout.flush()


@pytest.mark.timeout(600)
def test_c2_layout_rhs_and_steps_bitexact():
    n, m = 1000000, 4200000
    say("C2: N=%d M=%d, FastSatMaker seed 0" % (n, m))
    p = O.fastsat(n, m, 3, 0)
    pub = np.concatenate([np.random.default_rng(0).uniform(-1.0, 1.0, n), np.ones(m)]).astype(np.float32)
    sat = A.HipSat("f32")
    sat.SetProblem(A.SatProblem(p.literals, p.lengths))
    lay = sat.Layout()
    say("C2: layout %s" % lay)
    assert lay["clause_pass"] == 0 and lay["var_pass"] == 0 and lay["slot_spins"] == 1
    assert lay["groups"] == 1024 and lay["group_vars"] == 977
    # tiles sized to fill whole rounds of two workgroups per CU (256 CUs)
    assert lay["direct_primary"] == 1 and lay["tile_clauses"] == 4160
    assert lay["tiles"] == -(-m // 4160) == 1010
    assert lay["run_tables"] == 1 and lay["primary_words"] == 1
    # standalone RHS (spin gathers): bit-identical to the float oracle
    x, d = sat.MakeState(), sat.MakeState()
    x.CopyFrom(pub)
    sat.GetDerivatives(d, x, 0.0)
    xi = np.concatenate([[-1.0], pub]).astype(np.float32)
    want = O.rhs(p, xi, mode="mix")[1:]
    got = d.CopyTo()
    assert np.array_equal(got, want), np.abs(got - want).max()
    say("C2: RHS bit-identical")
    assert sat.GetClauseViolationCount(x) == O.violations(p, pub)
    # the integrator (slot spins carried between the passes)
    rk = A.HipRungeKutta()
    rk.SetEpsilon(1e-6)
    rk.Configure(sat)
    old, new = sat.MakeState(), sat.MakeState()
    old.CopyFrom(pub)
    with O.pow_mode("cr"):
        orc = O.RK(p, eps=1e-6, mode="mix")
        orc.set_state(pub)
        for s in range(1, 6):
            assert rk.Step(old, new, 1)
            orc.step(1)
            st = rk.stats()
            assert (st.hdid, st.time, st.attempts, st.rhscount) == (orc.hdid, orc.time, orc.attempts,
                                                                      orc.rhscount), s
            assert np.array_equal(old.CopyTo(), orc.get_state()), s
            assert sat.GetClauseViolationCount(old) == orc.violations(), s
            say("C2: step %d bit-identical (h %.3e, %d attempts)" % (s, st.hdid, st.attempts))


def _oracle_steps(p, pub, T):
    rk = O.RK(p, eps=1e-6, mode="mix")
    rk.set_state(pub)
    h = np.zeros(T)
    for s in range(T):
        assert rk.step(1)
        h[s] = rk.hdid
    return h


def _oracle_adjoint(p, pub, T):
    return O.RK(p, eps=1e-6, mode="mix").adjoint(pub, T)


@pytest.mark.timeout(900)
def test_c5_adjoint_full_size_vs_oracle():
    n, m, ntraj, T, C = 10000, 42500, 64, 500, 32
    say("C5: N=%d M=%d, %d trajectories, T=%d, checkpoint every %d" % (n, m, ntraj, T, C))
    p = O.fastsat(n, m, 3, 5)
    rng = np.random.default_rng(5)
    pubs = np.stack([np.concatenate([rng.uniform(-1, 1, n), np.ones(m)]) for _ in range(ntraj)]).astype(np.float32)
    adj = A.HipAdjoint(A.SatProblem(p.literals, p.lengths), eps=1e-6, steps=T, checkpoint_every=C)
    V = adj.forward(pubs)
    h = adj.steps()
    g = adj.backward()
    say("C5: GPU forward + backward done")
    assert h.shape == (ntraj, T) and np.all(h > 0)
    grad_traj, step_traj = (0, 63), (13, 29, 42, 50)
    with O.pow_mode("cr"), concurrent.futures.ThreadPoolExecutor(WORKERS) as ex:
        fa = {t: ex.submit(_oracle_adjoint, p, pubs[t], T) for t in grad_traj}
        fs = {t: ex.submit(_oracle_steps, p, pubs[t], T) for t in step_traj}
        for t, f in fs.items():
            assert np.array_equal(h[t], f.result()), t
        for t, f in fa.items():
            Vo, go, ho = f.result()
            assert np.array_equal(h[t], ho), t
            assert V[t] == pytest.approx(Vo, rel=1e-13)
            scale = np.abs(go).max()
            err = np.abs(g[t] - go).max() / scale
            say("C5: trajectory %d gradient max |diff| / max|grad| = %.2e" % (t, err))
            assert err <= 1e-12, (t, err)


def _c1_trajectories():
    out = []
    for seed in range(100):
        p = O.fastsat(50, 213, 3, seed)
        for ic in range(10):
            s0 = np.random.default_rng(1000 * seed + ic).uniform(-1.0, 1.0, 50)
            out.append((p, np.concatenate([s0, np.ones(213)]).astype(np.float32)))
    return out


def _oracle_solve(item):
    p, x = item
    rk = O.RK(p, eps=1e-6, mode="mix")
    rk.set_state(x)
    res, viol = rk.solve(max_time=1e300, max_steps=10000, batch=1, max_batches=10001)
    return res, rk.steps, int(viol.min())


@pytest.mark.timeout(900)
def test_c1_solve_distribution_1000_trajectories_vs_oracle():
    insts = _c1_trajectories()
    say("C1: %d trajectories, batched" % len(insts))
    b = A.HipBatch()
    for p, x in insts:
        b.add(A.SatProblem(p.literals, p.lengths), x)
    b.set_params(eps=1e-6, max_steps=10000, chunk_steps=256, check_every=1)
    b.set_tiling(block=0, lds_mode=-1, nbuckets=4)
    b.prepare(autotune=True)
    b.run(graph_chunks=8)
    res = b.results()
    say("C1: batched run done, %d solved" % sum(r["status"] == 1 for r in res))
    with O.pow_mode("cr"), concurrent.futures.ThreadPoolExecutor(WORKERS) as ex:
        want = list(ex.map(_oracle_solve, insts))
    solved_gpu = {i for i, r in enumerate(res) if r["status"] == 1}
    solved_orc = {i for i, w in enumerate(want) if w[0] == 0}
    say("C1: oracle solved %d" % len(solved_orc))
    assert solved_gpu == solved_orc
    # batch status <- SatResult: solved <- solution found, budget <- maximum
    # steps, underflow <- underflow
    status_of = {0: 1, 4: 2, 1: 3}
    for i, (r, w) in enumerate(zip(res, want)):
        assert r["status"] == status_of[w[0]], i
        assert r["steps"] == w[1], i
        if i in solved_gpu:
            assert r["sat_step"] == w[1], i
        assert r["min_violations"] == w[2], i
    # the distribution the benchmark reports: half of the instances at this
    # ratio are satisfiable, so more than 40% of the trajectories solve
    assert len(solved_gpu) > 400
