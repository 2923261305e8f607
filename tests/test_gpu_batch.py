"""GPU parity of the batched ragged solver.

Every trajectory of a batch must follow exactly the algorithm of the plain
FP32 integrator: states, h, attempts, SAT step and the minimum violated-clause
count are compared bit for bit / exactly with the oracle's mixed mode
(correctly rounded controller pow), stepping one accepted step at a time and
stopping where the batch stops (SAT or step budget).  Results must not depend
on the tile mapping (workgroup size, contributions in LDS or global scratch,
bucket count).
"""
import sys

import numpy as np
import pytest

import analogsat_amd as A
import oracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print("[gpu-batch]", msg, flush=True)
    sys.stdout.flush()


def make_instances(count, seed, nlo=20, nhi=300, scale=(0.6, 1.2)):
    """Random 3-/4-SAT instances with alpha = threshold_k * U(scale)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(rng.integers(nlo, nhi + 1))
        k = int(rng.choice([3, 4]))
        a = (4.26 if k == 3 else 9.93) * rng.uniform(*scale)
        m = max(1, int(round(a * n)))
        p = O.fastsat(n, m, k, 1000 * seed + i)
        pub = np.concatenate([rng.uniform(-1, 1, n), np.ones(m)]).astype(np.float32)
        out.append((p, pub))
    return out


def oracle_run(p, pub, max_steps, eps=1e-6):
    """Step the mixed-mode oracle like the batch: SAT check after every step."""
    with O.pow_mode("cr"):
        rk = O.RK(p, eps=eps, mode="mix")
        rk.set_state(pub)
        minv = None
        status, sat_step = 2, -1
        for _ in range(max_steps):
            if not rk.step(1):
                status = 3
                break
            v = rk.violations()
            minv = v if minv is None else min(minv, v)
            if v == 0:
                status, sat_step = 1, rk.steps
                break
        return rk, status, sat_step, minv


def check_against_oracle(batch, insts, max_steps):
    for i, (p, pub) in enumerate(insts):
        r = batch.result(i)
        rk, status, sat_step, minv = oracle_run(p, pub, max_steps)
        assert r["status"] == status, (i, r, status)
        assert r["steps"] == rk.steps and r["attempts"] == rk.attempts, (i, r["steps"], rk.steps)
        assert r["rhscount"] == rk.rhscount
        assert r["time"] == rk.time and r["hdid"] == rk.hdid, i
        assert r["sat_step"] == sat_step and r["min_violations"] == minv, (i, r, sat_step, minv)
        assert np.array_equal(batch.state(i), rk.get_state()), i


@pytest.mark.parametrize("block,lds", [(256, -1), (1024, 0), (512, 1)])
def test_batch_matches_oracle_per_trajectory(block, lds):
    say("batch 24 trajectories block=%d lds=%d" % (block, lds))
    insts = make_instances(24, 1)
    b = A.HipBatch()
    for p, pub in insts:
        b.add(A.SatProblem(p.literals, p.lengths), pub)
    b.set_params(eps=1e-6, max_steps=150, chunk_steps=7, check_every=1)
    b.set_tiling(block=block, lds_mode=lds, nbuckets=3)
    b.prepare(autotune=False)
    b.run(graph_chunks=2)
    check_against_oracle(b, insts, 150)


@pytest.mark.parametrize("block,chunk", [(256, 256), (1024, 512), (512, 0)])
def test_clause_chunked_kernel_matches_oracle(block, chunk):
    """lds_mode 3 (no contribution array: per chunk of clauses, r_m in LDS
    and each variable's terms formed from its var-major records), with
    chunks small enough that every instance spans many of them: bit-identical
    to the oracle."""
    say("clause-chunked kernel block=%d chunk=%d" % (block, chunk))
    insts = make_instances(12, 7, nlo=150, nhi=500)
    b = A.HipBatch()
    for p, pub in insts:
        b.add(A.SatProblem(p.literals, p.lengths), pub)
    b.set_params(eps=1e-6, max_steps=60, chunk_steps=11, check_every=1)
    b.set_tiling(block=block, lds_mode=3, nbuckets=2)
    b.set_chunk(chunk)
    b.prepare(autotune=False)
    bk = b.buckets()
    say("buckets: %s" % [(x["trajectories"], x["max_m"], x["block"], x["lds"]) for x in bk])
    assert all(x["lds"] == 3 for x in bk)
    b.run(graph_chunks=2)
    check_against_oracle(b, insts, 60)


def test_large_instances_chunked_and_global_scratch_agree_with_oracle():
    """Instances whose contributions do not fit in LDS (M = 40 000, k = 4):
    the clause-chunked kernel with its default chunk (several per instance)
    and the global-scratch kernel give the oracle's trajectory."""
    say("large instances: clause-chunked vs global scratch")
    insts = []
    for i in range(3):
        n, m = 5000, 40000
        p = O.fastsat(n, m, 4, 8100 + i)
        pub = np.concatenate([np.random.default_rng(40 + i).uniform(-1, 1, n), np.ones(m)]).astype(np.float32)
        insts.append((p, pub))
    for mode in (3, 0):
        b = A.HipBatch()
        for p, pub in insts:
            b.add(A.SatProblem(p.literals, p.lengths), pub)
        b.set_params(eps=1e-6, max_steps=8, chunk_steps=4, check_every=1)
        b.set_tiling(block=1024, lds_mode=mode, nbuckets=1)
        b.prepare(autotune=False)
        assert b.buckets()[0]["lds"] == mode
        b.run(graph_chunks=1)
        check_against_oracle(b, insts, 8)


def test_batch_autotune_and_buckets_do_not_change_results():
    say("batch autotune vs fixed tiling")
    insts = make_instances(40, 2, nlo=30, nhi=600)
    res = []
    for tiling in ((0, -1, 4), (1024, 0, 1)):
        b = A.HipBatch()
        for p, pub in insts:
            b.add(A.SatProblem(p.literals, p.lengths), pub)
        b.set_params(eps=1e-6, max_steps=120, chunk_steps=16)
        b.set_tiling(*tiling)
        b.prepare(autotune=tiling[0] == 0)
        b.run(graph_chunks=3)
        res.append(([b.result(i) for i in range(len(insts))], [b.state(i) for i in range(len(insts))]))
        if tiling[0] == 0:
            bk = b.buckets()
            say("buckets: %s" % bk)
            # workgroup sizes of the contribution kernels (256, 512, 1024)
            # and of the register-resident variants (also 768)
            assert all(x["block"] in (256, 512, 768, 1024) for x in bk)
    (r0, s0), (r1, s1) = res
    for i in range(len(insts)):
        for key in ("status", "steps", "attempts", "time", "min_violations", "sat_step"):
            assert r0[i][key] == r1[i][key], (i, key)
        assert np.array_equal(s0[i], s1[i])


def test_batch_maxsat_budget_and_min_violations():
    say("over-constrained batch (MAXSAT)")
    insts = make_instances(12, 3, nlo=40, nhi=120, scale=(2.0, 2.5))
    b = A.HipBatch()
    for p, pub in insts:
        b.add(A.SatProblem(p.literals, p.lengths), pub)
    b.set_params(eps=1e-6, max_steps=80, chunk_steps=25)
    b.set_tiling(block=256, lds_mode=-1, nbuckets=2)
    b.prepare(autotune=False)
    b.run(graph_chunks=1)
    res = b.results()
    assert sum(r["status"] == 2 and r["steps"] == 80 for r in res) >= 8
    assert all(r["min_violations"] >= 0 for r in res)
    check_against_oracle(b, insts, 80)


def test_batch_reset_reproduces():
    say("batch reset")
    insts = make_instances(8, 4)
    b = A.HipBatch()
    for p, pub in insts:
        b.add(A.SatProblem(p.literals, p.lengths), pub)
    b.set_params(eps=1e-6, max_steps=40, chunk_steps=40)
    b.set_tiling(block=256, lds_mode=-1, nbuckets=1)
    b.prepare(autotune=False)
    b.run()
    first = [b.state(i).copy() for i in range(len(insts))]
    b.reset()
    b.run()
    for i in range(len(insts)):
        assert np.array_equal(first[i], b.state(i))


def test_batch_runs_in_slices_like_one_run():
    """asat_batch_run with a chunk limit stops early and the next call goes
    on from there: one-chunk slices give the results and states of a single
    run (long budgets are run this way to report progress)."""
    say("batch run in slices")
    insts = make_instances(10, 8)
    b = A.HipBatch()
    for p, pub in insts:
        b.add(A.SatProblem(p.literals, p.lengths), pub)
    b.set_params(eps=1e-6, max_steps=70, chunk_steps=8)
    b.set_tiling(block=256, lds_mode=-1, nbuckets=2)
    b.prepare(autotune=False)
    b.run(graph_chunks=2)
    keys = ("status", "steps", "attempts", "rhscount", "time", "hdid", "min_violations", "sat_step")
    whole = [{k: b.result(i)[k] for k in keys} for i in range(len(insts))]
    states = [b.state(i).copy() for i in range(len(insts))]
    assert any(r["steps"] > 8 for r in whole)
    b.reset()
    slices = 0
    while any(r["status"] == 0 for r in b.results()):
        b.run(max_chunks=1, graph_chunks=1)
        slices += 1
        assert slices <= 20
    assert slices > 1
    for i in range(len(insts)):
        assert {k: b.result(i)[k] for k in keys} == whole[i], i
        assert np.array_equal(b.state(i), states[i]), i


def test_register_kernel_matches_oracle():
    """The register-resident kernel (lds_mode 2: state and stage derivatives
    in registers) on instance sizes that select its different variants,
    including the C3 size (N = 2000, M = 8500): bit-identical to the oracle."""
    say("register-resident kernel, mixed sizes")
    insts = make_instances(8, 5, nlo=20, nhi=60, scale=(0.6, 1.0))        # m <= 256: 256 threads x 1 clause
    insts += make_instances(6, 6, nlo=150, nhi=220, scale=(0.6, 1.0))      # m <= 1024 (k = 3) or 4096
    for i in range(2):
        p = O.fastsat(2000, 8500, 3, 7000 + i)
        pub = np.concatenate([np.random.default_rng(i).uniform(-1, 1, 2000), np.ones(8500)]).astype(np.float32)
        insts.append((p, pub))
    b = A.HipBatch()
    for p, pub in insts:
        b.add(A.SatProblem(p.literals, p.lengths), pub)
    b.set_params(eps=1e-6, max_steps=40, chunk_steps=9, check_every=1)
    b.set_tiling(block=0, lds_mode=2, nbuckets=4)
    b.prepare(autotune=False)
    bk = b.buckets()
    say("buckets: %s" % [(x["trajectories"], x["max_n"], x["max_m"], x["block"], x["lds"]) for x in bk])
    assert all(x["lds"] == 2 for x in bk)
    b.run(graph_chunks=2)
    check_against_oracle(b, insts, 40)
