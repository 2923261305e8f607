"""Benchmark: accepted RK45 trajectory-steps/s of the CTDS integrator on MI355X.

Workload (default): one uniform random 3-SAT instance, N = 1,000,000 variables,
M = 4,200,000 clauses (alpha = 4.2), FastSatMaker stream with seed 0; s(0)
uniform in [-1, 1) (numpy seed 0), a(0) = 1; FP32 state with FP64 error norm,
step control and time; eps = 1e-6; no SAT exit.  A "step" is one accepted
Cash-Karp step of the whole trajectory (all N + M state variables).

Timed region: exactly K accepted steps (one Step(K) call), inputs resident in
HBM, the device synchronised on both sides.  value = K / wall seconds.

Roofline: before the timed region each kernel kind of one attempt is timed
on its own with HIP events on the integrator's stream (asat_rk_profile,
back-to-back launches of one kind), which names the dominant kernel.  Inside
the timed region the integrator records a HIP event pair around that kernel
in every attempt (asat_rk_set_kernel_timer).  `achieved` = that kernel's
algorithmic bytes per launch / its in-sequence average duration, with the
algorithmic bytes of the RK45 attempt B_alg = 96 L + 136 D (L = K M literal
slots, D = N + M, FP32 state, int32 indices) split over the kernels:
  clause pass J : 8 L + 4 M (J + 1)    (literal read 4 + contribution write 4
                                        per slot; x and k_1..k_{J-1} read,
                                        k_J written per clause)
  var pass J    : 8 L + 4 N (J + 1)    (contribution read 4 + CSR read 4)
The six stages of both passes add up to 96 L + 108 D; the remaining 28 D of
B_alg is the final combine (x, k1, k3..k6 read, x_new written), which the
J = 6 passes fuse into stage 6.  The attempt-level fraction
B_alg x attempts/s / 8 TB/s and the floor B_floor = 24 L + 136 D are
reported beside it, and the bytes the implementation moves by design
(bytes_impl: sorted positions, ranks, region offsets, slot spins) next to
them.  `traffic` is the HBM bytes per launch of the dominant kernel read
from two rocprofv3 counter passes (FETCH_SIZE, WRITE_SIZE) over a short run
of the same instance, each scaled by what the counter reads for a plain copy
of known size in the same run; null where rocprofv3 is unavailable.

CPU baseline: the reference's own CPU solver (oracle/_ref, CpuSat<double> +
CpuRungeKutta<double>, single thread) on the same instance and start for a
bounded sample of accepted steps, when that build is present; otherwise the
oracle's C restatement.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def alg_kernel_bytes(n, m, nslots, esize=4):
    """Algorithmic bytes per launch of each kernel of one attempt, the RK45
    attempt's B_alg split by pass (int32 indices, state of esize bytes; for
    FP32 B_alg = 96 L + 136 D): clause pass J (4 + esize) L + esize M (J + 1),
    var pass J (esize + 4) L + esize N (J + 1).  The final combine's 7 vectors
    are fused into the J = 6 passes and not attributed."""
    L = nslots
    out = {}
    for j in range(1, 7):
        out["clause%d" % j] = (4 + esize) * L + esize * m * (j + 1)
        out["var%d" % j] = (esize + 4) * L + esize * n * (j + 1)
    out["control"] = 0
    return out


def alg_attempt_bytes(n, m, nslots, esize=4):
    """(B_alg, B_floor) of one RK45 attempt; FP32: 96 L + 136 D and 24 L + 136 D."""
    d = n + m
    return 12 * (4 + esize) * nslots + 34 * esize * d, 24 * nslots + 34 * esize * d


def kernel_regex(kind):
    """Pattern of the rocprofv3 kernel names that implement a kernel kind of
    asat_rk_profile's order (clause1..6, var1..6, control)."""
    import re
    if kind.startswith("clause"):
        j = int(kind[6:])
        return re.compile(r"k_clause_tile<\s*\w+\s*,\s*\d+\s*,\s*%d\s*,|k_clause_gen<\s*\w+\s*,\s*%d\s*>" % (j, j))
    if kind.startswith("var"):
        j = int(kind[3:])
        return re.compile(r"k_var(_region|_wide)?<\s*\w+\s*,\s*%d\s*>" % j)
    return re.compile(r"k_control\b")


def traffic_probe(args):
    """Child of measure_traffic, run under rocprofv3 --pmc: a few attempts of
    the benchmark instance, then one plain copy of 256 MiB (counter scale)."""
    import analogsat_amd as A
    n, k = args.n, args.k
    m = int(round(args.alpha * n))
    prob = A.fastsat(n, m, k, 0)
    pub0 = np.concatenate([np.random.default_rng(0).uniform(-1.0, 1.0, n), np.ones(m)])
    sat = A.HipSat(args.precision)
    sat.SetProblem(prob)
    rk = A.HipRungeKutta()
    rk.SetEpsilon(args.eps)
    rk.Configure(sat)
    old, new = sat.MakeState(), sat.MakeState()
    old.CopyFrom(pub0.astype(sat.dtype))
    rk.Step(old, new, 3)
    A.copy_bench(256 << 20, reps=1)
    print("probe done", flush=True)


def measure_traffic(args, kind, log):
    """HBM bytes per launch of kernel kind `kind`: two rocprofv3 --pmc passes
    (FETCH_SIZE, then WRITE_SIZE) over traffic_probe, each counter scaled by
    its reading for the 256 MiB copy (k_copy16 reads and writes 256 MiB).
    Returns (read_bytes, write_bytes, detail) or None with the reason."""
    import csv
    import glob
    import shutil
    import signal
    import subprocess
    import tempfile
    prof = shutil.which("rocprofv3") or ("/opt/rocm/bin/rocprofv3" if os.path.exists("/opt/rocm/bin/rocprofv3")
                                         else None)
    if prof is None:
        return None, "rocprofv3 not found"
    pat = kernel_regex(kind)
    copy_bytes = 256 << 20
    res = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix="asat_pmc_", dir="/tmp")
        cmd = [prof, "--pmc", counter, "--output-format", "csv", "-d", d, "-o", "t", "--",
               sys.executable, os.path.join(ROOT, "bench.py"), "--traffic-probe", "--n", str(args.n),
               "--alpha", str(args.alpha), "--k", str(args.k), "--eps", str(args.eps),
               "--precision", args.precision]
        env = dict(os.environ, TMPDIR="/tmp")
        log("traffic: rocprofv3 --pmc %s over a 3-step run" % counter)
        p = subprocess.Popen(cmd, cwd="/tmp", env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                             start_new_session=True)
        try:
            out = p.communicate(timeout=150)[0].decode(errors="replace")
        except subprocess.TimeoutExpired:
            os.killpg(p.pid, signal.SIGKILL)
            p.communicate()
            shutil.rmtree(d, ignore_errors=True)
            return None, "rocprofv3 --pmc %s timed out" % counter
        files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
        if not files:
            shutil.rmtree(d, ignore_errors=True)
            return None, "rocprofv3 --pmc %s wrote no counter file (rc %d): %s" % (counter, p.returncode,
                                                                                  out[-300:])
        vals, copy = [], []
        for fn in files:
            with open(fn) as f:
                for row in csv.DictReader(f):
                    if row.get("Counter_Name") != counter:
                        continue
                    name = row.get("Kernel_Name", "")
                    v = float(row["Counter_Value"])
                    if "k_copy16" in name:
                        copy.append(v)
                    elif pat.search(name):
                        vals.append(v)
        shutil.rmtree(d, ignore_errors=True)
        if not vals or not copy:
            return None, "counter file lacks the %s kernel or the copy kernel" % kind
        res[counter] = (float(np.mean(vals)), float(np.mean(copy)), len(vals))
    f_fetch = copy_bytes / res["FETCH_SIZE"][1]
    f_write = copy_bytes / res["WRITE_SIZE"][1]
    rd = res["FETCH_SIZE"][0] * f_fetch
    wr = res["WRITE_SIZE"][0] * f_write
    detail = {"read": int(rd), "write": int(wr), "launches": res["FETCH_SIZE"][2],
              "fetch_KiB": round(res["FETCH_SIZE"][0], 1), "write_KiB": round(res["WRITE_SIZE"][0], 1),
              "copy_fetch_KiB": round(res["FETCH_SIZE"][1], 1), "copy_write_KiB": round(res["WRITE_SIZE"][1], 1),
              "bytes_per_fetch_unit": round(f_fetch, 2), "bytes_per_write_unit": round(f_write, 2)}
    return (rd, wr, detail), None


def kernel_bytes(n, m, k, nslots, esize=4, layout=None):
    """Bytes each plain-solver kernel moves by design (every array the kernel
    touches, once; index arrays int32/uint16, state arrays of esize bytes),
    reported as bytes_impl beside the algorithmic figure.

    clause pass J (tiled, K = 3, 4, slot spins, direct primary slots): per
    slot a rank with the literal's sign (2) read; per slot other than its
    clause's primary one a sorted position (4) and the slot's stage spin
    (esize) read through it and one contribution (esize) stored; with run
    tables (layout["run_tables"]) the sorted positions are replaced by a
    word per variable group and tile (groups + 1 rounded up to 16) and the
    region starts (8 per group); with primary words (layout["primary_words"])
    the primary slots have no rank and a clause's primary variable takes 2
    bytes, plus 4 per tile; per clause
    its primary variable (4) read and its primary term (esize) stored, plus
    the N-element stage spin vector the primary spins come from (esize per
    variable); per clause A_J aux values (x_a and k_1..k_{J-1} read, k_J or
    x_new written).
    var pass J (LDS regions): per slot a contribution (esize) and a 16-bit
    offset (2) read, per non-primary slot the next stage's slot spin (esize)
    written; per variable a row pointer (4) and B_J state values.
    """
    L = nslots
    A = (2, 3, 4, 5, 6, 7)
    B = (3, 4, 5, 6, 7, 6)
    idx = 4 * (L - m)
    if layout and layout.get("run_tables"):
        g = layout["groups"]
        idx = 4 * layout["tiles"] * (-(-(g + 1) // 16) * 16) + 8 * (g + 1)
    rank, prim = 2 * L, 4 * m
    if layout and layout.get("primary_words"):
        rank, prim = 2 * (L - m), 2 * m + 4 * layout["tiles"]
    out = {}
    for j in range(6):
        out["clause%d" % (j + 1)] = (rank + idx + 2 * esize * (L - m) + prim + esize * m + esize * n
                                     + esize * m * A[j])
        out["var%d" % (j + 1)] = (2 + esize) * L + esize * (L - m) + 4 * (n + 1) + esize * n * B[j]
    out["control"] = 8 * (-(-m // 1024) + -(-n // 1024))
    return out


def adjoint_bytes(L, D):
    """Algorithmic HBM bytes of the C5 adjoint (FP32 forward, FP64 adjoint),
    L literal slots and D = N + M elements per trajectory.

    forward, per trajectory-attempt: SURVEY's 96 L + 136 D (one FP32 attempt).
    backward, per trajectory-step: the step's FP32 recompute (96 L + 136 D),
    then for each of the 6 stages the FP64 vector-Jacobian product of the
    RHS at the stage input (per slot: literal 4, FP64 term written 8 and read
    8, transposed-list entry 4; per element: stage input 4, adjoint input 8,
    product written 8), and the adjoint recursion over the stages (per
    element 20 FP64 vector reads and writes): 240 L + 416 D.
    """
    return 96 * L + 136 * D, 96 * L + 136 * D + 6 * (24 * L + 20 * D) + 160 * D


def attempt_bytes(n, m, k, nslots, esize=4, layout=None):
    """Bytes one attempt moves by design (sum over the 13 kernels)."""
    return sum(kernel_bytes(n, m, k, nslots, esize, layout).values())


def host_cores():
    """(cores used for the all-core CPU leg, nproc): the CPUs this process
    may run on, at most 16 (one GPU's share of the host on the MI355X box)."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return min(16, avail), os.cpu_count() or 1


def cpu_baseline(lits, lens, pub0, eps, steps):
    import oracle as O
    t_total = 0.0
    if O.ref_available():
        R = O.ref()
        h = R.ref_rk_create(lits, lens, len(lens), eps)
        R.ref_rk_set_state(h, pub0)
        t0 = time.perf_counter()
        R.ref_rk_step(h, steps)
        t_total = time.perf_counter() - t0
        R.ref_rk_destroy(h)
        kind = "reference"
    else:
        p = O.Problem(lits, lens)
        rk = O.RK(p, eps=eps)
        rk.set_state(pub0)
        t0 = time.perf_counter()
        rk.step(steps)
        t_total = time.perf_counter() - t0
        kind = "port"
    return {"value": steps / t_total, "unit": "trajectory-steps/s", "cores": 1, "kind": kind,
            "sample": "%d accepted steps of the same instance and start, FP64 CpuSat/CpuRungeKutta, "
                      "single thread, %.1f s" % (steps, t_total)}


def make_c3(count, n=2000, alpha=4.25):
    """C3: count independent 3-SAT instances, instance seed i, s(0) seed i, a(0) = 1."""
    import analogsat_amd as A
    m = int(alpha * n)
    out = []
    for i in range(count):
        p = A.fastsat(n, m, 3, i)
        pub = np.concatenate([np.random.default_rng(i).uniform(-1.0, 1.0, n), np.ones(m)]).astype(np.float32)
        out.append((p, pub))
    return out


def make_c4(count, seed=0):
    """C4: ragged batch, N ~ U{50..5000}, k ~ U{3,4}, alpha ~ U[0.9,1.3] x threshold_k, M = round(alpha N)."""
    import analogsat_amd as A
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(rng.integers(50, 5001))
        k = int(rng.integers(3, 5))
        alpha = rng.uniform(0.9, 1.3) * (4.26 if k == 3 else 9.93)
        m = int(round(alpha * n))
        p = A.fastsat(n, m, k, 100000 + i)
        pub = np.concatenate([rng.uniform(-1.0, 1.0, n), np.ones(m)]).astype(np.float32)
        out.append((p, pub))
    return out


def run_batch(insts, max_steps, eps, chunk_steps=64, graph_chunks=4, log=None, slice_chunks=0):
    """Batched solver over insts; returns (batch, wall_s, accepted steps, attempts).
    slice_chunks > 0 runs the batch in slices of that many chunks and logs
    progress after each (long budgets); wall_s sums the slices."""
    import analogsat_amd as A
    b = A.HipBatch()
    for p, pub in insts:
        b.add(p, pub)
    b.set_params(eps=eps, max_steps=max_steps, chunk_steps=chunk_steps, check_every=1)
    b.set_tiling(block=0, lds_mode=-1, nbuckets=4)
    t0 = time.perf_counter()
    b.prepare(autotune=True)
    if log:
        log("batch prepared+tuned in %.2f s: %s" % (time.perf_counter() - t0,
            [(x["trajectories"], x["max_slots"], x["block"], x["lds"]) for x in b.buckets()]))
    if slice_chunks > 0:
        wall = 0.0
        while True:
            t0 = time.perf_counter()
            b.run(max_chunks=slice_chunks, graph_chunks=graph_chunks)
            wall += time.perf_counter() - t0
            res = b.results()
            running = sum(r["status"] == 0 for r in res)
            if log:
                log("batch: %.1f s, %d accepted steps, %d trajectories running"
                    % (wall, sum(r["steps"] for r in res), running))
            if running == 0:
                break
    else:
        t0 = time.perf_counter()
        b.run(graph_chunks=graph_chunks)
        wall = time.perf_counter() - t0
    res = b.results()
    steps = sum(r["steps"] for r in res)
    attempts = sum(r["attempts"] for r in res)
    return b, wall, steps, attempts, res


def run_plain_per_instance(insts, max_steps, eps, batch=1):
    """The plain solver launched once per instance (SatSolver loop, SAT check every `batch` steps)."""
    import analogsat_amd as A
    total_steps, total_wall, tts = 0, 0.0, []
    for p, pub in insts:
        sat = A.HipSat("f32")
        rk = A.HipRungeKutta()
        rk.SetEpsilon(eps)
        solver = A.SatSolver(sat, rk)
        solver.SetProblem(p)
        solver.SetMaxTime(1e300)
        solver.SetMaxSteps(max_steps)
        solver.SetBatchSize(batch)
        solver.SetSatState(pub)
        t0 = time.perf_counter()
        res = solver.Solve()
        w = time.perf_counter() - t0
        total_wall += w
        total_steps += solver.GetStepCount()
        if res == A.SatResult.SAT_SOLUTION_FOUND:
            tts.append(w)
    return total_steps / total_wall, tts


def cpu_cores_rate(insts, eps, steps_each, threads):
    """The reference's CPU solver (oracle/_ref, FP64, single-threaded by
    design) on `threads` instances at once, one host thread each, for
    `steps_each` accepted steps: trajectory-steps/s over all cores."""
    import concurrent.futures
    import oracle as O
    if not O.ref_available():
        return None
    R = O.ref()

    def one(item):
        p, x = item
        lits = np.ascontiguousarray(p.literals, np.int32)
        lens = np.ascontiguousarray(p.lengths, np.int32)
        h = R.ref_rk_create(lits, lens, len(lens), eps)
        R.ref_rk_set_state(h, np.ascontiguousarray(x, np.float64))
        R.ref_rk_step(h, steps_each)
        n = R.ref_rk_steps(h)
        R.ref_rk_destroy(h)
        return n

    t0 = time.perf_counter()
    with concurrent.futures.ThreadPoolExecutor(threads) as ex:
        total = sum(ex.map(one, insts[:threads]))
    wall = time.perf_counter() - t0
    return {"value": round(total / wall, 1), "unit": "trajectory-steps/s", "cores": threads, "kind": "reference",
            "sample": "%d instances, %d accepted steps each, one thread per instance (FP64)" % (threads, steps_each)}


def batched_section(workload, count, max_steps, eps, plain_sample, log):
    insts = make_c3(count) if workload == "c3" else make_c4(count)
    b, wall, steps, attempts, res = run_batch(insts, max_steps, eps, log=log)
    solved = [r for r in res if r["status"] == 1]
    wall_tts = sorted(r["solved_wall_ms"] for r in solved)
    p50 = wall_tts[len(wall_tts) // 2] / 1e3 if len(wall_tts) * 2 >= len(res) else None
    plain_rate, plain_tts = run_plain_per_instance(insts[:plain_sample], max_steps, eps)
    out = {
        "workload": workload, "trajectories": len(res), "max_steps": max_steps,
        "value": round(steps / wall, 1), "unit": "trajectory-steps/s",
        "attempts_per_s": round(attempts / wall, 1), "wall_s": round(wall, 3),
        "solved": len(solved), "p50_time_to_solution_s": p50,
        "p50_analog_time_to_solution": (sorted(r["sat_time"] for r in solved)[len(solved) // 2] if solved else None),
        "plain_per_instance": {"value": round(plain_rate, 1), "unit": "trajectory-steps/s", "sample": plain_sample,
                               "p50_time_to_solution_s": (sorted(plain_tts)[len(plain_tts) // 2]
                                                          if plain_tts else None)},
        "batched_over_plain": round((steps / wall) / plain_rate, 1),
        "cpu_reference_all_cores": cpu_cores_rate(insts, eps, 400, host_cores()[0]),
        "buckets": [{k: v for k, v in x.items()} for x in b.buckets()],
    }
    if workload == "c4":
        out["min_violations"] = [r["min_violations"] for r in res[:64]]
    return out


def p50(xs):
    xs = sorted(xs)
    return xs[len(xs) // 2] if xs else None


def p50_all(solved_times, total):
    """Median time to solution over all `total` runs, an unsolved run counting
    as infinite: None unless at least half of the runs solved."""
    xs = sorted(solved_times) + [float("inf")] * (total - len(solved_times))
    v = xs[(total - 1) // 2] if total else float("inf")
    return None if v == float("inf") else v


def workload_c1(args, log):
    """C1 time-to-solution over seeds: N = 50, M = 213, k = 3, eps = 1e-6,
    SAT checked after every accepted step, 1e4-step cap.  Instance seeds
    0..S-1 x initial-condition seeds 0..9.  The batched solver (FP32 state)
    runs all trajectories at once; the plain solver (FP64, the reference's
    <double>) and the reference's CPU solver run a sample one at a time."""
    import analogsat_amd as A
    import oracle as O
    S, n, m = args.c1_seeds, 50, 213
    insts = []
    for seed in range(S):
        p = A.fastsat(n, m, 3, seed)
        for ic in range(10):
            s0 = np.random.default_rng(1000 * seed + ic).uniform(-1.0, 1.0, n)
            insts.append((p, np.concatenate([s0, np.ones(m)])))
    log("C1: %d trajectories (batched, FP32)" % len(insts))
    b, wall, steps, attempts, res = run_batch([(p, x.astype(np.float32)) for p, x in insts], 10000, 1e-6,
                                              chunk_steps=256, graph_chunks=8, log=log)
    solved = [r for r in res if r["status"] == 1]
    out = {"workload": "c1", "trajectories": len(res), "step_cap": 10000,
           "batched": {"dtype": "fp32-state/fp64-control", "wall_s": round(wall, 3),
                       "trajectory_steps_per_s": round(steps / wall, 1), "solved": len(solved),
                       "p50_time_to_solution_s": p50_all([r["solved_wall_ms"] / 1e3 for r in solved], len(res)),
                       "p50_steps_to_solution": p50_all([r["sat_step"] for r in solved], len(res)),
                       "p50_solved_only_s": p50([r["solved_wall_ms"] / 1e3 for r in solved])}}
    # the one-at-a-time legs take every start of instance seeds spread
    # evenly over the S seeds (unsolvable instances included)
    n_inst = max(1, min(S, args.c1_plain_sample // 10))
    pick = [(S * i) // n_inst * 10 + ic for i in range(n_inst) for ic in range(10)]
    sample = [insts[i] for i in pick]
    sres = [res[i] for i in pick]
    ssolved = [r for r in sres if r["status"] == 1]
    out["batched"]["sample"] = {"trajectories": len(pick), "solved": len(ssolved),
                                "p50_time_to_solution_s": p50_all([r["solved_wall_ms"] / 1e3 for r in ssolved],
                                                                  len(pick)),
                                "p50_steps_to_solution": p50_all([r["sat_step"] for r in ssolved], len(pick))}
    log("C1: plain FP64 solver, %d trajectories one at a time" % len(sample))
    tts, tts_steps, dev_ms, steps_p, wall_p = [], [], [], 0, 0.0
    # one solver for the whole sample, as a caller solving instance after
    # instance would keep it: SetProblem per instance, SetSatState per start;
    # the timed region is Solve() (SatSolver::GetElapsedWallTime's span)
    sat = A.HipSat("f64")
    rk = A.HipRungeKutta()
    rk.SetEpsilon(1e-6)
    solver = A.SatSolver(sat, rk)
    solver.SetMaxTime(1e300)
    solver.SetMaxSteps(10000)
    solver.SetBatchSize(1)
    cur = None
    for p, x in sample:
        if p is not cur:
            solver.SetProblem(p)
            cur = p
        solver.SetSatState(x)
        t0 = time.perf_counter()
        r = solver.Solve()
        w = time.perf_counter() - t0
        wall_p += w
        steps_p += solver.GetStepCount()
        if r == A.SatResult.SAT_SOLUTION_FOUND:
            tts.append(w)
            tts_steps.append(solver.GetStepCount())
            dev_ms.append(rk.last_step_ms())
    out["plain"] = {"dtype": "f64", "sample": len(sample), "trajectory_steps_per_s": round(steps_p / wall_p, 1),
                    "solved": len(tts), "p50_time_to_solution_s": p50_all(tts, len(sample)),
                    "p50_steps_to_solution": p50_all(tts_steps, len(sample)),
                    "p50_device_ms_solved": p50(dev_ms), "one_workgroup": rk.UsesOneWorkgroup()}
    if not args.no_cpu:
        kind = "reference" if O.ref_available() else "port"
        log("C1: CPU solver (%s), %d trajectories one at a time, 1 thread" % (kind, len(sample)))
        tts_c, tts_cs, steps_c, wall_c = [], [], 0, 0.0
        for p, x in sample:
            lits = np.ascontiguousarray(p.literals, np.int32)
            lens = np.ascontiguousarray(p.lengths, np.int32)
            x64 = np.ascontiguousarray(x, np.float64)
            t0 = time.perf_counter()
            if kind == "reference":
                R = O.ref()
                viol = np.zeros(10001, np.int32)
                nb, tout = ctypes.c_int(), ctypes.c_double()
                xo = np.zeros(len(x64))
                stats = np.zeros(3, np.int64)
                code = R.ref_solve(lits, lens, len(lens), x64, 1e-6, 1, 10000, 1e300, viol, len(viol),
                                   ctypes.byref(nb), xo, stats, ctypes.byref(tout))
                st = int(stats[0])
            else:
                rk = O.RK(O.Problem(lits, lens), eps=1e-6)
                rk.set_state(x64)
                code = rk.solve(max_time=1e300, max_steps=10000, batch=1, max_batches=10001)[0]
                st = rk.steps
            w = time.perf_counter() - t0
            wall_c += w
            steps_c += st
            if code == 0:
                tts_c.append(w)
                tts_cs.append(st)
        out["cpu"] = {"kind": kind, "cores": 1, "sample": len(sample),
                      "trajectory_steps_per_s": round(steps_c / wall_c, 1), "solved": len(tts_c),
                      "p50_time_to_solution_s": p50_all(tts_c, len(sample)),
                      "p50_steps_to_solution": p50_all(tts_cs, len(sample))}
    return out


def workload_c4(args, log):
    """C4: ragged batch, N ~ U{50..5000}, k ~ U{3,4}; budget of accepted steps
    per instance; exact minimum violation count per instance and the
    autotuned tile per size bucket."""
    insts = make_c4(args.batch_count)
    # budgets beyond a few thousand steps run for minutes: slices of 16 chunks
    # (1024 steps) with a progress line each
    b, wall, steps, attempts, res = run_batch(insts, args.batch_steps, args.eps, log=log,
                                              slice_chunks=16 if args.batch_steps > 2048 else 0)
    slots = sum(int(np.count_nonzero(p.literals)) for p, _ in insts)
    return {"workload": "c4", "trajectories": len(res), "step_budget": args.batch_steps, "slots": slots,
            "value": round(steps / wall, 1), "unit": "trajectory-steps/s", "wall_s": round(wall, 3),
            "attempts_per_s": round(attempts / wall, 1),
            "slot_attempts_per_s": round(sum(int(np.count_nonzero(p.literals)) * r["attempts"]
                                             for (p, _), r in zip(insts, res)) / wall, 1),
            "solved": sum(r["status"] == 1 for r in res),
            "min_violations_mean": round(float(np.mean([r["min_violations"] for r in res])), 2),
            "min_violations_max": int(max(r["min_violations"] for r in res)),
            "min_violations_first64": [r["min_violations"] for r in res[:64]],
            "buckets": b.buckets()}


def workload_c5(args, log):
    """C5: discrete adjoint, N = 1e4, M = 42 500, k = 3, 64 trajectories,
    T = 500 accepted steps, checkpoints every 32; gradient of
    V(T) = sum a K^2 with respect to s(0) and a(0)."""
    import analogsat_amd as A
    n, m, ntraj, T, C = 10000, 42500, 64, 500, 32
    p = A.fastsat(n, m, 3, 5)
    rng = np.random.default_rng(5)
    pubs = np.stack([np.concatenate([rng.uniform(-1, 1, n), np.ones(m)]) for _ in range(ntraj)]).astype(np.float32)
    adj = A.HipAdjoint(p, eps=1e-6, steps=T, checkpoint_every=C)
    log("C5: forward %d trajectories x %d steps" % (ntraj, T))
    t0 = time.perf_counter()
    adj.forward(pubs)
    t1 = time.perf_counter()
    g = adj.backward()
    t2 = time.perf_counter()
    fwd_ms, bwd_ms = adj.timing()
    traj_attempts, lockstep = adj.attempts()
    # central differences of V(T) along 256 random coordinates of trajectory
    # 0, with its step sizes frozen (FP32 replays of the perturbed starts,
    # 64 at a time); the step is the FP32-representable difference
    h0 = adj.steps()[0]
    nv = n + m
    coords = np.random.default_rng(6).choice(nv, 256, replace=False)
    x0 = pubs[0].astype(np.float64)
    starts, steps_fd = [], []
    for i in coords:
        xp, xm = x0.copy(), x0.copy()
        xp[i] += 1e-4
        xm[i] -= 1e-4
        xp32, xm32 = xp.astype(np.float32), xm.astype(np.float32)
        starts += [xp32, xm32]
        steps_fd.append(float(xp32[i]) - float(xm32[i]))
    starts = np.stack(starts)
    Vs = np.concatenate([adj.replay(starts[b:b + 64], np.tile(h0, (len(starts[b:b + 64]), 1)))[0]
                         for b in range(0, len(starts), 64)])
    fd = (Vs[0::2] - Vs[1::2]) / np.array(steps_fd)
    gi = g[0][coords]
    gmax = np.abs(g[0]).max()
    rel = np.abs(gi - fd) / gmax
    log("C5: FD (FP32 replays, h = 1e-4) on 256 coordinates: max |grad - FD| / max|grad| = %.2e, median %.2e"
        % (rel.max(), np.median(rel)))
    # the same coordinates on the FP64 frozen map at h = 1e-6 and 1e-8: in
    # its linear range the two agree, and they are the FP64 reference
    x64 = pubs[0].astype(np.float64)

    def fd_f64(step):
        starts64 = []
        for i in coords:
            xp, xm = x64.copy(), x64.copy()
            xp[i] += step
            xm[i] -= step
            starts64 += [xp, xm]
        starts64 = np.stack(starts64)
        V64 = np.concatenate([adj.replay_f64(starts64[b:b + 64], np.tile(h0, (len(starts64[b:b + 64]), 1)))
                              for b in range(0, len(starts64), 64)])
        return (V64[0::2] - V64[1::2]) / (2 * step)

    fd64, fd64b = fd_f64(1e-6), fd_f64(1e-8)
    rel64 = np.abs(gi - fd64) / gmax
    conv = np.abs(fd64 - fd64b) / gmax
    log("C5: FD (FP64 replays, h = 1e-6): max %.2e, median %.2e; h = 1e-6 vs 1e-8: max %.2e"
        % (rel64.max(), np.median(rel64), conv.max()))
    # HBM roofline of both halves (algorithmic bytes, see adjoint_bytes)
    L, D = 3 * m, n + m
    fb, bb = adjoint_bytes(L, D)
    fwd_bytes, bwd_bytes = fb * traj_attempts, bb * ntraj * T
    roof = {"bound": "hbm", "peak": 8000.0, "unit": "GB/s",
            "forward": {"trajectory_attempts": traj_attempts, "lockstep_attempts": lockstep,
                        "bytes_per_trajectory_attempt": fb, "achieved": round(fwd_bytes / (fwd_ms * 1e-3) / 1e9, 1),
                        "frac": round(fwd_bytes / (fwd_ms * 1e-3) / 8e12, 4)},
            "backward": {"trajectory_steps": ntraj * T, "bytes_per_trajectory_step": bb,
                         "achieved": round(bwd_bytes / (bwd_ms * 1e-3) / 1e9, 1),
                         "frac": round(bwd_bytes / (bwd_ms * 1e-3) / 8e12, 4)}}
    log("C5: forward %.1f GB/s (%.3f of 8 TB/s) over %d trajectory-attempts; backward %.1f GB/s (%.3f)"
        % (roof["forward"]["achieved"], roof["forward"]["frac"], traj_attempts, roof["backward"]["achieved"],
           roof["backward"]["frac"]))
    return {"workload": "c5", "trajectories": ntraj, "steps": T, "checkpoint_every": C,
            "forward_s": round(t1 - t0, 3), "backward_s": round(t2 - t1, 3),
            "forward_device_ms": round(fwd_ms, 2), "backward_device_ms": round(bwd_ms, 2),
            "roofline": roof,
            "backward_trajectory_steps_per_s": round(ntraj * T / (t2 - t1), 1),
            "grad_norm_first": float(np.linalg.norm(g[0])),
            "fd_check": {"coordinates": 256, "trajectory": 0, "frozen_steps": True,
                         "fp32_map_h_1e-4": {"max_abs_err_over_max_grad": float(rel.max()),
                                             "median_abs_err_over_max_grad": float(np.median(rel))},
                         "fp64_map_h_1e-6": {"max_abs_err_over_max_grad": float(rel64.max()),
                                             "median_abs_err_over_max_grad": float(np.median(rel64)),
                                             "max_change_to_h_1e-8_over_max_grad": float(conv.max())}}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=1000000)
    ap.add_argument("--alpha", type=float, default=4.2)
    ap.add_argument("--k", type=int, default=3)
    ap.add_argument("--eps", type=float, default=1e-6)
    ap.add_argument("--precision", default="f32")
    ap.add_argument("--cpu-steps", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-reps", type=int, default=20)
    ap.add_argument("--batch", choices=["none", "c3", "c4"], default="c3",
                    help="also measure a batched workload next to the plain solver")
    ap.add_argument("--batch-count", type=int, default=4096)
    ap.add_argument("--batch-steps", type=int, default=10000)
    ap.add_argument("--plain-sample", type=int, default=4)
    ap.add_argument("--workload", choices=["c2", "c1", "c4", "c5"], default="c2",
                    help="c2: the headline plain-solver line; c1/c4/c5: the other measured workloads")
    ap.add_argument("--c1-seeds", type=int, default=100)
    ap.add_argument("--c1-plain-sample", type=int, default=100,
                    help="C1 trajectories run one at a time on the plain solver and the CPU (whole instances of 10 starts)")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the two rocprofv3 counter passes that measure the dominant kernel's HBM traffic")
    ap.add_argument("--no-c1", action="store_true", help="skip the C1 time-to-solution leg of the default line")
    ap.add_argument("--traffic-probe", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="only run the plain copy kernel on this many MiB (HBM counter calibration) and exit")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (--gpus %d requested); "
              "multi-GPU sharding is not implemented" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import analogsat_amd as A

    if args.traffic_probe:
        traffic_probe(args)
        return

    if args.copy_calibration:
        nbytes = args.copy_calibration << 20
        us = A.copy_bench(nbytes, reps=5)
        print(json.dumps({"copy_bytes": nbytes, "us_per_copy": us, "GBps": 2 * nbytes / (us * 1e-6) / 1e9}))
        return

    log = lambda s: print("[bench] " + s, file=sys.stderr, flush=True)
    if args.workload != "c2":
        fn = {"c1": workload_c1, "c4": workload_c4, "c5": workload_c5}[args.workload]
        print(json.dumps(fn(args, log)), flush=True)
        return

    n, k = args.n, args.k
    m = int(round(args.alpha * n))
    log("instance N=%d M=%d k=%d (FastSatMaker seed 0)" % (n, m, k))
    prob = A.fastsat(n, m, k, 0)
    pub0 = np.concatenate([np.random.default_rng(0).uniform(-1.0, 1.0, n), np.ones(m)])

    sat = A.HipSat(args.precision)
    sat.SetProblem(prob)
    rk = A.HipRungeKutta()
    rk.SetEpsilon(args.eps)
    rk.Configure(sat)
    old, new = sat.MakeState(), sat.MakeState()
    old.CopyFrom(pub0.astype(sat.dtype))

    log("warmup %d steps" % args.warmup)
    rk.Step(old, new, args.warmup)
    log("profiling kernels one kind at a time (%d reps each)" % args.profile_reps)
    us = rk.profile_kernels(old, new, args.profile_reps)
    names = ["clause%d" % j for j in range(1, 7)] + ["var%d" % j for j in range(1, 7)] + ["control"]
    dom = max(names[:-1], key=lambda nm: us[names.index(nm)])
    rk.set_kernel_timer(names.index(dom))
    st0 = rk.stats()
    log("timed %d steps (%s timed in sequence)" % (args.steps, dom))
    t0 = time.perf_counter()
    ok = rk.Step(old, new, args.steps)
    wall = time.perf_counter() - t0
    st1 = rk.stats()
    dev_ms = rk.last_step_ms()
    t_ms, t_n = rk.kernel_timer()
    rk.set_kernel_timer(-1)
    if not ok:
        log("step size underflow during the timed region")
    steps = int(st1.steps - st0.steps)
    attempts = int(st1.attempts - st0.attempts)
    viol = sat.GetClauseViolationCount(old)
    log("done: %d steps, %d attempts, %.3f s wall, %.3f s device, t=%.4g, h=%.3g, violations=%d"
        % (steps, attempts, wall, dev_ms / 1e3, st1.time, st1.hdid, viol))

    nslots = int(np.count_nonzero(prob.literals)) if len(prob.literals) else 0
    esz = 4 if args.precision == "f32" else 8
    lay = sat.Layout()
    kb = kernel_bytes(n, m, k, nslots, esz, lay)
    kernels = {nm: {"us": round(float(us[i]), 2), "GBps": round(float(kb[nm] / (us[i] * 1e-6) / 1e9), 1)}
               for i, nm in enumerate(names)}
    log("kernels (one kind at a time): " + ", ".join("%s %.1fus %.0fGB/s" % (nm, v["us"], v["GBps"])
                                                      for nm, v in kernels.items()))
    dom_us = t_ms * 1e3 / max(t_n, 1)
    log("%s in sequence: %.2f us average over %d launches" % (dom, dom_us, t_n))
    ab = alg_kernel_bytes(n, m, nslots, esz)
    achieved = float(ab[dom] / (dom_us * 1e-6) / 1e9)
    attempt_us = float(us[:13].sum())
    att_bytes = attempt_bytes(n, m, k, nslots, esz, lay)
    b_alg, b_floor = alg_attempt_bytes(n, m, nslots, esz)
    att_rate = attempts / wall
    traffic, traffic_detail = None, None
    if not args.no_traffic:
        got, why = measure_traffic(args, dom, log)
        if got is None:
            traffic_detail = {"unavailable": why}
            log("traffic: " + why)
        else:
            traffic = int(got[0] + got[1])
            traffic_detail = got[2]
            log("traffic: %s reads %.1f MB and writes %.1f MB per launch (algorithmic %.1f MB)"
                % (dom, got[0] / 1e6, got[1] / 1e6, ab[dom] / 1e6))

    hbm = A.stream_bench(1 << 30, 5)
    log("practical HBM rates: copy %.0f GB/s, read %.0f GB/s" % (hbm["copy_GBps"], hbm["read_GBps"]))

    cpu = None
    if not args.no_cpu:
        log("cpu baseline: %d steps" % args.cpu_steps)
        cpu = cpu_baseline(prob.literals, prob.lengths, pub0, args.eps, args.cpu_steps)
        log("cpu baseline %.3f steps/s (%s)" % (cpu["value"], cpu["kind"]))

    batched = None
    if args.batch != "none":
        log("batched %s: %d trajectories, step cap %d" % (args.batch, args.batch_count, args.batch_steps))
        batched = batched_section(args.batch, args.batch_count, args.batch_steps, args.eps, args.plain_sample, log)
        log("batched %.0f traj-steps/s vs plain per instance %.0f (x%.1f)" % (
            batched["value"], batched["plain_per_instance"]["value"], batched["batched_over_plain"]))

    c1 = None
    if not args.no_c1:
        c1 = workload_c1(args, log)
        log("C1 p50 time to solution: batched %s s, plain FP64 %s s, CPU %s s" % (
            c1["batched"]["p50_time_to_solution_s"], c1["plain"]["p50_time_to_solution_s"],
            c1.get("cpu", {}).get("p50_time_to_solution_s")))

    value = steps / wall
    line = {
        "metric": "RK45 accepted trajectory-steps/s + p50 time-to-solution, batched vs plain",
        "value": round(value, 3),
        "unit": "trajectory-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / max(steps, 1), 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32-state/fp64-control" if args.precision == "f32" else "f64",
        "data": "synthetic (FastSatMaker seed 0, s0 numpy seed 0, a0=1)",
        "config": {"workload": "single random 3-SAT instance, plain solver", "n": n, "m": m, "k": k,
                   "alpha": args.alpha, "eps": args.eps, "trajectories": 1},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": dom,
                     "kernel_us": round(dom_us, 2), "kernel_launches": int(t_n),
                     "kernel_us_alone": round(float(us[names.index(dom)]), 2),
                     "alg_bytes_per_launch": int(ab[dom]), "bytes_impl": int(kb[dom]),
                     "impl_GBps": round(float(kb[dom] / (dom_us * 1e-6) / 1e9), 1),
                     "traffic_detail": traffic_detail,
                     "practical_GBps": {"copy": round(hbm["copy_GBps"], 1), "read": round(hbm["read_GBps"], 1)}},
        "attempt": {"attempts_per_s": round(att_rate, 2), "reject_ratio": round(1 - steps / max(attempts, 1), 4),
                    "B_alg": int(b_alg), "B_floor": int(b_floor),
                    "B_alg_GBps": round(b_alg * att_rate / 1e9, 1),
                    "frac_B_alg": round(b_alg * att_rate / 1e9 / HBM_PEAK_GBS, 4),
                    "frac_B_floor": round(b_floor * att_rate / 1e9 / HBM_PEAK_GBS, 4),
                    "kernel_us_sum": round(attempt_us, 1), "bytes_impl": int(att_bytes),
                    "impl_GBps_at_kernel_sum": round(att_bytes / (attempt_us * 1e-6) / 1e9, 1)},
        "kernels_alone": kernels,
        "device_ms": round(dev_ms, 3),
        "cpu_baseline": cpu,
        "batched": batched,
        "time_to_solution": c1,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
