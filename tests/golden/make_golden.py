"""Write the golden fixtures under tests/golden/ from the reference's own CPU path.

Runs in the build container only: it needs oracle/_ref/libasat_ref.so, which
oracle/Makefile compiles from the reference's CPU sources in place
(AnalogSAT/cpu_sat/CpuSat.cpp, CpuRungeKutta.cpp, CpuRandom.cpp,
AnalogSAT/problem/SatProblem.cpp, FastSatMaker.cpp, ...).  Every array written
here is an output of those sources; nothing comes from the C restatement.

Fixtures (numpy .npz, no pickles; DIMACS .cnf for instances):
  c1_n50_m213_s0.cnf   the C1 instance: FastSatMaker(seed 0), k=3, N=50, M=213
  c1_traj.npz          FP64 trajectory of CpuRungeKutta<double> on c1 with
                       s(0) = 2u-1 from CpuRandom<double>(seed 1), a(0) = 1,
                       eps = 1e-6: per-step hdid, time, attempts, violations
                       for 1000 steps and the states at steps 1, 10, 100, 1000
  c1_solve.npz         SatSolver<double>::Solve on the same start, batch 1
  rhs_cases.npz        single RHS evaluations (k = 3, 4, mixed lengths with
                       padding, k = 12 generic path, bias b = 2.5, N = 2000)
  gen_cases.npz        FastSatMaker literal streams and CpuRandom uniforms
  tanh_cases.npz       CpuSatTanh<double> and CpuSatTanh<float> RHS evaluations
                       (the cases of rhs_cases.npz, each with its own q)
  c1_traj_ld.npz       CpuSat<long double> + CpuRungeKutta<long double> from the
                       c1_traj start: states after 10, 100 and 1000 steps and the
                       1000 step sizes (x87 extended precision)
  rhs_cases_f32.npz    CpuSat<float> RHS evaluations on the cases of
                       rhs_cases.npz, the state rounded to float
  randomsat_cases.npz  RandomSatMaker literal streams, including instances
                       where most draws repeat an earlier clause
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle as O  # noqa: E402

R = O.ref()
if R is None:
    raise SystemExit("oracle/_ref/libasat_ref.so is not built (make -C oracle ref)")


def ref_fastsat(n, m, k, seed):
    out = np.zeros(m * k, dtype=np.int32)
    R.ref_fastsat(n, m, k, seed, out)
    return out


def ref_uniform(seed, n):
    out = np.zeros(n, dtype=np.float64)
    R.ref_uniform_f64(seed, out, n)
    return out


def write_cnf(path, n, lits, lens):
    with open(path, "w") as f:
        f.write("p cnf %d %d\n" % (n, len(lens)))
        off = 0
        for L in lens:
            f.write(" ".join(str(int(v)) for v in lits[off:off + L]) + " 0\n")
            off += L


def ref_rhs(lits, lens, x, b=0.0):
    out = np.zeros_like(x)
    R.ref_rhs_f64(lits, lens, len(lens), b, x, out)
    return out


def main():
    # ---- C1 instance and trajectory -------------------------------------
    n, m, k = 50, 213, 3
    lits = ref_fastsat(n, m, k, 0)
    lens = np.full(m, k, dtype=np.int32)
    write_cnf(os.path.join(HERE, "c1_n50_m213_s0.cnf"), n, lits, lens)
    u = ref_uniform(1, n + m)
    pub0 = np.concatenate([2.0 * u[:n] - 1.0, np.ones(m)])
    h = R.ref_rk_create(lits, lens, m, 1e-6)
    R.ref_rk_set_state(h, pub0)
    T = 1000
    hdid = np.zeros(T)
    tm = np.zeros(T)
    att = np.zeros(T, dtype=np.int64)
    viol = np.zeros(T, dtype=np.int32)
    snaps = {}
    for s in range(1, T + 1):
        assert R.ref_rk_step(h, 1) == 1
        hdid[s - 1] = R.ref_rk_hdid(h)
        tm[s - 1] = R.ref_rk_time(h)
        att[s - 1] = R.ref_rk_attempts(h)
        viol[s - 1] = R.ref_rk_violations(h)
        if s in (1, 10, 100, 1000):
            st = np.zeros(n + m)
            R.ref_rk_get_state(h, st)
            snaps["state_%d" % s] = st
    rhs_count = R.ref_rk_rhscount(h)
    R.ref_rk_destroy(h)
    np.savez(os.path.join(HERE, "c1_traj.npz"), pub0=pub0, eps=np.float64(1e-6), hdid=hdid, time=tm,
             attempts=att, violations=viol, rhscount=np.int64(rhs_count), **snaps)

    # ---- C1 solve (SatSolver, batch 1, until SAT or 1e4 steps) ------------
    nb = ctypes.c_int()
    vb = np.zeros(20000, dtype=np.int32)
    outp = np.zeros(n + m)
    stats = np.zeros(3, dtype=np.int64)
    tout = ctypes.c_double()
    res = R.ref_solve(lits, lens, m, pub0, 1e-6, 1, 10000, 1e8, vb, len(vb), ctypes.byref(nb), outp, stats,
                      ctypes.byref(tout))
    np.savez(os.path.join(HERE, "c1_solve.npz"), result=np.int32(res), steps=stats[0], rhscount=stats[1],
             time=np.float64(tout.value), batch_violations=vb[: nb.value], final=outp)

    # ---- single RHS evaluations --------------------------------------------
    cases = {}
    rng = np.random.default_rng(12345)

    def add(name, lits, lens, nvar, b=0.0, seed=3):
        mm = len(lens)
        uu = ref_uniform(seed, nvar + mm)
        x = np.concatenate([[-1.0], 2.0 * uu[:nvar] - 1.0, uu[nvar:] + 1.0])
        cases[name + "_lits"] = np.asarray(lits, dtype=np.int32)
        cases[name + "_lens"] = np.asarray(lens, dtype=np.int32)
        cases[name + "_x"] = x
        cases[name + "_b"] = np.float64(b)
        cases[name + "_rhs"] = ref_rhs(np.asarray(lits, dtype=np.int32), np.asarray(lens, dtype=np.int32), x, b)

    add("k3", lits, lens, n)
    l4 = ref_fastsat(100, 990, 4, 5)
    add("k4", l4, np.full(990, 4, dtype=np.int32), 100)
    # mixed lengths 1..5 with zero literals inside clauses (dropped by AddClauses)
    mix_l, mix_n = [], []
    for j in range(300):
        L = int(rng.integers(1, 6))
        vs = rng.choice(np.arange(1, 61), size=L, replace=False)
        sg = np.where(rng.random(L) < 0.5, 1, -1)
        c = list(vs * sg)
        if j % 17 == 0:
            c.insert(int(rng.integers(0, len(c) + 1)), 0)
        mix_l += c
        mix_n.append(len(c))
    add("mixed", mix_l, mix_n, 60)
    l12 = ref_fastsat(40, 30, 12, 9)
    l12b = list(l12) + [5, -7, 0]     # one short clause: padding on the generic path
    add("k12", l12b, [12] * 30 + [3], 40)
    add("bias", lits, lens, n, b=2.5)
    lbig = ref_fastsat(2000, 8500, 3, 11)
    add("n2000", lbig, np.full(8500, 3, dtype=np.int32), 2000)
    np.savez(os.path.join(HERE, "rhs_cases.npz"), **cases)

    # ---- tanh family: same inputs, CpuSatTanh in FP64 and FP32 ------------
    tanh = {}
    for name, q in (("k3", 1.1), ("k4", 2.0), ("mixed", 1.1), ("k12", 0.7), ("bias", 1.5), ("n2000", 1.1)):
        lt, ln = cases[name + "_lits"], cases[name + "_lens"]
        x = cases[name + "_x"]
        b = float(cases[name + "_b"])
        out64 = np.zeros_like(x)
        R.ref_rhs_tanh_f64(lt, ln, len(ln), b, q, x, out64)
        x32 = x.astype(np.float32)
        out32 = np.zeros_like(x32)
        R.ref_rhs_tanh_f32(lt, ln, len(ln), b, q, x32, out32)
        tanh[name + "_q"] = np.float64(q)
        tanh[name + "_rhs64"] = out64
        tanh[name + "_rhs32"] = out32
    np.savez(os.path.join(HERE, "tanh_cases.npz"), **tanh)

    # ---- extended precision trajectory -------------------------------------
    ld = {"pub0": pub0.astype(np.longdouble)}
    for steps in (10, 100, 1000):
        out = np.zeros(len(pub0), np.longdouble)
        hs = np.zeros(steps, np.longdouble)
        done = R.ref_traj_ld(lits, lens, len(lens), 1e-6, ld["pub0"], steps, out, hs)
        assert done == steps
        ld["x%d" % steps] = out
        if steps == 1000:
            ld["hdid"] = hs
    np.savez(os.path.join(HERE, "c1_traj_ld.npz"), **ld)

    # ---- generator / uniform streams ------------------------------------
    gen = {}
    for (nn, mm, kk, sd) in [(50, 213, 3, 0), (50, 213, 3, 99), (2000, 8500, 3, 4), (300, 2979, 4, 2),
                             (10, 40, 5, 1)]:
        gen["fastsat_%d_%d_%d_%d" % (nn, mm, kk, sd)] = ref_fastsat(nn, mm, kk, sd)
    for sd in (0, 1, 2, 4294967295):
        gen["uniform_%d" % sd] = ref_uniform(sd, 2048)
    np.savez(os.path.join(HERE, "gen_cases.npz"), **gen)

    # ---- CpuSat<float> on the RHS cases ------------------------------------
    f32 = {}
    for name in ("k3", "k4", "mixed", "k12", "bias", "n2000"):
        lt, ln = cases[name + "_lits"], cases[name + "_lens"]
        x32 = cases[name + "_x"].astype(np.float32)
        out32 = np.zeros_like(x32)
        R.ref_rhs_f32(lt, ln, len(ln), float(cases[name + "_b"]), x32, out32)
        f32[name + "_rhs32"] = out32
    np.savez(os.path.join(HERE, "rhs_cases_f32.npz"), **f32)

    # ---- RandomSatMaker streams ------------------------------------------
    # (n, m, k, seed): C(6,3) 2^3 = 160 and C(5,4) 2^4 = 80 distinct clauses,
    # so most late draws repeat; k = 1 with m = 2n takes every clause
    rs = {}
    for (nn, mm, kk, sd) in [(50, 213, 3, 0), (6, 150, 3, 1), (8, 400, 3, 2), (5, 70, 4, 3), (4, 8, 1, 0),
                             (300, 2979, 4, 2)]:
        out = np.zeros(mm * kk, dtype=np.int32)
        R.ref_randomsat(nn, mm, kk, sd, out)
        rs["randomsat_%d_%d_%d_%d" % (nn, mm, kk, sd)] = out
    np.savez(os.path.join(HERE, "randomsat_cases.npz"), **rs)
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    main()
