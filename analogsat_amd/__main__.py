"""Command line: ``python -m analogsat_amd run|bench|speedtest [options]``.

Mirrors the reference Frontend's ``run`` command for the GPU solver
(Frontend/src/main.cpp:201-245 defaults, Frontend/src/cnf_runner.cpp:247-307):
read a DIMACS file, solve it with the CTDS integrator, print one progress line
per batch (steps, analog time, last step, violations, minimum violations), the
SatResult message and the reference's "Performance: ... RK steps/sec" line,
and write ``<resultfolder>/<name>.out`` in MiniSat-compatible form
(SAT / INDET and the signed assignment).

Options (same names and ranges as the reference's parser,
Frontend/src/main.cpp:46-199): -problem, -resultfolder, -eps, -tmax, -stepmax,
-batch, -bias, -tanh (the tanh CTDS family, Frontend/src/cnf_runner.cpp:44-54),
-trajectory (write <name>_traj_orig.dat or <name>_traj_tanh.dat), plus -float
(FP32 state with FP64 control) and -seed (initial condition).

``bench`` and ``speedtest`` follow Frontend/src/bench_runner.cpp:37-117,
351-499: problem sizes N = round(10 * 2^p) for p = nstart, nstart + nstep,
.. nend; samples <problemfolder>/<k>a<int(100 alpha)>/N<N>/pNNNN.cnf, made
with FastSatMaker (M = int(alpha N)) when missing and kept once found SAT;
one result line per run appended to
<resultfolder>/perf_K<k>_a<int(100 alpha)>_N<N>.dat (sample, device,
solver, result, wall s, cpu s, RHS count); ``bench`` skips samples already
listed there unless -rerun; ``speedtest`` cycles through the existing
samples with the -timeout wall limit until 5 s of total solve time (at most
200 runs) per N.  -config FILE reads further options from a file, one
"-option value" per line, recursively.

``make_ramsey`` and ``run_ramsey`` follow Frontend/src/make_ramsey_problems.cpp
and Frontend/src/ramsey_runner.cpp (``analogsat_amd.ramsey``): for node counts
N = nstart .. nend, -ramsey 3,4,5 (the digits of R(3,4,5); the reference's
help text calls it -R, its parser -ramsey) and -ramseycircular (block-circulant
adjacency), write <problemfolder>/ramsey_<regular|circular>_<digits>_N<N>.cnf,
or solve those files samplestart .. sampleend - 1 times each, appending result
lines to <resultfolder>/perf_<digits>_N<N>.dat and writing a solved coloring to
<resultfolder>/coloring_<digits>_N<N>.dat.  Options of the reference this
path does not cover (-nogpu, -minisat) are rejected.
"""
import math
import os
import sys

import numpy as np


# Frontend/src/main.cpp:205-245 defaults and :166-192 ranges
DEFAULTS = {"eps": 1e-6, "tmax": 1e8, "bias": 0.0, "batch": 50, "stepmax": 50000000, "resultfolder": ".",
            "problem": None, "trajectory": False, "float": False, "tanh": False, "seed": None,
            "k": 3, "alpha": 4.25, "samplestart": 0, "sampleend": 100, "nstart": 1.0, "nend": 5.0, "nstep": 0.5,
            "problemfolder": ".", "timeout": 3600.0, "rerun": False, "force_save_cnf": False, "version": 1,
            "usegpu": 0, "ramsey": "", "ramseycircular": False}
RANGES = {"eps": (1e-8, 1e-1), "tmax": (0, 1e80), "bias": (0.0, 1e6), "batch": (1, 2**31 - 1),
          "stepmax": (0, 2**31 - 1), "seed": (0, 2**63), "k": (2, 256), "alpha": (0.0, 1e10),
          "samplestart": (0, 9999), "sampleend": (0, 10000), "nstart": (1.0, 100.0), "nend": (1.0, 100.0),
          "nstep": (0.0, 100.0), "timeout": (0, 1e80), "version": (1, 3), "usegpu": (0, 99)}
INTS = ("batch", "stepmax", "seed", "k", "samplestart", "sampleend", "version", "usegpu")
FLAGS = ("trajectory", "float", "tanh", "rerun", "force_save_cnf", "ramseycircular")
STRINGS = ("problem", "resultfolder", "problemfolder", "ramsey")
COMMANDS = ("run", "bench", "speedtest", "make_ramsey", "run_ramsey")
UNSUPPORTED = ("-nogpu", "-minisat")


def parse_options(args, conf, depth=0):
    """Frontend/src/main.cpp:133-194 ParseOptions, -config recursion included."""
    i = 0
    while i < len(args):
        a = args[i]
        if a in UNSUPPORTED:
            raise SystemExit("%s: not supported by the MI355X integrator path" % a)
        key = a.lstrip("-")
        if key == "config":
            if i + 1 >= len(args):
                raise SystemExit("-config needs a file name")
            if depth > 16:
                raise SystemExit("-config nested too deeply")
            sub = []
            with open(args[i + 1]) as f:
                for line in f:
                    line = line.rstrip("\r\n")
                    pos = line.find(" ")
                    sub += [line[:pos], line[pos + 1:]] if pos >= 0 else [line]
            parse_options([x for x in sub if x != ""], conf, depth + 1)
            i += 2
            continue
        if key in FLAGS:
            conf[key] = True
            i += 1
            continue
        if key not in conf or i + 1 >= len(args):
            raise SystemExit("unknown or incomplete option: %s" % a)
        val = args[i + 1]
        if key in STRINGS:
            conf[key] = val
        else:
            num = int(val) if key in INTS else float(val)
            lo, hi = RANGES[key]
            if not lo <= num <= hi:
                raise SystemExit("%s out of range [%g, %g]" % (a, lo, hi))
            conf[key] = num
        i += 2
    return conf


def parse(argv):
    if not argv or argv[0] not in COMMANDS:
        raise SystemExit("usage: python -m analogsat_amd run|bench|speedtest|make_ramsey|run_ramsey "
                         "[-config FILE] [-problem FILE.cnf] [-ramsey 3,4,5] [-ramseycircular] "
                         "[-eps E] [-batch B] [-stepmax S] [-tmax T] [-timeout S] [-resultfolder DIR] "
                         "[-problemfolder DIR] [-k K] [-alpha A] [-nstart P] [-nend P] [-nstep P] "
                         "[-samplestart I] [-sampleend I] [-bias B] [-tanh] [-trajectory] [-float] [-seed S] "
                         "[-rerun] [-force_save_cnf]")
    conf = dict(DEFAULTS)
    conf["command"] = argv[0]
    parse_options(argv[1:], conf)
    if conf["command"] == "run" and not conf["problem"]:
        raise SystemExit("-problem is required")
    if conf["command"] in ("make_ramsey", "run_ramsey"):
        from .ramsey import parse_ramsey
        try:
            R = parse_ramsey(conf["ramsey"])
        except ValueError as e:
            raise SystemExit(str(e))
        if not 2 <= len(R) <= 8 or min(R) < 2:
            raise SystemExit("-ramsey needs 2 to 8 comma-separated digits, each at least 2 (e.g. -ramsey 3,4,5)")
        if int(conf["nstart"]) < 2:
            raise SystemExit("-nstart: a Ramsey graph needs at least 2 nodes")
    if conf["usegpu"] != 0:
        raise SystemExit("-usegpu: this build runs on device 0 only")
    return conf


def run(conf, out=sys.stdout):
    import analogsat_amd as A
    os.makedirs(conf["resultfolder"], exist_ok=True)
    problem = A.SatProblem.read_cnf(conf["problem"])
    prec = "f32" if conf["float"] else "f64"
    sat = A.HipSatTanh(prec) if conf["tanh"] else A.HipSat(prec)
    sat.Set_B(conf["bias"])
    rk = A.HipRungeKutta()
    rk.SetEpsilon(conf["eps"])
    solver = A.SatSolver(sat, rk)
    solver.SetProblem(problem)
    solver.SetMaxTime(conf["tmax"])
    solver.SetMaxSteps(max(1, conf["stepmax"]))
    solver.SetBatchSize(conf["batch"])
    solver.SetRandomInitialState(A.HipRandom(conf["seed"]))
    n, m = sat.Get_N(), sat.Get_M()
    name = os.path.splitext(os.path.basename(conf["problem"]))[0]
    traj = None
    if conf["trajectory"]:
        traj = open(os.path.join(conf["resultfolder"], "%s_traj_%s.dat" % (name, "tanh" if conf["tanh"] else "orig")),
                    "w")

    def write_state(s):
        x = s.GetSatState().CopyTo()
        traj.write("%f" % s.GetElapsedTime() + "".join(" %f" % v for v in x) + "\n")

    if traj:
        write_state(solver)
    min_v = [m]

    def callback(s):
        v = s.GetLastClauseViolationCount()
        min_v[0] = min(min_v[0], v)
        out.write("  %d\t%5.4f\t%e\t%d\t%d\n" % (s.GetStepCount(), s.GetElapsedTime(), s.GetLastStepSize(), v,
                                                min_v[0]))
        if traj:
            write_state(s)
        return True

    solver.SetCallback(callback)
    result = solver.Solve()
    if traj:
        traj.close()
    last = solver.GetSatState().CopyTo()
    out.write("Performance: %f RK steps/sec\n" % (solver.GetRhsCount() / 6.0 / max(solver.GetElapsedWallTime(), 1e-12)))
    out.write(A.GetSatResultMessage(result) + "\n")
    with open(os.path.join(conf["resultfolder"], name + ".out"), "w") as f:
        f.write("SAT\n" if result == A.SatResult.SAT_SOLUTION_FOUND else "INDET\n")
        f.write(" ".join(str(i + 1 if last[i] > 0 else -(i + 1)) for i in range(n)) + " 0\n")
    return result


def n_values(conf):
    """Problem sizes of a series (Frontend/src/main.cpp:402-404, bench_runner.cpp:38-49)."""
    if conf["nstep"] <= 0:
        raise SystemExit("-nstep must be positive for bench and speedtest")
    powers, p = [], conf["nstart"]
    while p < conf["nend"] + 1e-3:
        powers.append(p)
        p += conf["nstep"]
    return [int(math.floor(10.0 * 2.0 ** p + 0.5)) for p in powers]


class BenchRunner:
    """Frontend/src/bench_runner.cpp:351-499 with the MI355X solver.

    ``configure(N, sample)`` names the sample and loads it if its CNF exists;
    ``run_sample()`` makes it first if not (one FastSatMaker stream per
    runner), solves it, appends the result line and saves the CNF when it
    was found SAT (or always with -force_save_cnf).
    """

    def __init__(self, conf, out=sys.stdout):
        import analogsat_amd as A
        self.A = A
        self.conf = conf
        self.out = out
        os.makedirs(conf["resultfolder"], exist_ok=True)
        os.makedirs(conf["problemfolder"], exist_ok=True)
        self.N = None
        self.maker = A.FastSatMaker(conf["seed"])
        self.rng = None
        self.done = []
        self.last_solve_time = 0.0
        self.problem = None

    def configure(self, n, sample):
        conf = self.conf
        same = self.N == n
        self.N = n
        self.M = int(conf["alpha"] * n)
        a = int(conf["alpha"] * 100)
        folder = os.path.join(conf["problemfolder"], "%da%d" % (conf["k"], a), "N%d" % n)
        os.makedirs(folder, exist_ok=True)
        self.infile = os.path.join(folder, "p%04d.cnf" % sample)
        self.problem_name = "%da%d/N%d/p%04d.cnf" % (conf["k"], a, n, sample)
        self.out.write("Problem: %s\n" % self.problem_name)
        self.cnf_exists = os.path.isfile(self.infile)
        self.problem = self.A.SatProblem.read_cnf(self.infile) if self.cnf_exists else None
        self.outfile = os.path.join(conf["resultfolder"], "perf_K%d_a%d_N%d.dat" % (conf["k"], a, n))
        if same:
            return
        self.done = []
        if os.path.isfile(self.outfile):
            with open(self.outfile) as f:
                for line in f:
                    if "\t" in line:
                        self.done.append(line.split("\t", 1)[0])

    def is_done(self):
        return self.problem_name in self.done

    def run_sample(self):
        conf = self.conf
        if not self.cnf_exists:
            self.maker.Set_K(conf["k"])
            self.maker.Set_N(self.N)
            self.maker.Set_M(self.M)
            self.out.write("Making K=%d, N=%d, M=%d\n" % (conf["k"], self.N, self.M))
            self.problem = self.maker.MakeSatProblem()
        self.out.write("ANALOGSAT GPUv%d:\n" % conf["version"])
        solved = self._run_gpu()
        if (solved and not self.cnf_exists) or conf["force_save_cnf"]:
            self.out.write("saving cnf\n")
            self.cnf_exists = True
            self.problem.write_cnf(self.infile)
        if solved:
            self.done.append(self.problem_name)
        return solved

    def _run_gpu(self):
        if self.rng is None:
            self.rng = self.A.HipRandom(self.conf["seed"])
        solved, solver = solve_on_gpu(self.conf, self.problem, self.rng, self.out, self.problem_name, self.outfile)
        self.last_solve_time = solver.GetElapsedWallTime()
        return solved


def solve_on_gpu(conf, problem, rng, out, problem_name, outfile):
    """One solve from a fresh random start, as the reference's runners do it
    (Frontend/src/bench_runner.cpp:425-499, Frontend/src/ramsey_runner.cpp:
    183-257): progress every 100 steps, the -timeout wall limit, then one
    result line appended to outfile.  Returns (solved, solver)."""
    import analogsat_amd as A
    prec = "f32" if conf["float"] else "f64"
    sat = A.HipSatTanh(prec) if conf["tanh"] else A.HipSat(prec)
    rk = A.HipRungeKutta()
    rk.SetEpsilon(conf["eps"])
    solver = A.SatSolver(sat, rk)
    sat.Set_B(conf["bias"])
    solver.SetProblem(problem)
    solver.SetMaxTime(conf["tmax"])
    solver.SetMaxSteps(max(1, conf["stepmax"]))
    solver.SetBatchSize(conf["batch"])
    solver.SetRandomInitialState(rng)
    min_v = [problem.Get_M()]

    def callback(s):
        v = s.GetLastClauseViolationCount()
        min_v[0] = min(min_v[0], v)
        if s.GetStepCount() % 100 == 0:
            out.write("  %d\t%5.4f\t%e\t%d\t%d\n" % (s.GetStepCount(), s.GetElapsedTime(), s.GetLastStepSize(), v,
                                                    min_v[0]))
        if s.GetElapsedWallTime() > conf["timeout"]:
            out.write("  walltime limit reached\n")
            return False
        return True

    solver.SetCallback(callback)
    res = solver.Solve()
    msg = A.GetSatResultMessage(res)
    out.write(msg + "\n")
    with open(outfile, "a") as f:
        f.write("%s\t%s\tanalogsat%d%s\t%s\t%f\t%f\t%d\n" % (
            problem_name, A.GetGpuName(), conf["version"], "tanh" if conf["tanh"] else "orig", msg,
            solver.GetElapsedWallTime(), solver.GetElapsedCpuTime(), solver.GetRhsCount()))
    return res == A.SatResult.SAT_SOLUTION_FOUND, solver


def run_bench_series(conf, out=sys.stdout):
    """RunBenchSeries (Frontend/src/bench_runner.cpp:52-72)."""
    bench = BenchRunner(conf, out)
    for n in n_values(conf):
        for s in range(conf["samplestart"], conf["sampleend"]):
            bench.configure(n, s)
            if conf["rerun"] or not bench.is_done():
                bench.run_sample()


def run_speedtest(conf, walltimeout=5.0, out=sys.stdout):
    """RunSpeedtest (Frontend/src/bench_runner.cpp:79-117): per N, cycle the
    existing samples until their solve walltimes add up to ``walltimeout``
    seconds, at most 200 runs."""
    bench = BenchRunner(conf, out)
    for n in n_values(conf):
        s, total, count = 0, 0.0, 0
        while total < walltimeout and count < 200:
            bench.configure(n, s)
            if bench.cnf_exists:
                bench.run_sample()
                total += bench.last_solve_time
                count += 1
                s += 1
            elif s == 0:
                raise SystemExit("no cnf samples exist with the given configuration")
            else:
                s = 0


def main(argv=None):
    conf = parse(sys.argv[1:] if argv is None else argv)
    if conf["command"] == "run":
        run(conf)
    elif conf["command"] == "bench":
        run_bench_series(conf)
    elif conf["command"] == "make_ramsey":
        from .ramsey import make_ramsey_problems
        make_ramsey_problems(conf)
    elif conf["command"] == "run_ramsey":
        from .ramsey import run_ramsey_series
        run_ramsey_series(conf)
    else:
        run_speedtest(conf)
    print("All done.")


if __name__ == "__main__":
    main()
