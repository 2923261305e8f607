"""Command line: ``python -m analogsat_amd run -problem X.cnf [options]``.

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
(FP32 state with FP64 control) and -seed (initial condition).  Options of the
reference this path does not cover (-nogpu, -minisat, -ramsey,
bench/speedtest/make) are rejected.
"""
import os
import sys

import numpy as np


DEFAULTS = {"eps": 1e-6, "tmax": 1e8, "bias": 0.0, "batch": 50, "stepmax": 50000000, "resultfolder": ".",
            "problem": None, "trajectory": False, "float": False, "tanh": False, "seed": None}
RANGES = {"eps": (1e-8, 1e-1), "tmax": (0, 1e80), "bias": (0.0, 1e6), "batch": (1, 2**31 - 1),
          "stepmax": (0, 2**31 - 1), "seed": (0, 2**63)}
UNSUPPORTED = ("-nogpu", "-minisat", "-ramsey", "-ramseycircular", "-config", "-usegpu", "-k", "-alpha",
               "-samplestart", "-sampleend", "-nstart", "-nend", "-nstep", "-problemfolder", "-version", "-timeout",
               "-rerun")


def parse(argv):
    if not argv or argv[0] != "run":
        raise SystemExit("usage: python -m analogsat_amd run -problem FILE.cnf [-eps E] [-batch B] [-stepmax S] "
                         "[-tmax T] [-resultfolder DIR] [-bias B] [-tanh] [-trajectory] [-float] [-seed S]")
    conf = dict(DEFAULTS)
    i = 1
    while i < len(argv):
        a = argv[i]
        if a in UNSUPPORTED:
            raise SystemExit("%s: not supported by the MI355X integrator path" % a)
        key = a.lstrip("-")
        if key in ("trajectory", "float", "tanh"):
            conf[key] = True
            i += 1
            continue
        if key not in conf or i + 1 >= len(argv):
            raise SystemExit("unknown or incomplete option: %s" % a)
        val = argv[i + 1]
        if key in ("problem", "resultfolder"):
            conf[key] = val
        else:
            num = int(val) if key in ("batch", "stepmax", "seed") else float(val)
            lo, hi = RANGES[key]
            if not lo <= num <= hi:
                raise SystemExit("%s out of range [%g, %g]" % (a, lo, hi))
            conf[key] = num
        i += 2
    if not conf["problem"]:
        raise SystemExit("-problem is required")
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


def main(argv=None):
    conf = parse(sys.argv[1:] if argv is None else argv)
    run(conf)


if __name__ == "__main__":
    main()
