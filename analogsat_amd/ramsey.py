"""Ramsey graph-coloring workloads: CNF encoding, problem files and the solve
series of the reference Frontend's ``make_ramsey`` / ``run_ramsey`` commands.

``RamseyClauseMaker`` follows Frontend/src/ramsey_maker.cpp:25-382:
- edge e of the complete graph on N nodes has ``colorDepth`` bit variables
  e*colorDepth + b + 1 (1 bit for 2 colors, 2 for 3-4, 3 for 5-8); a variable
  that is true (spin > 0) sets its bit of the edge's color;
- for color c with Ramsey digit R[c], every c-colored K_R[c] is excluded: one
  clause per R[c]-clique (cliques in lexicographic node order), whose literals
  run over the color bits b and, inside, the clique's edges in (j1 < j2)
  order, each literal asserting that the edge's bit b differs from bit b of c;
- colors from len(R) up to the next power of two are excluded on every edge;
- with the circulant constraint (``circular``), the adjacency matrix is
  block-circulant: each block's first column holds fresh edge ids, and every
  other entry copies the entry one row up, one column left (cyclic in the
  block), so a coloring is fixed by those first columns.
The file and result names follow Frontend/src/make_ramsey_problems.cpp:27-60
and Frontend/src/ramsey_runner.cpp:28-56, 291-310.
"""
import itertools
import math
import os

import numpy as np

__all__ = ["RamseyClauseMaker", "ramsey_digits", "ramsey_file_name", "parse_ramsey", "make_ramsey_cnf",
           "make_ramsey_problems", "RamseyRunner", "run_ramsey_series"]


def parse_ramsey(text):
    """"3,4,5" -> [3, 4, 5] (the reference splits on commas, atoi per token)."""
    try:
        R = [int(tok) for tok in str(text).split(",") if tok.strip() != ""]
    except ValueError:
        raise ValueError("invalid Ramsey digits %r: expected comma-separated integers" % (text,)) from None
    return R


def ramsey_digits(R):
    """GetRamseyDigits (Frontend/src/make_ramsey_problems.cpp:27-39)."""
    if len(R) > 8:
        raise RuntimeError("invalid ramsey digits")
    return "".join("%d" % r for r in R)


def ramsey_file_name(R, N, circular):
    """GetRamseyFileName (Frontend/src/make_ramsey_problems.cpp:41-60)."""
    return "ramsey_%s_%s_N%d.cnf" % ("circular" if circular else "regular", ramsey_digits(R), N)


class RamseyClauseMaker:
    """Frontend/src/ramsey_maker.h:35-106.  ``circular`` selects
    RAMSEY_CIRCULAR; ``blocksize`` 0 means N."""

    def __init__(self, N, R, circular=False, blocksize=0):
        R = [int(r) for r in R]
        if len(R) < 2 or len(R) > 8:
            raise ValueError("The number of Ramsey digits must be between 2 and 8")
        if any(r < 2 for r in R):
            raise ValueError("Ramsey digits must be at least 2")
        if N < 2:
            raise ValueError("a Ramsey graph needs at least 2 nodes")
        self.N = int(N)
        self.R = R
        self.blocksize = self.N if blocksize == 0 else int(blocksize)
        s = len(R)
        self.colorDepth, smax = (1, 2) if s == 2 else ((2, 4) if s <= 4 else (3, 8))
        self.E_regular, self.edgeCount_regular = self._regular_adjacency()
        self.E_circular, self.edgeCount_circular = self._circular_adjacency(self.blocksize)
        if circular:
            self.E, self.edgeCount = self.E_circular, self.edgeCount_circular
        else:
            self.E, self.edgeCount = self.E_regular, self.edgeCount_regular
        # clause blocks: 2-D int32 arrays, one clause per row, in emission order
        self._blocks = []
        for color, k in enumerate(R):
            self._exclude_monochromatic_clique(k, color)
        for color in range(s, smax):
            self._exclude_color(color)

    # -- adjacency matrices (ramsey_maker.cpp:298-380) ----------------------
    def _regular_adjacency(self):
        N = self.N
        E = np.full((N, N), -1, np.int64)
        iu = np.triu_indices(N, 1)
        ids = np.arange(len(iu[0]), dtype=np.int64)
        E[iu] = ids
        E[(iu[1], iu[0])] = ids
        return E, len(ids)

    def _circular_adjacency(self, bs):
        N = self.N
        if bs <= 0 or N % bs != 0:
            raise ValueError("the circular block size must be a divisor of N")
        E = np.full((N, N), -1, np.int64)
        count = 0
        nb = N // bs
        for a in range(nb):
            for b in range(a, nb):
                for i in range(1 if a == b else 0, bs):
                    row, col = a * bs + i, b * bs
                    E[row, col] = count
                    E[col, row] = count
                    count += 1
                for j in range(1, bs):
                    for i in range(bs):
                        row, col = a * bs + i, b * bs + j
                        if E[row, col] >= 0:
                            continue
                        ee = E[a * bs + (bs + i - 1) % bs, b * bs + j - 1]
                        E[row, col] = ee
                        E[col, row] = ee
        return E, count

    # -- clauses ------------------------------------------------------------
    def _exclude_monochromatic_clique(self, k, color):
        N, cd = self.N, self.colorDepth
        if k > N:
            return
        nodes = np.array(list(itertools.combinations(range(N), k)), dtype=np.int64).reshape(-1, k)
        pairs = [(j1, j2) for j1 in range(k) for j2 in range(j1 + 1, k)]
        edges = np.stack([self.E[nodes[:, j1], nodes[:, j2]] for j1, j2 in pairs], axis=1)
        if (edges == edges[:, :1]).all(axis=1).any():
            print("Warning: Monochromatic clique inavoidable")
        cols = []
        for b in range(cd):
            lit = edges * cd + b + 1
            cols.append(-lit if (color >> b) & 1 else lit)
        self._blocks.append(np.concatenate(cols, axis=1).astype(np.int32))

    def _exclude_color(self, color):
        cd = self.colorDepth
        e = np.arange(self.edgeCount, dtype=np.int64)[:, None]
        bits = np.arange(cd, dtype=np.int64)[None, :]
        lit = e * cd + bits + 1
        sign = np.array([-1 if (color >> b) & 1 else 1 for b in range(cd)], dtype=np.int64)[None, :]
        self._blocks.append((lit * sign).astype(np.int32))

    def SetSpecificColor(self, row, column, color):
        """Unit clauses fixing one edge's color (ramsey_maker.cpp:75-89)."""
        e = int(self.E[row, column])
        cd = self.colorDepth
        unit = [[(e * cd + b + 1) if (color >> b) & 1 else -(e * cd + b + 1)] for b in range(cd)]
        self._blocks.append(np.array(unit, dtype=np.int32).reshape(cd, 1))

    def clause_arrays(self):
        """(literals, lengths) of every clause in emission order."""
        if not self._blocks:
            return np.zeros(0, np.int32), np.zeros(0, np.int32)
        lits = np.concatenate([b.reshape(-1) for b in self._blocks])
        lens = np.concatenate([np.full(b.shape[0], b.shape[1], np.int32) for b in self._blocks])
        return lits, lens

    def GetClauses(self):
        """The clauses as lists of literals (ramsey_maker.cpp:73)."""
        return [list(map(int, row)) for b in self._blocks for row in b]

    def problem(self):
        """The clauses as an analogsat_amd.SatProblem."""
        from . import SatProblem
        lits, lens = self.clause_arrays()
        return SatProblem(lits, lens)

    def SetInitialCondition(self, row, column, color, ic):
        """Write one edge's color bits as +-1 into the state list ic
        (ramsey_maker.cpp:91-100); ic grows as needed."""
        e = int(self.E[row, column])
        for b in range(self.colorDepth):
            var = e * self.colorDepth + b
            if var >= len(ic):
                ic.extend([0.0] * (var + 1 - len(ic)))
            ic[var] = 1.0 if (color >> b) & 1 else -1.0
        return ic

    def color_matrix(self, solution):
        """Colors + 1 of every edge of the state (0 on the diagonal)."""
        N, cd = self.N, self.colorDepth
        sol = np.asarray(solution)
        M = np.zeros((N, N), np.int64)
        for r in range(N):
            for c in range(N):
                if r == c:
                    continue
                e = int(self.E[r, c])
                color = 0
                for b in range(cd):
                    if sol[e * cd + b] > 0:
                        color |= 1 << b
                M[r, c] = color + 1
        return M

    def WriteColorMatrix(self, filename, solution):
        """ramsey_maker.cpp:102-136: one row per node, space separated."""
        M = self.color_matrix(solution)
        with open(filename, "w") as f:
            for row in M:
                f.write(" ".join("%d" % v for v in row) + "\n")
        return M

    def ApplyCircularToRegular(self, instate):
        """State of the circulant encoding -> state of the regular one
        (ramsey_maker.cpp:138-162): edge (i, j), i < j, in regular id order
        takes the bits of its circulant id."""
        cd, N = self.colorDepth, self.N
        src = np.asarray(instate)
        out = np.zeros(self.edgeCount_regular * cd, dtype=src.dtype)
        e = 0
        for i in range(N):
            for j in range(i + 1, N):
                es = int(self.E_circular[i, j])
                out[e * cd:(e + 1) * cd] = src[es * cd:(es + 1) * cd]
                e += 1
        return out

    def WriteCNF(self, filename, verbose=False, out=None):
        """ramsey_maker.cpp:164-180."""
        p = self.problem()
        if verbose:
            w = (out.write if out is not None else lambda s: print(s, end=""))
            w("  %d-SAT\n" % p.Get_K())
            w("  clauses: %d\n" % p.Get_M())
            w("  variables: %d\n" % p.Get_N())
            w("  alpha: %f\n" % (float(np.float32(p.Get_M()) / np.float32(p.Get_N()))))
        p.write_cnf(filename)
        return p


def make_ramsey_cnf(R, N, circular, folder, out=None):
    """MakeRamseyCNF (Frontend/src/make_ramsey_problems.cpp:63-77)."""
    rcm = RamseyClauseMaker(N, R, circular, N)
    fname = ramsey_file_name(R, N, circular)
    (out.write if out is not None else lambda s: print(s, end=""))("Saving %s\n" % fname)
    rcm.WriteCNF(os.path.join(folder, fname), True, out)
    return os.path.join(folder, fname)


def _node_range(conf):
    return range(int(conf["nstart"]), int(conf["nend"]) + 1)


def make_ramsey_problems(conf, out=None):
    """MakeRamseyProblems (Frontend/src/make_ramsey_problems.cpp:88-112)."""
    R = parse_ramsey(conf["ramsey"])
    os.makedirs(conf["problemfolder"], exist_ok=True)
    return [make_ramsey_cnf(R, N, conf["ramseycircular"], conf["problemfolder"], out) for N in _node_range(conf)]


class RamseyRunner:
    """RamseyRunner (Frontend/src/ramsey_runner.h:31-76, ramsey_runner.cpp:
    GPU path) with the MI355X solver: ``Configure(R, N)`` reads the problem
    file made by make_ramsey; ``RunSample()`` solves it from a fresh random
    start, appends the result line to <resultfolder>/perf_<digits>_N<N>.dat
    and keeps the final state in ``GetSolution()``."""

    def __init__(self, conf, out):
        from . import HipRandom
        self.conf = conf
        self.out = out
        os.makedirs(conf["resultfolder"], exist_ok=True)
        os.makedirs(conf["problemfolder"], exist_ok=True)
        self.rng = HipRandom(conf["seed"])
        self.solution = None
        self.last_solve_time = 0.0

    def Configure(self, R, N):
        from . import SatProblem
        conf = self.conf
        self.problem_name = ramsey_file_name(R, N, conf["ramseycircular"])
        infile = os.path.join(conf["problemfolder"], self.problem_name)
        self.outfile = os.path.join(conf["resultfolder"], "perf_%s_N%d.dat" % (ramsey_digits(R), N))
        if not os.path.isfile(infile):
            raise RuntimeError("ramsey input cnf file '%s' does not exist" % infile)
        self.problem = SatProblem.read_cnf(infile)

    def RunSample(self):
        from .__main__ import solve_on_gpu
        p = self.problem
        self.out.write("Running Ramsey CNF with K=%d, N=%d, M=%d\n" % (p.Get_K(), p.Get_N(), p.Get_M()))
        self.out.write("ANALOGSAT GPUv%d:\n" % self.conf["version"])
        self.out.write("Running on GPU\n  %d-SAT\n  clauses: %d\n  variables: %d\n\n"
                       % (p.Get_K(), p.Get_M(), p.Get_N()))
        # RamseyRunner::RunAnalogsatGpu always instantiates the <double>
        # solver (Frontend/src/ramsey_runner.cpp:219-305), so -float is ignored
        conf = dict(self.conf, float=False)
        solved, solver = solve_on_gpu(conf, p, self.rng, self.out, self.problem_name, self.outfile)
        self.solution = solver.GetSatState().CopyTo()
        self.last_solve_time = solver.GetElapsedWallTime()
        return solved

    def GetSolution(self):
        return self.solution

    def GetLastSolveTime(self):
        return self.last_solve_time


def run_ramsey_series(conf, out=None):
    """RunRamseySeries (Frontend/src/ramsey_runner.cpp:28-56): for every N,
    every sample is solved; a solved sample's coloring is written to
    <resultfolder>/coloring_<digits>_N<N>.dat."""
    import sys
    out = out if out is not None else sys.stdout
    R = parse_ramsey(conf["ramsey"])
    rr = RamseyRunner(conf, out)
    solved_any = []
    for N in _node_range(conf):
        rr.Configure(R, N)
        rcm = RamseyClauseMaker(N, R, conf["ramseycircular"], N)
        matrixname = os.path.join(conf["resultfolder"], "coloring_%s_N%d.dat" % (ramsey_digits(R), N))
        for _ in range(conf["samplestart"], conf["sampleend"]):
            if rr.RunSample():
                rcm.WriteColorMatrix(matrixname, rr.GetSolution())
                solved_any.append((N, matrixname))
    return solved_any
