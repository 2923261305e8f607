"""Ramsey workloads (Frontend/src/ramsey_maker.cpp, make_ramsey_problems.cpp,
ramsey_runner.cpp): the CNF encoding, the problem files and the run series.

The reference's maker is part of the Frontend, whose umbrella header pulls in
the CUDA headers, so it cannot be compiled here as a checker.  The encoding is
pinned instead by what the mathematics fixes:
- for random colorings, the CNF is satisfied exactly when no color c holds a
  monochromatic K_R[c] and no edge uses a color beyond the digits, checked by
  an independent clique search on the decoded coloring (asymmetric digits and
  three colors, so a flipped literal sign, a wrong bit order or a dropped
  color-exclusion clause fails);
- R(3,3) = 6: exactly the 12 two-colorings of K5 without a monochromatic
  triangle satisfy the N = 5 problem and none the N = 6 one (brute force);
- clause counts and lengths in closed form; the circulant adjacency's
  defining recurrence; a circulant solution maps to a regular one.
"""
import itertools
import math
from math import comb
import os

import numpy as np
import pytest

from analogsat_amd import __main__ as cli
from analogsat_amd.ramsey import RamseyClauseMaker, parse_ramsey, ramsey_file_name


def satisfied(clauses, x):
    """x: array of +-1 per variable (index = variable - 1)."""
    return all(any((lit > 0) == (x[abs(lit) - 1] > 0) for lit in c) for c in clauses)


def valid_coloring(colors, R):
    """No edge colored beyond the digits and no K_R[c] of color c."""
    N = colors.shape[0]
    if (colors[~np.eye(N, dtype=bool)] >= len(R)).any():
        return False
    for c, k in enumerate(R):
        for nodes in itertools.combinations(range(N), k):
            if all(colors[a, b] == c for a, b in itertools.combinations(nodes, 2)):
                return False
    return True


@pytest.mark.parametrize("N,R", [(7, (3, 4)), (6, (3, 3, 3)), (6, (4, 3))])
def test_cnf_means_no_monochromatic_clique(N, R):
    m = RamseyClauseMaker(N, R)
    clauses = m.GetClauses()
    nvar = m.edgeCount * m.colorDepth
    rng = np.random.default_rng(N * 100 + len(R))
    hits = 0
    for t in range(400):
        # bias towards valid colorings: start from a random coloring with
        # colors below len(R), sometimes perturbed to an unused color
        colors = rng.integers(0, len(R), size=(N, N))
        colors = np.triu(colors, 1) + np.triu(colors, 1).T
        if t % 5 == 0 and m.colorDepth > 1:
            colors[0, 1] = colors[1, 0] = (1 << m.colorDepth) - 1
        x = -np.ones(nvar)
        for i in range(N):
            for j in range(i + 1, N):
                e = int(m.E[i, j])
                for b in range(m.colorDepth):
                    if (colors[i, j] >> b) & 1:
                        x[e * m.colorDepth + b] = 1.0
        np.fill_diagonal(colors, -1)
        want = valid_coloring(colors, R)
        assert satisfied(clauses, x) == want, (t, colors)
        hits += want
        # the color matrix writer decodes the same coloring (colors + 1)
        M = m.color_matrix(x)
        assert np.array_equal(M, colors + 1)
    assert hits > 0


def test_r33_brute_force():
    five = RamseyClauseMaker(5, [3, 3]).GetClauses()
    n5 = sum(satisfied(five, [1 if (b >> v) & 1 else -1 for v in range(10)]) for b in range(1 << 10))
    assert n5 == 12
    six = RamseyClauseMaker(6, [3, 3]).GetClauses()
    assert not any(satisfied(six, [1 if (b >> v) & 1 else -1 for v in range(15)]) for b in range(1 << 15))


@pytest.mark.parametrize("N,R", [(8, (3, 3)), (9, (3, 4, 5)), (7, (3, 3, 3, 3)), (6, (3, 3, 3, 3, 3))])
def test_clause_counts_and_lengths(N, R):
    m = RamseyClauseMaker(N, R)
    lits, lens = m.clause_arrays()
    s = len(R)
    cd, smax = (1, 2) if s == 2 else ((2, 4) if s <= 4 else (3, 8))
    assert m.colorDepth == cd
    E = N * (N - 1) // 2
    assert m.edgeCount == E
    want = [math.comb(N, k) for k in R] + [E] * (smax - s)
    got, pos = [], 0
    for k, n in zip(list(R) + [None] * (smax - s), want):
        L = cd * k * (k - 1) // 2 if k is not None else cd
        assert (lens[pos:pos + n] == L).all()
        got.append(n)
        pos += n
    assert pos == len(lens) and got == want
    assert lens.sum() == len(lits) and np.abs(lits).max() == E * cd


def test_literal_order_of_one_clause():
    """First clause of R(3,3,3) excludes color 0 on triangle (0, 1, 2): bit 0
    over edges (0,1), (0,2), (1,2), then bit 1, all positive (both bits 0)."""
    m = RamseyClauseMaker(4, [3, 3, 3])
    c = m.GetClauses()
    e01, e02, e12 = m.E[0, 1], m.E[0, 2], m.E[1, 2]
    assert c[0] == [int(e01) * 2 + 1, int(e02) * 2 + 1, int(e12) * 2 + 1, int(e01) * 2 + 2, int(e02) * 2 + 2,
                    int(e12) * 2 + 2]
    # color 1 (bits 1, 0): bit 0 literals negated
    n0 = math.comb(4, 3)
    assert c[n0] == [-(int(e01) * 2 + 1), -(int(e02) * 2 + 1), -(int(e12) * 2 + 1), int(e01) * 2 + 2,
                     int(e02) * 2 + 2, int(e12) * 2 + 2]
    # color 3 is excluded on every edge: (-bit0, -bit1)
    assert c[3 * n0] == [-1, -2]


@pytest.mark.parametrize("N,bs", [(5, 5), (6, 3), (8, 4), (9, 3)])
def test_circular_adjacency(N, bs):
    m = RamseyClauseMaker(N, [3, 3], circular=True, blocksize=bs)
    E = m.E_circular
    assert (np.diag(E) == -1).all() and (E == E.T).all()
    off = ~np.eye(N, dtype=bool)
    assert (E[off] >= 0).all() and E.max() == m.edgeCount - 1
    nb = N // bs
    assert m.edgeCount == nb * (bs - 1) + nb * (nb - 1) // 2 * bs
    # every entry off the block's first column repeats the entry one row up,
    # one column left (cyclic within the block), or is the diagonal
    for a in range(nb):
        for b in range(a, nb):
            for j in range(1, bs):
                for i in range(bs):
                    r, c = a * bs + i, b * bs + j
                    if r != c:
                        assert E[r, c] == E[a * bs + (i - 1) % bs, b * bs + j - 1] or E[c, r] == E[r, c]
    with pytest.raises(ValueError):
        RamseyClauseMaker(N, [3, 3], circular=True, blocksize=bs + 1 if N % (bs + 1) else 4 * N)


def test_circular_solutions_map_to_regular_solutions():
    circ = RamseyClauseMaker(5, [3, 3], circular=True)
    reg = RamseyClauseMaker(5, [3, 3]).GetClauses()
    cc = circ.GetClauses()
    nv = circ.edgeCount
    found = 0
    for b in range(1 << nv):
        x = np.array([1.0 if (b >> v) & 1 else -1.0 for v in range(nv)])
        if satisfied(cc, x):
            found += 1
            assert satisfied(reg, circ.ApplyCircularToRegular(x))
    assert found >= 1


def test_specific_color_and_initial_condition():
    m = RamseyClauseMaker(5, [3, 3, 3])
    before = len(m.GetClauses())
    m.SetSpecificColor(1, 3, 2)
    c = m.GetClauses()[before:]
    e = int(m.E[1, 3])
    assert c == [[-(e * 2 + 1)], [e * 2 + 2]]
    ic = m.SetInitialCondition(1, 3, 2, [])
    assert len(ic) == e * 2 + 2 and ic[e * 2] == -1.0 and ic[e * 2 + 1] == 1.0


def test_make_ramsey_files(tmp_path):
    conf = cli.parse(["make_ramsey", "-ramsey", "3,4", "-nstart", "5", "-nend", "6", "-problemfolder",
                      str(tmp_path), "-ramseycircular"])
    from analogsat_amd.ramsey import make_ramsey_problems
    import io
    out = io.StringIO()
    paths = make_ramsey_problems(conf, out)
    assert [os.path.basename(p) for p in paths] == ["ramsey_circular_34_N5.cnf", "ramsey_circular_34_N6.cnf"]
    assert "Saving ramsey_circular_34_N5.cnf" in out.getvalue() and "clauses:" in out.getvalue()
    from analogsat_amd import SatProblem
    for N, p in zip((5, 6), paths):
        lits, lens = RamseyClauseMaker(N, [3, 4], circular=True).clause_arrays()
        q = SatProblem.read_cnf(p)
        ql, qn = q.literals, q.lengths
        assert np.array_equal(np.asarray(qn), lens) and np.array_equal(np.asarray(ql)[np.asarray(ql) != 0], lits)
    assert ramsey_file_name([5, 5], 43, False) == "ramsey_regular_55_N43.cnf"
    assert parse_ramsey("3,3,3,3") == [3, 3, 3, 3]
    for bad in (["make_ramsey", "-ramsey", "3"], ["run_ramsey", "-ramsey", "3,1"], ["make_ramsey", "-ramsey", "x,y"],
                ["make_ramsey", "-ramsey", "3,3", "-nstart", "1"]):
        with pytest.raises(SystemExit):
            cli.parse(bad)


@pytest.mark.gpu
def test_run_ramsey_series_on_gpu(tmp_path):
    """run_ramsey on R(3,3) with N = 5 (satisfiable, with bias) and N = 6
    (unsatisfiable: stops at the step limit): result lines, and the written
    coloring has no monochromatic triangle."""
    import io
    from analogsat_amd.ramsey import make_ramsey_problems, run_ramsey_series
    pf, rf = tmp_path / "p", tmp_path / "r"
    conf = cli.parse(["make_ramsey", "-ramsey", "3,3", "-nstart", "5", "-nend", "6", "-problemfolder", str(pf)])
    make_ramsey_problems(conf, io.StringIO())
    conf = cli.parse(["run_ramsey", "-ramsey", "3,3", "-nstart", "5", "-nend", "6", "-problemfolder", str(pf),
                      "-resultfolder", str(rf), "-samplestart", "0", "-sampleend", "2", "-stepmax", "3000",
                      "-bias", "0.5", "-seed", "11"])
    print("[ramsey] run_ramsey R(3,3) N=5..6", flush=True)
    out = io.StringIO()
    solved = run_ramsey_series(conf, out)
    assert [n for n, _ in solved] == [5, 5], out.getvalue()
    for N in (5, 6):
        lines = open(rf / ("perf_33_N%d.dat" % N)).read().splitlines()
        assert len(lines) == 2
        for line in lines:
            f = line.split("\t")
            assert f[0] == "ramsey_regular_33_N%d.cnf" % N and f[2] == "analogsat1orig" and len(f) == 7
            assert (f[3] == "solution found") == (N == 5), f
    M = np.loadtxt(rf / "coloring_33_N5.dat", dtype=int)
    assert M.shape == (5, 5) and (np.diag(M) == 0).all() and (M == M.T).all()
    colors = M - 1
    assert valid_coloring(colors, (3, 3))


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("N", [16, 20, 23, 24])
def test_dense_ramsey_problem_bitexact_vs_oracle(precision, N):
    """R(5,5) encodings have few variables with long rows (every variable has
    20 C(N,5) / C(N,2) contributions: 728, 1632, 2660 and 3080 at N = 16, 20,
    23, 24), too long for the LDS var pass: the one-workgroup-per-variable var
    pass.  Its rows pass through LDS in chunks of 2048 values; at N = 23 and
    24 a row spans two chunks (the second staged while the first is summed),
    and the last chunk (612, 1032 values) is not a multiple of the 8-vector
    read group.  RHS, 20 steps, attempts and RHS counts bit-identical to the
    oracle."""
    import analogsat_amd as A
    import oracle as O
    print("[ramsey] dense R(5,5) N=%d %s" % (N, precision), flush=True)
    lits, lens = RamseyClauseMaker(N, [5, 5]).clause_arrays()
    p = O.Problem(lits, lens)
    sat = A.HipSat(precision)
    sat.SetProblem(A.SatProblem(lits, lens))
    n, m = sat.Get_N(), sat.Get_M()
    assert sat.Layout()["var_pass"] == 2          # one workgroup per variable
    rows = np.bincount(np.abs(lits), minlength=n + 1)[1:]
    assert rows.min() == rows.max() == 20 * comb(N, 5) // comb(N, 2)
    mode = "f64" if precision == "f64" else "mix"
    dt = np.float64 if precision == "f64" else np.float32
    pub = O.random_state_f64(N, n, m).astype(dt)
    x, d = sat.MakeState(), sat.MakeState()
    x.CopyFrom(pub)
    sat.GetDerivatives(d, x, 0.0)
    want = O.rhs(p, np.concatenate([[-1.0], pub]).astype(dt), mode=mode)[1:]
    assert np.array_equal(d.CopyTo(), want)
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
        st = rk.stats()
        assert (st.hdid, st.attempts, st.rhscount) == (orc.hdid, orc.attempts, orc.rhscount)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_dense_ramsey_tanh_with_bias(precision):
    """The reference's Ramsey runs use the tanh family with a bias term
    (Frontend/Config/run_ramsey_55_tanh.cfg: -tanh -bias 500).  On the dense
    R(5,5) N = 16 encoding (one-workgroup-per-variable var pass), against the
    oracle with the GPU's definition of the bias mean: RHS within 16 ulp of
    the largest per-variable sum of |slot term| (the device tanh is not
    glibc's bit for bit), and 20 steps with the same attempts and states
    within 1e-9 (FP64) / 1e-4 (FP32) relative."""
    import analogsat_amd as A
    import oracle as O
    print("[ramsey] dense R(5,5) N=16 tanh bias 500 %s" % precision, flush=True)
    q, b = 1.1, 500.0
    lits, lens = RamseyClauseMaker(16, [5, 5]).clause_arrays()
    p = O.Problem(lits, lens)
    sat = A.HipSatTanh(precision, q=q)
    sat.SetProblem(A.SatProblem(lits, lens))
    sat.Set_B(b)
    n, m = sat.Get_N(), sat.Get_M()
    mode = "f64" if precision == "f64" else "mix"
    dt = np.float64 if precision == "f64" else np.float32
    pub = O.random_state_f64(16, n, m).astype(dt)
    x, d = sat.MakeState(), sat.MakeState()
    x.CopyFrom(pub)
    sat.GetDerivatives(d, x, 0.0)
    xi = np.concatenate([[-1.0], pub]).astype(dt)
    with O.mean_mode("double"):
        want = O.rhs(p, xi, mode=mode, b=b, family="tanh", q=q)[1:].astype(np.float64)
    got = d.CopyTo().astype(np.float64)
    # largest per-variable sum of |slot term| (bounds the rounding of the sums)
    k = int(lens.max())
    C = np.zeros((m, k), np.int64)
    off = np.concatenate([[0], np.cumsum(lens)])
    for j in range(m):
        C[j, :lens[j]] = lits[off[j]:off[j + 1]]
    xs = np.asarray(xi, np.float64)
    s = np.where(C != 0, xs[np.abs(C)], -1.0)
    f = np.where(C < 0, 1.0 + s, 1.0 - s)
    a = xs[n + 1:n + 1 + m]
    tot = np.zeros(n + 1)
    for qi in range(k):
        prod = np.prod(np.delete(f, qi, axis=1), axis=1)
        term = 2.0 ** -k * a * np.abs(prod) * np.abs(np.tanh(q * f[:, qi]))
        np.add.at(tot, np.abs(C[:, qi]), np.where(C[:, qi] != 0, term, 0.0))
    scale = max(tot[1:].max(), np.abs(want[:n]).max())
    ulp = 2.0 ** -52 if precision == "f64" else 2.0 ** -23
    err = np.abs(got - want).max()
    print("[ramsey]   max |diff| %.3e = %.2f ulp of scale" % (err, err / (ulp * scale)), flush=True)
    assert err <= 16 * ulp * scale
    rk = A.HipRungeKutta()
    rk.SetEpsilon(1e-6)
    rk.Configure(sat)
    old, new = sat.MakeState(), sat.MakeState()
    old.CopyFrom(pub)
    with O.pow_mode("cr"), O.mean_mode("double"):
        orc = O.RK(p, eps=1e-6, mode=mode, b=b, family="tanh", q=q)
        orc.set_state(pub)
        rk.Step(old, new, 20)
        orc.step(20)
    got_x, ref = old.CopyTo().astype(np.float64), orc.get_state().astype(np.float64)
    rel = np.abs(got_x - ref).max() / np.abs(ref).max()
    print("[ramsey]   rel diff %.3e, attempts %d vs %d" % (rel, rk.stats().attempts, orc.attempts), flush=True)
    assert rk.stats().steps == 20 and rk.stats().attempts == orc.attempts
    assert rel <= (1e-9 if precision == "f64" else 1e-4)
