"""CLI mirror of the reference Frontend's ``run``, ``bench`` and ``speedtest``
commands: option and -config parsing, problem-size series, sample naming and
the bench/speedtest bookkeeping on the CPU (solver replaced by a stub); a
solve of the C1 fixture and a short bench + speedtest series on the GPU."""
import glob
import io
import os

import numpy as np
import pytest

from analogsat_amd import __main__ as cli


def test_parse_defaults_and_ranges():
    c = cli.parse(["run", "-problem", "x.cnf"])
    assert (c["eps"], c["batch"], c["stepmax"], c["tmax"]) == (1e-6, 50, 50000000, 1e8)
    c = cli.parse(["run", "-problem", "x.cnf", "-eps", "1e-4", "-batch", "7", "-float"])
    assert c["eps"] == 1e-4 and c["batch"] == 7 and c["float"]
    for bad in (["run", "-problem", "x", "-eps", "1"], ["run", "-problem", "x", "-batch", "0"],
                ["run", "-problem", "x", "-nogpu"], ["run"], ["make_ramsey"], ["bench", "-minisat"],
                ["bench", "-version", "4"], ["bench", "-k", "1"], ["bench", "-samplestart", "10000"],
                ["bench", "-nend", "0.5"], ["bench", "-usegpu", "1"], ["bench", "-alpha"], ["bench", "-unknown", "1"]):
        with pytest.raises(SystemExit):
            cli.parse(bad)
    c = cli.parse(["bench"])     # Frontend/src/main.cpp:205-245 defaults
    assert (c["k"], c["alpha"], c["samplestart"], c["sampleend"], c["timeout"]) == (3, 4.25, 0, 100, 3600.0)
    assert (c["nstart"], c["nend"], c["nstep"], c["version"]) == (1.0, 5.0, 0.5, 1)
    assert not c["rerun"] and not c["force_save_cnf"]


def test_config_files_recursive(tmp_path):
    """-config: one option per line, split at the first space; files nest;
    later options override earlier ones (Frontend/src/main.cpp:140-163)."""
    inner = tmp_path / "inner.cfg"
    inner.write_text("-alpha 3.8\r\n-tanh\n-problemfolder my problems\n-nstart 2.5")
    outer = tmp_path / "outer.cfg"
    outer.write_text("-k 4\n-config %s\n-timeout 1.5\n-rerun\n" % inner)
    c = cli.parse(["speedtest", "-alpha", "2.0", "-config", str(outer), "-sampleend", "7"])
    assert c["k"] == 4 and c["alpha"] == 3.8 and c["tanh"] and c["rerun"]
    assert c["problemfolder"] == "my problems" and c["nstart"] == 2.5 and c["timeout"] == 1.5
    assert c["sampleend"] == 7 and c["command"] == "speedtest"
    bad = tmp_path / "bad.cfg"
    bad.write_text("-alpha 3.8\n-nogpu\n")
    with pytest.raises(SystemExit):
        cli.parse(["bench", "-config", str(bad)])
    loop = tmp_path / "loop.cfg"
    loop.write_text("-config %s\n" % loop)
    with pytest.raises(SystemExit):
        cli.parse(["bench", "-config", str(loop)])


REF_CONFIGS = "/root/reference/Frontend/Config"


@pytest.mark.skipif(not os.path.isdir(REF_CONFIGS), reason="reference config files not present")
def test_reference_gpu_configs_parse():
    """Every GPU config file the reference ships (Ramsey ones included)
    parses with the same values; the CPU-solver and MiniSat ones are refused."""
    n_gpu = n_ramsey = 0
    for path in sorted(glob.glob(os.path.join(REF_CONFIGS, "*.cfg"))):
        text = open(path).read()
        if any(o in text for o in ("-nogpu", "-minisat")):
            with pytest.raises(SystemExit):
                cli.parse(["bench", "-config", path])
            continue
        ramsey = "-ramsey" in text
        c = cli.parse(["run_ramsey" if ramsey else "speedtest", "-config", path])
        for line in text.splitlines():
            opt, _, val = line.partition(" ")
            if val:
                want = val if opt in ("-problemfolder", "-resultfolder", "-problem", "-ramsey") else float(val)
                assert c[opt[1:]] == want, (path, opt)
            elif opt:
                assert c[opt[1:]] is True, (path, opt)
        n_gpu += 1
        n_ramsey += ramsey
    assert n_gpu >= 10 and n_ramsey >= 6


def test_series_sizes():
    """N = round(10 * 2^p), p from nstart to nend (+1e-3) by nstep."""
    c = cli.parse(["bench"])
    assert cli.n_values(c) == [20, 28, 40, 57, 80, 113, 160, 226, 320]
    c = cli.parse(["bench", "-nstart", "1", "-nend", "1.9995", "-nstep", "1"])
    assert cli.n_values(c) == [20, 40]
    c = cli.parse(["bench", "-nstart", "12.5", "-nend", "12.5"])
    assert cli.n_values(c) == [57926]
    with pytest.raises(SystemExit):
        cli.n_values(cli.parse(["bench", "-nstep", "0"]))


class _StubRunner(cli.BenchRunner):
    """Host bookkeeping only: the solve is a stub that reports SAT for even
    sample numbers and writes a result line of the reference's form."""

    def _run_gpu(self):
        solved = int(self.problem_name[-8:-4]) % 2 == 0
        with open(self.outfile, "a") as f:
            f.write("%s\tstub\tanalogsat1orig\t%s\t0.1\t0.1\t60\n"
                    % (self.problem_name, "solution found" if solved else "user interrupt"))
        self.last_solve_time = 0.5
        self.calls = getattr(self, "calls", 0) + 1
        return solved


def test_bench_bookkeeping(tmp_path, monkeypatch):
    pf, rf = tmp_path / "problems", tmp_path / "results"
    conf = cli.parse(["bench", "-nstart", "1", "-nend", "1.5", "-sampleend", "4", "-seed", "5",
                      "-problemfolder", str(pf), "-resultfolder", str(rf)])
    monkeypatch.setattr(cli, "BenchRunner", _StubRunner)
    out = io.StringIO()
    cli.run_bench_series(conf, out=out)
    for n, m in ((20, 85), (28, 119)):
        lines = (rf / ("perf_K3_a425_N%d.dat" % n)).read_text().splitlines()
        assert [ln.split("\t")[0] for ln in lines] == ["3a425/N%d/p%04d.cnf" % (n, s) for s in range(4)]
        saved = sorted(os.listdir(pf / "3a425" / ("N%d" % n)))
        assert saved == ["p0000.cnf", "p0002.cnf"]           # kept once found SAT
        import analogsat_amd as A
        p = A.SatProblem.read_cnf(str(pf / "3a425" / ("N%d" % n) / "p0000.cnf"))
        assert (p.Get_M(), p.Get_K()) == (m, 3) and p.Get_N() <= n
    assert "Making K=3, N=20, M=85" in out.getvalue()
    # second pass: every sample with a line in the result file (solved or not) is skipped
    cli.run_bench_series(conf, out=io.StringIO())
    assert len((rf / "perf_K3_a425_N20.dat").read_text().splitlines()) == 4
    # a sample whose line is gone runs again, from a fresh instance (its CNF was never saved)
    lines = (rf / "perf_K3_a425_N20.dat").read_text().splitlines()
    (rf / "perf_K3_a425_N20.dat").write_text("\n".join(lines[:3]) + "\n")
    out = io.StringIO()
    cli.run_bench_series(conf, out=out)
    lines = (rf / "perf_K3_a425_N20.dat").read_text().splitlines()
    assert len(lines) == 4 and lines[3].startswith("3a425/N20/p0003.cnf\t")
    assert "Making K=3, N=20, M=85" in out.getvalue()
    # -rerun runs every sample; existing CNFs are loaded, not remade
    before = (pf / "3a425" / "N20" / "p0000.cnf").read_text()
    cli.run_bench_series(dict(conf, rerun=True), out=io.StringIO())
    assert len((rf / "perf_K3_a425_N20.dat").read_text().splitlines()) == 8
    assert (pf / "3a425" / "N20" / "p0000.cnf").read_text() == before
    # -force_save_cnf keeps every sample
    cli.run_bench_series(dict(conf, force_save_cnf=True, sampleend=6, samplestart=4), out=io.StringIO())
    assert sorted(os.listdir(pf / "3a425" / "N20"))[-2:] == ["p0004.cnf", "p0005.cnf"]


def test_speedtest_bookkeeping(tmp_path, monkeypatch):
    """Speedtest cycles the existing samples (0, 2, 4 .. as saved) until the
    summed solve time reaches the limit, and refuses an empty sample set."""
    pf, rf = tmp_path / "problems", tmp_path / "results"
    conf = cli.parse(["bench", "-nstart", "1", "-nend", "1", "-sampleend", "5", "-seed", "2",
                      "-problemfolder", str(pf), "-resultfolder", str(rf)])
    monkeypatch.setattr(cli, "BenchRunner", _StubRunner)
    cli.run_bench_series(conf, out=io.StringIO())          # saves p0000, p0002, p0004
    (rf / "perf_K3_a425_N20.dat").unlink()
    cli.run_speedtest(conf, walltimeout=3.0, out=io.StringIO())
    names = [ln.split("\t")[0][-9:] for ln in (rf / "perf_K3_a425_N20.dat").read_text().splitlines()]
    assert names == ["p0000.cnf", "p0000.cnf", "p0000.cnf", "p0000.cnf", "p0000.cnf", "p0000.cnf"]
    with pytest.raises(SystemExit):
        cli.run_speedtest(dict(conf, alpha=3.0), out=io.StringIO())


@pytest.mark.gpu
def test_cli_solves_c1(tmp_path, golden_dir):
    conf = cli.parse(["run", "-problem", os.path.join(golden_dir, "c1_n50_m213_s0.cnf"), "-batch", "5",
                      "-resultfolder", str(tmp_path), "-seed", "3", "-stepmax", "20000"])
    buf = io.StringIO()
    res = cli.run(conf, out=buf)
    lines = (tmp_path / "c1_n50_m213_s0.out").read_text().splitlines()
    assert lines[0] == ("SAT" if int(res) == 0 else "INDET")
    assert "RK steps/sec" in buf.getvalue()
    if lines[0] == "SAT":
        assign = {abs(int(v)): int(v) > 0 for v in lines[1].split()[:-1]}
        lits = [int(t) for ln in open(os.path.join(golden_dir, "c1_n50_m213_s0.cnf")) if not ln.startswith("p")
                for t in ln.split()]
        clause, ok = [], True
        for l in lits:
            if l == 0:
                ok &= any(assign[abs(x)] == (x > 0) for x in clause)
                clause = []
            else:
                clause.append(l)
        assert ok


@pytest.mark.gpu
def test_cli_bench_and_speedtest_gpu(tmp_path):
    """A short bench series (N = 20, 28; three samples each) and a speedtest
    over the saved samples, with the MI355X solver."""
    import analogsat_amd as A
    pf, rf = tmp_path / "problems", tmp_path / "results"
    conf = cli.parse(["bench", "-nstart", "1", "-nend", "1.5", "-sampleend", "3", "-seed", "11", "-batch", "10",
                      "-stepmax", "20000", "-timeout", "5", "-force_save_cnf",
                      "-problemfolder", str(pf), "-resultfolder", str(rf)])
    out = io.StringIO()
    cli.run_bench_series(conf, out=out)
    msgs = set(A.GetSatResultMessage(r) for r in A.SatResult)
    gpu = A.GetGpuName()
    for n in (20, 28):
        lines = (rf / ("perf_K3_a425_N%d.dat" % n)).read_text().splitlines()
        assert len(lines) == 3
        for s, ln in enumerate(lines):
            f = ln.split("\t")
            assert f[0] == "3a425/N%d/p%04d.cnf" % (n, s) and f[1] == gpu and f[2] == "analogsat1orig"
            assert f[3] in msgs and float(f[4]) > 0 and float(f[5]) >= 0 and int(f[6]) >= 6
            p = A.SatProblem.read_cnf(str(pf / "3a425" / ("N%d" % n) / ("p%04d.cnf" % s)))
            assert p.Get_M() == int(4.25 * n)
    assert "solution found" in out.getvalue()
    cli.run_speedtest(dict(conf, nend=1.0, tanh=True), walltimeout=0.3, out=io.StringIO())
    lines = (rf / "perf_K3_a425_N20.dat").read_text().splitlines()
    assert len(lines) > 3 and all(ln.split("\t")[2] == "analogsat1tanh" for ln in lines[3:])
