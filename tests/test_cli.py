"""CLI mirror of the reference Frontend's ``run`` command: option parsing on
the CPU; a solve of the C1 fixture on the GPU (result file in MiniSat form,
the assignment satisfies every clause)."""
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
                ["run", "-problem", "x", "-nogpu"], ["run"], ["bench"]):
        with pytest.raises(SystemExit):
            cli.parse(bad)


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
