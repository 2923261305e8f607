"""C-ABI checks that need no GPU: the library loads, exports every symbol the
headers in include/ declare, and its host-only logic behaves like the
reference (FastSatMaker stream, argument validation, error codes)."""
import ctypes
import json
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import analogsat_amd as A

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(asat_\w+)\s*\(", text))
    return names


@pytest.fixture(scope="module")
def L():
    return A.lib()


def test_exports_every_declared_symbol(L):
    names = declared_symbols()
    assert len(names) > 30
    out = subprocess.run(["nm", "-D", "--defined-only", A.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (asat_\w+)", out))
    missing = names - exported
    assert not missing, missing
    for n in names:
        assert hasattr(L, n)


def test_library_is_gfx950_code_object():
    # the offload bundle of the library built here names its target triple
    blob = open(A.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob


def test_one_hip_runtime_with_torch():
    """Loading the library and then importing torch maps one HIP runtime and
    one HSA runtime (torch's wheel carries its own copies, which it would
    otherwise load beside ours).  Loading and importing make no GPU call."""
    pytest.importorskip("torch")
    code = ("import sys; sys.path.insert(0, %r)\n"
            "import analogsat_amd as A; A.lib()\n"
            "import torch\n"
            "import json; print(json.dumps([sorted(A._mapped('libamdhip64.so')), sorted(A._mapped('libhsa-runtime64.so'))]))\n"
            % os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    hip, hsa = json.loads(out.stdout.strip().splitlines()[-1])
    assert len(hip) == 1 and len(hsa) == 1, out.stdout


def test_version_and_device_count(L):
    assert b"gfx950" in L.asat_version()
    assert L.asat_device_count() >= 0


def test_fastsat_matches_golden_stream(golden_dir):
    d = np.load(os.path.join(golden_dir, "gen_cases.npz"), allow_pickle=False)
    for key in d.files:
        if not key.startswith("fastsat_"):
            continue
        n, m, k, sd = map(int, key.split("_")[1:])
        p = A.fastsat(n, m, k, sd)
        assert np.array_equal(p.literals, d[key]), key


def test_fastsat_stream_continues_like_one_generator():
    # two problems from one maker == the first 2m clauses of one long stream
    mk = A.FastSatMaker(17)
    mk.Set_N(40); mk.Set_M(50); mk.Set_K(3)
    p1 = mk.MakeSatProblem()
    p2 = mk.MakeSatProblem()
    both = A.fastsat(40, 100, 3, 17)
    assert np.array_equal(np.concatenate([p1.literals, p2.literals]), both.literals)


def test_fastsat_argument_errors(L):
    out = np.zeros(30, np.int32)
    assert L.asat_problem_fastsat(5, 10, 6, 1, out.ctypes.data) == A.ASAT_ERR_INVALID_ARGUMENT
    assert b"fastsat" in L.asat_last_error()


def test_rk_validation_without_gpu():
    rk = A.HipRungeKutta()
    with pytest.raises(ValueError):
        rk.SetEpsilon(2.0)
    with pytest.raises(ValueError):
        rk.SetEpsilon(1e-17)
    rk.SetEpsilon(1e-6)
    assert rk.GetEpsilon() == 1e-6
    with pytest.raises(RuntimeError, match="configured"):
        rk.stats()


def test_null_arguments_rejected(L):
    assert L.asat_sys_create(0, None) == A.ASAT_ERR_INVALID_ARGUMENT
    assert L.asat_sys_create(7, ctypes.byref(ctypes.c_void_p())) == A.ASAT_ERR_INVALID_ARGUMENT
    assert L.asat_rk_step(None, None, None, 1) == A.ASAT_ERR_INVALID_ARGUMENT


def test_cnf_roundtrip(tmp_path, golden_dir):
    p = A.SatProblem.read_cnf(os.path.join(golden_dir, "c1_n50_m213_s0.cnf"))
    assert (p.Get_N(), p.Get_M(), p.Get_K()) == (50, 213, 3)
    q = A.SatProblem.from_clauses([[1, 0, -3], [], [2, 5], [0]])
    assert (q.Get_N(), q.Get_M(), q.Get_K()) == (5, 2, 2)
    f = tmp_path / "x.cnf"
    q.write_cnf(str(f))
    r = A.SatProblem.read_cnf(str(f))
    assert list(r.literals) == [1, -3, 2, 5]
    assert list(r.lengths) == [2, 2]


def test_solver_setters_validate():
    s = A.SatSolver(None, None)
    with pytest.raises(ValueError):
        s.SetBatchSize(0)
    with pytest.raises(ValueError):
        s.SetMaxTime(0)
    with pytest.raises(ValueError):
        s.SetMaxSteps(0)
    assert (s.GetMaxTime(), s.GetMaxSteps(), s.GetBatchSize()) == (100.0, 100000, 20)
    assert A.GetSatResultMessage(A.SatResult.SAT_SOLUTION_FOUND) == "solution found"


def test_randomsat_matches_golden_stream(golden_dir):
    d = np.load(os.path.join(golden_dir, "randomsat_cases.npz"), allow_pickle=False)
    for key in d.files:
        n, m, k, sd = map(int, key.split("_")[1:])
        p = A.randomsat(n, m, k, sd)
        assert np.array_equal(p.literals, d[key]), key


def test_randomsat_maker_stream_and_errors():
    # one maker making two problems continues one stream, as the reference object does
    mk = A.RandomSatMaker(7)
    mk.Set_N(30); mk.Set_M(120); mk.Set_K(3)
    a, b = mk.MakeSatProblem(), mk.MakeSatProblem()
    assert not np.array_equal(a.literals, b.literals)
    assert np.array_equal(a.literals, A.randomsat(30, 120, 3, 7).literals)
    with pytest.raises(ValueError):
        A.randomsat(4, 9, 1, 0)      # more clauses than the 8 distinct ones
