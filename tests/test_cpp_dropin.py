"""The C++ drop-in headers: the reference caller's solve sequence compiles
against include/asat/ and links against libanalogsat_hip.so (CPU), and on a
GPU reproduces the reference CPU solver's solve of the C1 fixture (gpu)."""
import os
import subprocess

import numpy as np
import pytest

import analogsat_amd as A

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "dropin_run.cpp")


def build_dropin(tmpdir):
    exe = os.path.join(str(tmpdir), "dropin_run")
    libdir = os.path.dirname(A.LIB_PATH)
    cmd = ["g++", "-std=c++14", "-O1", "-Wall", "-Werror", "-I" + os.path.join(ROOT, "include"), SRC, "-o", exe,
           "-L" + libdir, "-lanalogsat_hip", "-Wl,-rpath," + libdir]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe


def test_dropin_compiles_and_links(tmp_path):
    if not os.path.exists(os.path.join(os.path.dirname(A.LIB_PATH), "libanalogsat_hip.so")):
        pytest.skip("library not built")
    exe = build_dropin(tmp_path)
    assert os.path.exists(exe)


@pytest.mark.gpu
def test_dropin_solves_c1_like_reference(tmp_path, golden_dir):
    exe = build_dropin(tmp_path)
    g = np.load(os.path.join(golden_dir, "c1_traj.npz"), allow_pickle=False)
    d = np.load(os.path.join(golden_dir, "c1_solve.npz"), allow_pickle=False)
    start = tmp_path / "start.txt"
    start.write_text("\n".join(repr(float(v)) for v in g["pub0"]))
    out = subprocess.run([exe, os.path.join(golden_dir, "c1_n50_m213_s0.cnf"), "1e-6", "1", "10000", str(start)],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.strip().splitlines()
    assert lines[0] == "solution found"
    steps, rhs, viol, minv, batches = map(int, lines[1].split())
    assert (steps, rhs, viol) == (int(d["steps"]), int(d["rhscount"]), 0)
    assert batches == len(d["batch_violations"])
    signs = [int(v) for v in lines[2].split()[:-1]]
    want = [i + 1 if x > 0 else -(i + 1) for i, x in enumerate(d["final"][:50])]
    assert signs == want


@pytest.mark.gpu
def test_dropin_without_callback_solves_c1_like_reference(tmp_path, golden_dir):
    """The same caller without a callback: SatSolver hands the loop to the
    integrator (asat_rk_solve), which runs it on the device for this size."""
    exe = build_dropin(tmp_path)
    g = np.load(os.path.join(golden_dir, "c1_traj.npz"), allow_pickle=False)
    d = np.load(os.path.join(golden_dir, "c1_solve.npz"), allow_pickle=False)
    start = tmp_path / "start.txt"
    start.write_text("\n".join(repr(float(v)) for v in g["pub0"]))
    out = subprocess.run([exe, os.path.join(golden_dir, "c1_n50_m213_s0.cnf"), "1e-6", "1", "10000", str(start),
                          "nocallback"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.strip().splitlines()
    assert lines[0] == "solution found"
    steps, rhs, viol, minv, batches = map(int, lines[1].split())
    assert (steps, rhs, viol, minv, batches) == (int(d["steps"]), int(d["rhscount"]), 0, -1, -1)
    signs = [int(v) for v in lines[2].split()[:-1]]
    want = [i + 1 if x > 0 else -(i + 1) for i, x in enumerate(d["final"][:50])]
    assert signs == want
