"""Build libanalogsat_hip.so in-tree with hipcc for gfx950.

Every source is compiled to an object of its own (in parallel, under
analogsat_amd/_build/), then the objects are linked into the shared library.
All HIP sources are compiled with
  -ffp-contract=off   no fused multiply-add, so the FP64 path rounds like the
                      reference's plain -O3 host build
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      keep the compiler from emitting packed FP32 instructions
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libanalogsat_hip.so")
OBJDIR = os.path.join(HERE, "_build")

SOURCES = ["asat_plain.hip", "asat_batch.hip", "asat_adjoint.hip", "asat_noalloc.hip", "asat_problem.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("ASAT_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    "--offload-arch=" + ARCH,
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    "-I" + os.path.join(ROOT, "include"),
    "-I" + CSRC,
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed: %s" % " ".join(cmd[-3:]))


def build(verbose=False):
    os.makedirs(OBJDIR, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    objs = [os.path.join(OBJDIR, os.path.splitext(s)[0] + ".o") for s in srcs]
    jobs = [[HIPCC] + FLAGS + ["-c", "-o", o, os.path.join(CSRC, s)] for s, o in zip(srcs, objs)]
    with concurrent.futures.ThreadPoolExecutor(max(1, min(len(jobs), os.cpu_count() or 1))) as ex:
        for f in [ex.submit(_run, j, verbose) for j in jobs]:
            f.result()
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs, verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
