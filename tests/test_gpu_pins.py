"""GPU checks of the device random stream and of the FP32 RHS against the
reference's float instantiation (through the C ABI).

* asat_rng_uniform (ISatRandom::GenerateUniform, replacing CudaRandom's cuRAND
  stream, AnalogSAT/cuda_base/CudaRandom.cu:24-60) equals the host Philox4x32-10
  stream of the oracle, whose round function is pinned by the published
  Random123 known-answer vectors: same counter/key packing, same conversion to
  double and float, and consecutive draws continue the stream
* asat_sys_set_random_state (CpuSat::SetRandomState, AnalogSAT/cpu_sat/CpuSat.cpp:122-133)
  maps that stream to s = 2u - 1, a = u + 1 element for element, in the
  public clause order
* the FP32 RHS equals the reference's CpuSat<float> outputs bit for bit
"""
import os
import sys

import numpy as np
import pytest

import analogsat_amd as A
import oracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print("[gpu-pins]", msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_rng_uniform_equals_host_philox(dtype):
    torch = pytest.importorskip("torch")
    say("rng uniform %s" % np.dtype(dtype).name)
    seed = (1 << 37) + 12345
    rng = A.HipRandom(seed)
    tdt = torch.float64 if dtype == np.float64 else torch.float32
    prec = "f64" if dtype == np.float64 else "f32"
    a = torch.zeros(100003, dtype=tdt, device="cuda")
    b = torch.zeros(777, dtype=tdt, device="cuda")
    rng.GenerateUniform(a.data_ptr(), a.numel(), prec)
    rng.GenerateUniform(b.data_ptr(), b.numel(), prec)
    got = np.concatenate([a.cpu().numpy(), b.cpu().numpy()])
    want = O.philox_uniform(seed, 0, len(got), dtype)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_set_random_state_maps_the_stream(precision):
    say("set random state %s" % precision)
    p = O.fastsat(300, 1278, 3, 4)
    sat = A.HipSat(precision)
    sat.SetProblem(A.SatProblem(p.literals, p.lengths))
    dt = np.float64 if precision == "f64" else np.float32
    rng = A.HipRandom(99)
    first, second = sat.MakeState(), sat.MakeState()
    sat.SetRandomState(first, rng)
    sat.SetRandomState(second, rng)          # continues the stream
    D = p.n + p.m
    u = O.philox_uniform(99, 0, 2 * D, dt)
    for st, uu in ((first, u[:D]), (second, u[D:])):
        want = np.concatenate([uu[:p.n] * dt(2.0) - dt(1.0), uu[p.n:] + dt(1.0)]).astype(dt)
        assert np.array_equal(st.CopyTo(), want)


@pytest.mark.parametrize("case", ["k3", "k4", "mixed", "k12", "n2000"])
def test_rhs_f32_matches_reference_float_instantiation(golden_dir, case):
    say("rhs f32 vs CpuSat<float> case=%s" % case)
    d = np.load(os.path.join(golden_dir, "rhs_cases.npz"), allow_pickle=False)
    f = np.load(os.path.join(golden_dir, "rhs_cases_f32.npz"), allow_pickle=False)
    sat = A.HipSat("f32")
    sat.SetProblem(A.SatProblem(d[case + "_lits"], d[case + "_lens"]))
    x, dx = sat.MakeState(), sat.MakeState()
    x.CopyFrom(d[case + "_x"][1:].astype(np.float32))
    sat.GetDerivatives(dx, x, 0.0)
    assert np.array_equal(dx.CopyTo(), f[case + "_rhs32"][1:])
