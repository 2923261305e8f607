"""GPU discrete adjoint vs the oracle (C ABI through the Python mirror).

* forward: logged step sizes == the oracle's mixed-mode adaptive run, bit for bit
* replay with the logged steps reproduces V(T) and x(T) of the forward exactly
* gradient == the oracle's FP32-forward / FP64-adjoint on the same steps
  (max relative difference <= 1e-12; the only difference is the order of the
  final V tree-sum, not used by the gradient)
* gradient vs central finite differences of the FP64 oracle with the same
  frozen step sequence (h = 1e-4): agreement within the FP32-forward error
* the checkpoint interval does not change the gradient
* torch.autograd.Function returns the C ABI gradient
"""
import os
import sys

import numpy as np
import pytest

import analogsat_amd as A
import oracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print("[gpu-adjoint]", msg, flush=True)
    sys.stdout.flush()


def setup(n=60, m=255, k=3, seed=7, ntraj=4):
    p = O.fastsat(n, m, k, seed)
    rng = np.random.default_rng(seed)
    pubs = np.stack([np.concatenate([rng.uniform(-1, 1, n), rng.uniform(1, 2, m)]) for _ in range(ntraj)])
    return p, pubs.astype(np.float32)


def test_adjoint_matches_oracle_and_replay():
    say("adjoint N=60 T=40 C=8, 4 trajectories")
    p, pubs = setup()
    T = 40
    adj = A.HipAdjoint(A.SatProblem(p.literals, p.lengths), eps=1e-6, steps=T, checkpoint_every=8)
    V = adj.forward(pubs)
    h = adj.steps()
    g = adj.backward()
    with O.pow_mode("cr"):
        for t in range(len(pubs)):
            rk = O.RK(p, eps=1e-6, mode="mix")
            Vo, go, ho = rk.adjoint(pubs[t], T)
            assert np.array_equal(h[t], ho), t
            assert V[t] == pytest.approx(Vo, rel=1e-13)
            scale = np.abs(go).max()
            assert np.abs(g[t] - go).max() <= 1e-12 * scale, (t, np.abs(g[t] - go).max() / scale)
    Vr, xT = adj.replay(pubs, h)
    assert np.array_equal(Vr, V)
    for t in range(len(pubs)):
        rk = O.RK(p, eps=1e-6, mode="mix")
        rk.set_state(pubs[t])
        with O.pow_mode("cr"):
            rk.step(T)
        assert np.array_equal(xT[t], rk.get_state())


def test_checkpoint_interval_invariance():
    say("checkpoint interval invariance")
    p, pubs = setup(seed=3, ntraj=2)
    grads, Vs = [], []
    for C in (1, 7, 30, 64):      # 7: a short last segment; 64 > T: a single one
        adj = A.HipAdjoint(A.SatProblem(p.literals, p.lengths), eps=1e-6, steps=30, checkpoint_every=C)
        Vs.append(adj.forward(pubs))
        grads.append(adj.backward())
    for i in range(1, len(grads)):
        assert np.array_equal(Vs[0], Vs[i]), i
        assert np.array_equal(grads[0], grads[i]), i


def test_gradient_vs_fd_of_fp64_frozen_map():
    say("FD of the FP64 frozen map")
    n, m, T = 50, 212, 30
    p, pubs = setup(n, m, 3, seed=11, ntraj=1)
    adj = A.HipAdjoint(A.SatProblem(p.literals, p.lengths), eps=1e-6, steps=T, checkpoint_every=16)
    adj.forward(pubs)
    h = adj.steps()[0]
    g = adj.backward()[0]
    rk = O.RK(p, eps=1e-6, mode="f64")
    x0 = pubs[0].astype(np.float64)
    rng = np.random.default_rng(0)
    errs = []
    for i in rng.choice(n + m, 32, replace=False):
        e = np.zeros(n + m)
        e[i] = 1e-4
        fd = (rk.frozen_energy(x0 + e, h) - rk.frozen_energy(x0 - e, h)) / 2e-4
        errs.append(abs(g[i] - fd))
    scale = np.abs(g).max()
    say("max |grad - FD| / max|grad| = %.2e" % (max(errs) / scale))
    assert max(errs) <= 1e-3 * scale


def test_fp64_replay_matches_oracle_frozen_map():
    """The FP64 frozen-step replay (the reference arithmetic in double)
    equals the oracle's FP64 frozen map: V(T) to rounding of the final sum."""
    say("FP64 frozen replay vs oracle")
    n, m, T = 50, 212, 30
    p, pubs = setup(n, m, 3, seed=13, ntraj=3)
    adj = A.HipAdjoint(A.SatProblem(p.literals, p.lengths), eps=1e-6, steps=T, checkpoint_every=8)
    adj.forward(pubs)
    h = adj.steps()
    x64 = pubs.astype(np.float64)
    x64[1, 3] += 1e-6
    V = adj.replay_f64(x64, h)
    for t in range(3):
        rk = O.RK(p, eps=1e-6, mode="f64")
        want = rk.frozen_energy(x64[t], h[t])
        assert abs(V[t] - want) <= 1e-12 * abs(want), (t, V[t], want)


def test_torch_autograd_function():
    torch = pytest.importorskip("torch")
    say("torch.autograd.Function")
    n, m = 40, 170
    p, pubs = setup(n, m, 3, seed=5, ntraj=3)
    adj = A.HipAdjoint(A.SatProblem(p.literals, p.lengths), eps=1e-6, steps=20, checkpoint_every=4)
    s0 = torch.tensor(pubs[:, :n], dtype=torch.float32, requires_grad=True)
    a0 = torch.tensor(pubs[:, n:], dtype=torch.float32, requires_grad=True)
    f = A.energy_function()
    V = f.apply(s0, a0, adj)
    (V * torch.tensor([1.0, 2.0, -0.5], dtype=torch.float64)).sum().backward()
    adj2 = A.HipAdjoint(A.SatProblem(p.literals, p.lengths), eps=1e-6, steps=20, checkpoint_every=4)
    adj2.forward(pubs)
    g = adj2.backward() * np.array([1.0, 2.0, -0.5])[:, None]
    np.testing.assert_allclose(s0.grad.numpy(), g[:, :n].astype(np.float32), rtol=1e-6)
    np.testing.assert_allclose(a0.grad.numpy(), g[:, n:].astype(np.float32), rtol=1e-6)


def test_fp32_with_fp64_steps_within_1e5_of_fp64_oracle(golden_dir):
    """FP32 state error alone: the C1 instance and start, the FP64 oracle's
    accepted step sizes replayed with FP32 state (asat_adj_replay).  Without
    the divergence of the step sequences that chaotic trajectories bring, the
    FP32 state stays within 1e-5 relative of the FP64 oracle after 10, 100
    and 1000 steps."""
    say("FP32 frozen-step replay vs FP64 oracle, 1000 steps")
    g = np.load(os.path.join(golden_dir, "c1_traj.npz"), allow_pickle=False)
    p = O.fastsat(50, 213, 3, 0)
    pub = g["pub0"].astype(np.float64)
    for T in (10, 100, 1000):
        with O.pow_mode("cr"):
            orc = O.RK(p, eps=1e-6, mode="f64")
            orc.set_state(pub)
            hs = []
            for _ in range(T):
                orc.step(1)
                hs.append(orc.hdid)
            x64 = orc.get_state().astype(np.float64)
        adj = A.HipAdjoint(A.SatProblem(p.literals, p.lengths), eps=1e-6, steps=T, checkpoint_every=32)
        _, xT = adj.replay(pub.astype(np.float32)[None], np.array(hs)[None])
        rel = np.abs(xT[0].astype(np.float64) - x64).max() / np.abs(x64).max()
        say("  T = %d: max rel %.3e" % (T, rel))
        assert rel <= 1e-5, (T, rel)
