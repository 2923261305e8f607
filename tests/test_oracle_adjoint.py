"""Pin the oracle's adjoint (CPU only).

The reference has no adjoint; the oracle's is pinned by mathematics:
  * the RHS VJP equals finite-difference Jacobian-vector products
    (J^T w . v == w . J v for random v, w)
  * grad V equals finite differences of V, and -grad_s V equals the spin RHS
  * the discrete adjoint of T frozen Cash-Karp steps equals central finite
    differences of V(T) with the same frozen step sequence (h = 1e-4 relative
    coordinate perturbation, FP64), on random coordinates
"""
import numpy as np
import pytest

import oracle as O


def internal(p, rng):
    return np.concatenate([[-1.0], rng.uniform(-1, 1, p.n), rng.uniform(1, 2, p.m)])


@pytest.mark.parametrize("k", [3, 4])
def test_vjp_matches_fd_jvp(k):
    p = O.fastsat(30, 100 if k == 3 else 250, k, 4)
    rng = np.random.default_rng(k)
    Y = internal(p, rng)
    D = len(Y)
    w = rng.normal(size=D); w[0] = 0
    v = rng.normal(size=D); v[0] = 0
    u = O.vjp(p, Y, w)
    eps = 1e-6
    jv = (O.rhs(p, Y + eps * v) - O.rhs(p, Y - eps * v)) / (2 * eps)
    jv[0] = 0
    assert np.dot(u, v) == pytest.approx(np.dot(w, jv), rel=1e-7)


def test_vjp_padding_and_mixed_lengths():
    p = O.Problem.from_clauses([[1, -2], [2, 3, -4], [-1, 4, 5], [3]])
    rng = np.random.default_rng(9)
    Y = internal(p, rng)
    w = rng.normal(size=len(Y)); w[0] = 0
    u = O.vjp(p, Y, w)
    # dense Jacobian by central differences
    J = np.zeros((len(Y), len(Y)))
    for i in range(1, len(Y)):
        e = np.zeros(len(Y)); e[i] = 1e-6
        J[:, i] = (O.rhs(p, Y + e) - O.rhs(p, Y - e)) / 2e-6
    J[0, :] = 0
    np.testing.assert_allclose(u[1:], (J.T @ w)[1:], rtol=1e-6, atol=1e-9)


def test_energy_gradient():
    p = O.fastsat(25, 106, 3, 2)
    rng = np.random.default_rng(1)
    x = internal(p, rng)
    V, g = O.energy_grad(p, x)
    d = O.rhs(p, x)
    np.testing.assert_allclose(g[1:p.n + 1], -d[1:p.n + 1], rtol=1e-12, atol=1e-15)
    for i in rng.choice(np.arange(1, len(x)), 12, replace=False):
        e = np.zeros(len(x)); e[i] = 1e-6
        fd = (O.energy_grad(p, x + e)[0] - O.energy_grad(p, x - e)[0]) / 2e-6
        assert g[i] == pytest.approx(fd, rel=1e-6, abs=1e-9 * max(1.0, abs(V)))


def test_discrete_adjoint_matches_frozen_fd():
    n, m, T = 60, 255, 40
    p = O.fastsat(n, m, 3, 7)
    rng = np.random.default_rng(5)
    pub = np.concatenate([rng.uniform(-1, 1, n), rng.uniform(1, 2, m)])
    rk = O.RK(p, eps=1e-6)
    V, g, h = rk.adjoint(pub, T)
    assert np.isfinite(V) and np.all(h > 0)
    assert rk.frozen_energy(pub, h) == pytest.approx(V, rel=1e-12)
    for i in rng.choice(n + m, 24, replace=False):
        step = 1e-4 * max(1.0, abs(pub[i]))
        e = np.zeros(n + m); e[i] = step
        fd = (rk.frozen_energy(pub + e, h) - rk.frozen_energy(pub - e, h)) / (2 * step)
        assert g[i] == pytest.approx(fd, rel=2e-4, abs=1e-7 * max(1.0, abs(V))), i


def test_adaptive_forward_equals_frozen_replay():
    n, m, T = 40, 170, 30
    p = O.fastsat(n, m, 3, 11)
    rng = np.random.default_rng(4)
    pub = np.concatenate([rng.uniform(-1, 1, n), np.ones(m)])
    for mode, dt in (("f64", np.float64), ("mix", np.float32)):
        rk = O.RK(p, eps=1e-6, mode=mode)
        V, g, h = rk.adjoint(pub.astype(dt), T)
        V2, g2 = rk.adjoint_frozen(pub.astype(dt), h)
        assert V == V2 and np.array_equal(g, g2)
        # the replay reproduces the adaptive run's final state exactly
        rk.set_state(pub.astype(dt)); rk.reset(); rk.step(T)
        x = np.concatenate([[-1.0], rk.get_state()]).astype(dt)
        assert O.energy_grad(p, x, mode=mode)[0] == V


def test_mixed_adjoint_close_to_f64_on_same_steps():
    """FP32 forward / FP64 adjoint vs all-FP64, same frozen step sequence."""
    n, m, T = 50, 212, 40
    p = O.fastsat(n, m, 3, 3)
    rng = np.random.default_rng(2)
    pub = np.concatenate([rng.uniform(-1, 1, n), np.ones(m)]).astype(np.float32)
    _, _, h = O.RK(p, eps=1e-6, mode="mix").adjoint(pub, T)
    V64, g64 = O.RK(p, eps=1e-6).adjoint_frozen(pub.astype(np.float64), h)
    Vmx, gmx = O.RK(p, eps=1e-6, mode="mix").adjoint_frozen(pub, h)
    assert Vmx == pytest.approx(V64, rel=1e-5)
    assert np.abs(gmx - g64).max() <= 1e-4 * np.abs(g64).max()
