"""CPU oracle for the AnalogSat integrator -- test infrastructure only.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` import this package, and only to check results.  It wraps two C
libraries built by ``oracle/Makefile``:

* ``libasat_oracle.so`` -- the C restatement (``asat_oracle.c``) of the
  reference CPU path: ``CpuSat`` (AnalogSAT/cpu_sat/CpuSat.cpp:64-241),
  ``CpuRungeKutta`` (AnalogSAT/cpu_sat/CpuRungeKutta.cpp:65-267),
  ``SatSolver::DoSolve`` (AnalogSAT/solver/SatSolver.h:273-355), the problem
  container (AnalogSAT/problem/SatProblem.cpp:64-95), ``FastSatMaker``
  (AnalogSAT/problem/FastSatMaker.cpp:59-107) and ``CpuRandom``
  (AnalogSAT/cpu_sat/CpuRandom.cpp:25-52).  Two precision modes: ``f64``
  (the reference's double instantiation) and ``mix`` (float state, double step
  control -- the GPU FP32 mode, which has no reference counterpart).
* ``_ref/libasat_ref.so`` -- the reference's own CPU sources compiled in place
  (optional; present only where it was built).

Pinning status of each function is listed in the module docstrings of
``tests/test_oracle_golden.py`` and ``tests/test_oracle_math.py``.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libasat_oracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libasat_ref.so")

_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_ldp = np.ctypeslib.ndpointer(np.longdouble, flags="C_CONTIGUOUS")
_vp = ctypes.c_void_p

_lib = None
_ref = None

RESULT_NAMES = {0: "solution found", 1: "underflow", 2: "user interrupt",
                3: "maximum time reached", 4: "maximum steps reached", 5: "unknown result"}


def build():
    """Compile the restatement (and the reference checker where its sources exist)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    if os.path.isdir("/root/reference/AnalogSAT"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        for sfx, fp, ftype in (("_f64", _f64p, ctypes.c_double), ("_mix", _f32p, ctypes.c_float),
                               ("_ld", _ldp, ctypes.c_longdouble)):
            getattr(L, "orc_rhs" + sfx).argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _i32p, ftype, fp, fp]
            getattr(L, "orc_rhs_tanh" + sfx).argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _i32p, ftype, ftype,
                                                        fp, fp]
            getattr(L, "orc_rk_set_family" + sfx).argtypes = [_vp, ctypes.c_int, ctypes.c_double]
            getattr(L, "orc_violations" + sfx).argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _i32p, fp]
            getattr(L, "orc_violations" + sfx).restype = ctypes.c_int
            ctype = ctypes.c_longdouble if sfx == "_ld" else ctypes.c_double   # the control type CT
            f = getattr(L, "orc_rk_create" + sfx)
            f.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _i32p, ctype]
            f.restype = _vp
            getattr(L, "orc_rk_destroy" + sfx).argtypes = [_vp]
            getattr(L, "orc_rk_reset" + sfx).argtypes = [_vp]
            getattr(L, "orc_rk_set_b" + sfx).argtypes = [_vp, ctypes.c_double]
            getattr(L, "orc_rk_set_floor" + sfx).argtypes = [_vp, ctypes.c_double]
            getattr(L, "orc_rk_set_state" + sfx).argtypes = [_vp, fp]
            getattr(L, "orc_rk_get_state" + sfx).argtypes = [_vp, fp]
            getattr(L, "orc_rk_step" + sfx).argtypes = [_vp, ctypes.c_int]
            getattr(L, "orc_rk_step" + sfx).restype = ctypes.c_int
            getattr(L, "orc_rk_violations" + sfx).argtypes = [_vp]
            getattr(L, "orc_rk_violations" + sfx).restype = ctypes.c_int
            for nm in ("time", "hdid", "hnext", "errmax"):
                g = getattr(L, "orc_rk_" + nm + sfx)
                g.argtypes = [_vp]
                g.restype = ctypes.c_double
            for nm in ("steps", "attempts", "rhscount"):
                g = getattr(L, "orc_rk_" + nm + sfx)
                g.argtypes = [_vp]
                g.restype = ctypes.c_long
            getattr(L, "orc_rk_err_vector" + sfx).argtypes = [_vp, _f64p]
            getattr(L, "orc_vjp" + sfx).argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _i32p, fp, _f64p, _f64p]
            getattr(L, "orc_energy_grad" + sfx).argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _i32p, fp, _f64p]
            getattr(L, "orc_energy_grad" + sfx).restype = ctypes.c_double
            getattr(L, "orc_frozen_energy" + sfx).argtypes = [_vp, fp, _f64p, ctypes.c_int]
            getattr(L, "orc_frozen_energy" + sfx).restype = ctypes.c_double
            getattr(L, "orc_adjoint" + sfx).argtypes = [_vp, fp, ctypes.c_int, _f64p, _f64p]
            getattr(L, "orc_adjoint" + sfx).restype = ctypes.c_double
            getattr(L, "orc_adjoint_frozen" + sfx).argtypes = [_vp, fp, _f64p, ctypes.c_int, _f64p]
            getattr(L, "orc_adjoint_frozen" + sfx).restype = ctypes.c_double
            f = getattr(L, "orc_solve" + sfx)
            f.argtypes = [_vp, ctype, ctypes.c_int, ctypes.c_int, _i32p, ctypes.c_int,
                          ctypes.POINTER(ctypes.c_int)]
            f.restype = ctypes.c_int
        L.orc_problem_build.argtypes = [_i32p, _i32p, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                        ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), _vp]
        L.orc_uniform_f64.argtypes = [ctypes.c_uint, _f64p, ctypes.c_int]
        L.orc_random_state_f64.argtypes = [ctypes.c_uint, ctypes.c_int, ctypes.c_int, _f64p]
        L.orc_fastsat.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_ulong, _i32p]
        L.orc_randomsat.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_ulong, _i32p]
        L.orc_randomsat.restype = ctypes.c_int
        _u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
        L.orc_philox4x32_10.argtypes = [_u32p, _u32p, _u32p]
        L.orc_philox_uniform.argtypes = [ctypes.c_ulonglong, ctypes.c_ulonglong, ctypes.c_int, ctypes.c_int, _vp]
        L.orc_set_pow_mode.argtypes = [ctypes.c_int]
        L.orc_set_mean_mode.argtypes = [ctypes.c_int]
        L.orc_get_mean_mode.restype = ctypes.c_int
        L.orc_get_pow_mode.restype = ctypes.c_int
        L.orc_pow_cr.argtypes = [ctypes.c_double, ctypes.c_int]
        L.orc_pow_cr.restype = ctypes.c_double
        _lib = L
    return _lib


def ref_available():
    return os.path.exists(REF_PATH)


def ref():
    """The reference's own CPU path (oracle/_ref), or None where it was not built."""
    global _ref
    if _ref is None and ref_available():
        R = ctypes.CDLL(REF_PATH)
        R.ref_problem_sizes.argtypes = [_i32p, _i32p, ctypes.c_int] + [ctypes.POINTER(ctypes.c_int)] * 3
        R.ref_rhs_f64.argtypes = [_i32p, _i32p, ctypes.c_int, ctypes.c_double, _f64p, _f64p]
        R.ref_rhs_f32.argtypes = [_i32p, _i32p, ctypes.c_int, ctypes.c_double, _f32p, _f32p]
        R.ref_rhs_tanh_f64.argtypes = [_i32p, _i32p, ctypes.c_int, ctypes.c_double, ctypes.c_double, _f64p, _f64p]
        R.ref_rhs_tanh_f32.argtypes = [_i32p, _i32p, ctypes.c_int, ctypes.c_double, ctypes.c_double, _f32p, _f32p]
        R.ref_violations_f64.argtypes = [_i32p, _i32p, ctypes.c_int, _f64p]
        R.ref_traj_ld.argtypes = [_i32p, _i32p, ctypes.c_int, ctypes.c_double, _ldp, ctypes.c_int, _ldp, _ldp]
        R.ref_traj_ld.restype = ctypes.c_int
        R.ref_violations_f64.restype = ctypes.c_int
        R.ref_rk_create.argtypes = [_i32p, _i32p, ctypes.c_int, ctypes.c_double]
        R.ref_rk_create.restype = _vp
        R.ref_rk_destroy.argtypes = [_vp]
        R.ref_rk_set_state.argtypes = [_vp, _f64p]
        R.ref_rk_get_state.argtypes = [_vp, _f64p]
        R.ref_rk_step.argtypes = [_vp, ctypes.c_int]
        R.ref_rk_step.restype = ctypes.c_int
        for nm in ("time", "hdid"):
            getattr(R, "ref_rk_" + nm).argtypes = [_vp]
            getattr(R, "ref_rk_" + nm).restype = ctypes.c_double
        for nm in ("steps", "attempts", "rhscount"):
            getattr(R, "ref_rk_" + nm).argtypes = [_vp]
            getattr(R, "ref_rk_" + nm).restype = ctypes.c_long
        R.ref_rk_violations.argtypes = [_vp]
        R.ref_rk_violations.restype = ctypes.c_int
        R.ref_solve.argtypes = [_i32p, _i32p, ctypes.c_int, _f64p, ctypes.c_double, ctypes.c_int,
                                ctypes.c_int, ctypes.c_double, _i32p, ctypes.c_int,
                                ctypes.POINTER(ctypes.c_int), _f64p,
                                np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS"),
                                ctypes.POINTER(ctypes.c_double)]
        R.ref_solve.restype = ctypes.c_int
        R.ref_fastsat.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_ulong, _i32p]
        R.ref_fastsat.restype = ctypes.c_int
        R.ref_randomsat.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_ulong, _i32p]
        R.ref_randomsat.restype = ctypes.c_int
        R.ref_uniform_f64.argtypes = [ctypes.c_uint, _f64p, ctypes.c_int]
        _ref = R
    return _ref


# ----------------------------------------------------------------------------
# problem helpers
# ----------------------------------------------------------------------------
class Problem:
    """Padded clause-major matrix as the reference's CpuSat stores it.

    ``literals``/``lengths`` are the flat DIMACS-style clause list
    (SatProblem::AddClauses input); ``C`` is the m*k matrix with zero padding.
    """

    def __init__(self, literals, lengths):
        self.literals = np.ascontiguousarray(literals, dtype=np.int32)
        self.lengths = np.ascontiguousarray(lengths, dtype=np.int32)
        n, m, k = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        L = lib()
        L.orc_problem_build(self.literals, self.lengths, len(self.lengths), ctypes.byref(n),
                            ctypes.byref(m), ctypes.byref(k), None)
        self.n, self.m, self.k = n.value, m.value, k.value
        self.C = np.zeros(max(self.m * self.k, 1), dtype=np.int32)
        L.orc_problem_build(self.literals, self.lengths, len(self.lengths), ctypes.byref(n),
                            ctypes.byref(m), ctypes.byref(k), self.C.ctypes.data)

    @classmethod
    def from_clauses(cls, clauses):
        lits = [l for c in clauses for l in c]
        return cls(np.array(lits, dtype=np.int32), np.array([len(c) for c in clauses], dtype=np.int32))

    @property
    def nclauses(self):
        return len(self.lengths)


def fastsat(n, m, k, seed):
    """FastSatMaker(seed) instance as a Problem (restated generator)."""
    out = np.zeros(m * k, dtype=np.int32)
    lib().orc_fastsat(n, m, k, seed, out)
    return Problem(out, np.full(m, k, dtype=np.int32))


def randomsat(n, m, k, seed):
    """RandomSatMaker(seed) instance as a Problem (restated generator: the
    FastSatMaker draw with exact-duplicate clauses drawn again)."""
    out = np.zeros(m * k, dtype=np.int32)
    if lib().orc_randomsat(n, m, k, seed, out) != 0:
        raise ValueError("more clauses than distinct %d-clauses over %d variables" % (k, n))
    return Problem(out, np.full(m, k, dtype=np.int32))


def philox4x32_10(ctr, key):
    """One Philox4x32-10 block: four 32-bit counter words, two key words."""
    out = np.zeros(4, np.uint32)
    lib().orc_philox4x32_10(np.ascontiguousarray(ctr, np.uint32), np.ascontiguousarray(key, np.uint32), out)
    return out


def philox_uniform(seed, offset, length, dtype=np.float64):
    """The library's ISatRandom stream: `length` uniforms at stream offset `offset`."""
    out = np.zeros(length, dtype=dtype)
    lib().orc_philox_uniform(seed, offset, length, 1 if dtype == np.float32 else 0, out.ctypes.data)
    return out


def uniform_f64(seed, length):
    out = np.zeros(length, dtype=np.float64)
    lib().orc_uniform_f64(seed, out, length)
    return out


def random_state_f64(seed, n, m):
    out = np.zeros(n + m, dtype=np.float64)
    lib().orc_random_state_f64(seed, n, m, out)
    return out


class pow_mode:
    """Context manager: 'glibc' (reference behaviour, default) or 'cr'
    (correctly rounded controller pow, the GPU's definition)."""

    def __init__(self, mode):
        self.mode = {"glibc": 0, "cr": 1}[mode]

    def __enter__(self):
        self.prev = lib().orc_get_pow_mode()
        lib().orc_set_pow_mode(self.mode)
        return self

    def __exit__(self, *a):
        lib().orc_set_pow_mode(self.prev)


class mean_mode:
    """Context manager for the bias term's sum of a: 'serial' (reference
    behaviour, default) or 'double' (accumulated in double and rounded once,
    the GPU's definition)."""

    def __init__(self, mode):
        self.mode = {"serial": 0, "double": 1}[mode]

    def __enter__(self):
        self.prev = lib().orc_get_mean_mode()
        lib().orc_set_mean_mode(self.mode)
        return self

    def __exit__(self, *a):
        lib().orc_set_mean_mode(self.prev)


def pow_cr(x, q):
    return lib().orc_pow_cr(float(x), int(q))


def _dtype(mode):
    return {"f64": np.float64, "mix": np.float32, "ld": np.longdouble}[mode]


def rhs(problem, x, mode="f64", b=0.0, family="ctds", q=1.1):
    """One RHS on the internal layout x = [-1, s, a] (length n+m+1);
    family 'ctds' (CpuSat) or 'tanh' (CpuSatTanh with parameter q)."""
    dt = _dtype(mode)
    x = np.ascontiguousarray(x, dtype=dt)
    out = np.zeros_like(x)
    if family == "tanh":
        getattr(lib(), "orc_rhs_tanh_" + mode)(problem.n, problem.m, problem.k, problem.C, b, q, x, out)
    else:
        getattr(lib(), "orc_rhs_" + mode)(problem.n, problem.m, problem.k, problem.C, b, x, out)
    return out


def vjp(problem, Y, w, mode="f64"):
    """u = J(Y)^T w on the internal layout (Y in the state type, w/u double)."""
    dt = np.float64 if mode == "f64" else np.float32
    Y = np.ascontiguousarray(Y, dtype=dt)
    w = np.ascontiguousarray(w, dtype=np.float64)
    u = np.zeros(len(Y))
    getattr(lib(), "orc_vjp_" + mode)(problem.n, problem.m, problem.k, problem.C, Y, w, u)
    return u


def energy_grad(problem, x, mode="f64"):
    """(V, grad V) at the internal-layout state x; V = sum_m a_m K_m^2."""
    dt = np.float64 if mode == "f64" else np.float32
    x = np.ascontiguousarray(x, dtype=dt)
    g = np.zeros(len(x))
    V = getattr(lib(), "orc_energy_grad_" + mode)(problem.n, problem.m, problem.k, problem.C, x, g)
    return V, g


def violations(problem, pub, mode="f64"):
    dt = _dtype(mode)
    x = np.concatenate([[-1.0], np.asarray(pub, dtype=np.float64)]).astype(dt)
    return getattr(lib(), "orc_violations_" + mode)(problem.n, problem.m, problem.k, problem.C, x)


class RK:
    """Restated CpuRungeKutta bound to one trajectory: mode 'f64' (the
    reference's <double>), 'mix' (FP32 state, FP64 control: the GPU's FP32
    definition) or 'ld' (the reference's <long double>)."""

    def __init__(self, problem, eps=1e-4, mode="f64", b=0.0, floor=None, family="ctds", q=1.1):
        self.problem = problem
        self.mode = mode
        self.dt = _dtype(mode)
        self._L = lib()
        self._sfx = "_" + mode
        self._h = getattr(self._L, "orc_rk_create" + self._sfx)(problem.n, problem.m, problem.k, problem.C, eps)
        if b:
            self._call("orc_rk_set_b", b)
        if floor is not None:
            self._call("orc_rk_set_floor", floor)
        if family == "tanh":
            self._call("orc_rk_set_family", 1, q)

    def _call(self, name, *args):
        return getattr(self._L, name + self._sfx)(self._h, *args)

    def __del__(self):
        if getattr(self, "_h", None):
            getattr(self._L, "orc_rk_destroy" + self._sfx)(self._h)
            self._h = None

    def set_state(self, pub):
        self._call("orc_rk_set_state", np.ascontiguousarray(pub, dtype=self.dt))

    def get_state(self):
        out = np.zeros(self.problem.n + self.problem.m, dtype=self.dt)
        self._call("orc_rk_get_state", out)
        return out

    def step(self, count=1):
        return bool(self._call("orc_rk_step", count))

    def reset(self):
        self._call("orc_rk_reset")

    def violations(self):
        return self._call("orc_rk_violations")

    time = property(lambda s: s._call("orc_rk_time"))
    hdid = property(lambda s: s._call("orc_rk_hdid"))
    hnext = property(lambda s: s._call("orc_rk_hnext"))
    errmax = property(lambda s: s._call("orc_rk_errmax"))
    steps = property(lambda s: s._call("orc_rk_steps"))
    attempts = property(lambda s: s._call("orc_rk_attempts"))
    rhscount = property(lambda s: s._call("orc_rk_rhscount"))

    def err_vector(self):
        """|xerr/xscal| per element of the last attempt, internal layout [0, s, a]."""
        out = np.zeros(self.problem.n + self.problem.m + 1)
        self._call("orc_rk_err_vector", out)
        return out

    def adjoint(self, pub, T):
        """Adaptive forward of T accepted steps from pub, then the discrete
        adjoint: returns (V(T), dV/d pub (double), h sequence)."""
        pub = np.ascontiguousarray(pub, dtype=self.dt)
        h = np.zeros(T)
        g = np.zeros(self.problem.n + self.problem.m)
        V = self._call("orc_adjoint", pub, T, h, g)
        return V, g, h

    def adjoint_frozen(self, pub, h):
        """(V(T), dV/d pub) through len(h) Cash-Karp steps with step sizes h."""
        pub = np.ascontiguousarray(pub, dtype=self.dt)
        h = np.ascontiguousarray(h, dtype=np.float64)
        g = np.zeros(self.problem.n + self.problem.m)
        V = self._call("orc_adjoint_frozen", pub, h, len(h), g)
        return V, g

    def frozen_energy(self, pub, h):
        """V after len(h) Cash-Karp steps with the given step sizes (no control)."""
        pub = np.ascontiguousarray(pub, dtype=self.dt)
        h = np.ascontiguousarray(h, dtype=np.float64)
        return self._call("orc_frozen_energy", pub, h, len(h))

    def solve(self, max_time=100.0, max_steps=100000, batch=20, max_batches=100000):
        viol = np.zeros(max_batches, dtype=np.int32)
        nb = ctypes.c_int()
        res = self._call("orc_solve", max_time, max_steps, batch, viol, max_batches, ctypes.byref(nb))
        return res, viol[: min(nb.value, max_batches)]
