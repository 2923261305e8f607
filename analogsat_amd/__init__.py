"""AnalogSat CTDS integrator for AMD Instinct MI355X (gfx950).

Python mirror of the reference's plugin interface over the C ABI in
``include/analogsat_hip.h`` (library ``analogsat_amd/libanalogsat_hip.so``):

=====================  =====================================================
this module            reference (AnalogSAT/...)
=====================  =====================================================
``HipSat``             ``ISat<TFloat,TState>`` / ``CudaSat3`` (solver/ISat.h:31-87)
``HipState``           ``ISatState`` / ``CudaSatState2`` (solver/ISatState.h:31-161)
``HipRungeKutta``      ``IODEInt`` / ``CudaRungeKutta`` (solver/IODEInt.h:28-60)
``HipRandom``          ``ISatRandom`` / ``CudaRandom`` (solver/ISatRandom.h:25-33)
``SatSolver``          ``SatSolver<TFloat,TState>`` (solver/SatSolver.h:41-356)
``SatProblem``         ``SatProblem`` (problem/SatProblem.h:32-73)
``FastSatMaker``       ``FastSatMaker`` (problem/FastSatMaker.h:29-61)
``SatResult``          ``SatResult`` (solver/SatResult.h:24-32)
=====================  =====================================================

Methods keep the reference's names (``SetProblem``, ``GetDerivatives``,
``Step``, ``Solve`` ...) and raise the reference's exception types
(``ValueError`` for ``std::invalid_argument``, ``RuntimeError`` for
``std::runtime_error``).  The library is required: importing works without
it, but every object constructor raises if ``libanalogsat_hip.so`` is missing.
"""
import ctypes
import enum
import os
import time as _time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libanalogsat_hip.so")

ASAT_OK, ASAT_ERR_INVALID_ARGUMENT, ASAT_ERR_NOT_CONFIGURED, ASAT_ERR_HIP, ASAT_UNDERFLOW, ASAT_ERR_UNSUPPORTED = range(6)
ASAT_F64, ASAT_F32 = 0, 1

_lib = None


class SatResult(enum.IntEnum):
    """AnalogSAT/solver/SatResult.h:24-32."""
    SAT_SOLUTION_FOUND = 0
    SAT_UNDERFLOW = 1
    SAT_ODE_INTERRUPTED = 2
    SAT_MAXTIME_REACHED = 3
    SAT_MAXITER_REACHED = 4
    SAT_UNKNOWN = 5


_RESULT_MSG = {0: "solution found", 1: "underflow", 2: "user interrupt", 3: "maximum time reached",
               4: "maximum steps reached", 5: "unknown result"}


def GetSatResultMessage(result):
    """AnalogSAT/solver/SatResult.cpp:20-32."""
    return _RESULT_MSG.get(int(result), "unknown result")


class BatchResult(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("steps", ctypes.c_longlong), ("attempts", ctypes.c_longlong),
                ("rhscount", ctypes.c_longlong), ("time", ctypes.c_double), ("hdid", ctypes.c_double),
                ("hnext", ctypes.c_double), ("min_violations", ctypes.c_int), ("violations", ctypes.c_int),
                ("sat_step", ctypes.c_int), ("sat_time", ctypes.c_double), ("solved_wall_ms", ctypes.c_double)]


class BatchBucket(ctypes.Structure):
    _fields_ = [("trajectories", ctypes.c_int), ("max_n", ctypes.c_int), ("max_m", ctypes.c_int),
                ("max_slots", ctypes.c_int), ("block", ctypes.c_int), ("lds", ctypes.c_int),
                ("grid", ctypes.c_int), ("lds_bytes", ctypes.c_longlong),
                ("tuned_us_per_attempt", ctypes.c_double)]


class RkStats(ctypes.Structure):
    _fields_ = [("time", ctypes.c_double), ("hdid", ctypes.c_double), ("hnext", ctypes.c_double),
                ("errmax", ctypes.c_double), ("steps", ctypes.c_longlong), ("attempts", ctypes.c_longlong),
                ("rhscount", ctypes.c_longlong)]


class SolveResult(ctypes.Structure):
    """asat_solve_result_t: outcome of SatSolver's loop run by asat_rk_solve."""
    _fields_ = [("result", ctypes.c_int), ("last_violations", ctypes.c_int), ("steps", ctypes.c_longlong),
                ("on_device", ctypes.c_int), ("pad_", ctypes.c_int)]


class LayoutInfo(ctypes.Structure):
    """asat_layout_t: the passes and launch shapes SetProblem chose."""
    _fields_ = [(nm, ctypes.c_int) for nm in ("clause_pass", "var_pass", "groups", "group_vars", "tiles",
                                              "clause_grid", "var_grid", "var_lds_bytes", "slot_spins",
                                              "direct_primary", "tile_clauses", "run_tables",
                                              "primary_words")]

    def as_dict(self):
        return {nm: getattr(self, nm) for nm, _ in self._fields_}


def build():
    from . import builder
    return builder.build()


def _mapped(stem):
    """Real paths of the mapped shared objects whose file name starts with stem."""
    paths = set()
    try:
        with open("/proc/self/maps") as f:
            for line in f:
                fields = line.split(None, 5)   # address perms offset dev inode [path]
                if len(fields) == 6 and "/" + stem in fields[5]:
                    paths.add(fields[5].strip().removesuffix(" (deleted)"))
    except OSError:
        return set()
    return {os.path.realpath(p) for p in paths if os.path.basename(p).startswith(stem)}


def _share_hip_runtime():
    """Keep one HIP runtime in the process.

    PyTorch-ROCm wheels carry their own libamdhip64 / libhsa-runtime64 and
    their libraries ask for them as ``libamdhip64.so`` / ``libhsa-runtime64.so``
    with an RPATH of their own directory.  The dynamic loader reuses an object
    only under a name it has already registered, so without this a later
    ``import torch`` maps a second HIP and HSA runtime next to the one this
    library was linked against, and the second can then report no GPU.
    Registering the unversioned names for the runtime already mapped (only
    when the name resolves to that same file) makes torch reuse it.  If torch
    was imported first, this library has already bound to torch's runtime and
    nothing is done."""
    for stem in ("libamdhip64", "libhsa-runtime64"):
        have = _mapped(stem + ".so")
        if len(have) != 1:
            continue
        target = next(iter(have))
        name = stem + ".so"
        if os.path.realpath(os.path.join(os.path.dirname(target), name)) != target:
            continue
        dirs = [d for d in os.environ.get("LD_LIBRARY_PATH", "").split(":") if d]
        first = next((os.path.join(d, name) for d in dirs if os.path.exists(os.path.join(d, name))), None)
        if first is not None and os.path.realpath(first) != target:
            continue
        try:
            ctypes.CDLL(name, mode=ctypes.RTLD_GLOBAL)
        except OSError:
            continue


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libanalogsat_hip.so is not built: run python -m analogsat_amd.builder")
        L = ctypes.CDLL(LIB_PATH)
        _share_hip_runtime()
        vp, ip, i = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.c_int
        pvp = ctypes.POINTER(ctypes.c_void_p)
        sig = {
            "asat_last_error": ([], ctypes.c_char_p),
            "asat_version": ([], ctypes.c_char_p),
            "asat_device_count": ([], i),
            "asat_problem_fastsat": ([i, i, i, ctypes.c_ulong, vp], i),
            "asat_problem_seed_engine": ([ctypes.c_ulong, ctypes.POINTER(ctypes.c_ulonglong)], i),
            "asat_problem_fastsat_stream": ([i, i, i, ctypes.POINTER(ctypes.c_ulonglong), vp], i),
            "asat_problem_randomsat": ([i, i, i, ctypes.c_ulong, vp], i),
            "asat_problem_randomsat_stream": ([i, i, i, ctypes.POINTER(ctypes.c_ulonglong), vp], i),
            "asat_sys_create": ([i, pvp], i),
            "asat_sys_destroy": ([vp], i),
            "asat_sys_set_problem": ([vp, vp, vp, i], i),
            "asat_sys_get_nmk": ([vp, ip, ip, ip], i),
            "asat_sys_set_b": ([vp, ctypes.c_double], i),
            "asat_sys_get_b": ([vp, ctypes.POINTER(ctypes.c_double)], i),
            "asat_sys_rhs": ([vp, vp, vp, ctypes.c_double], i),
            "asat_sys_violations": ([vp, vp, ip], i),
            "asat_sys_set_random_state": ([vp, vp, vp], i),
            "asat_sys_clause_order": ([vp, vp], i),
            "asat_sys_layout": ([vp, ctypes.POINTER(LayoutInfo)], i),
            "asat_sys_set_family": ([vp, ctypes.c_int], i),
            "asat_sys_get_family": ([vp, ip], i),
            "asat_sys_set_q": ([vp, ctypes.c_double], i),
            "asat_sys_get_q": ([vp, ctypes.POINTER(ctypes.c_double)], i),
            "asat_state_create": ([vp, pvp], i),
            "asat_state_destroy": ([vp], i),
            "asat_state_set_public": ([vp, vp], i),
            "asat_state_get_public": ([vp, vp], i),
            "asat_state_get_boolean": ([vp, vp], i),
            "asat_state_copy": ([vp, vp], i),
            "asat_state_set_zero": ([vp], i),
            "asat_state_alloc_size": ([vp, ip], i),
            "asat_state_device_ptr": ([vp, pvp, ip], i),
            "asat_rk_create": ([pvp], i),
            "asat_rk_destroy": ([vp], i),
            "asat_rk_set_eps": ([vp, ctypes.c_double], i),
            "asat_rk_get_eps": ([vp, ctypes.POINTER(ctypes.c_double)], i),
            "asat_rk_set_xscal_floor": ([vp, ctypes.c_double], i),
            "asat_batch_set_xscal_floor": ([vp, ctypes.c_double], i),
            "asat_rk_configure": ([vp, vp], i),
            "asat_rk_step": ([vp, vp, vp, i], i),
            "asat_solve_noalloc": ([i, vp, vp, i, vp, ctypes.c_double, i, ctypes.c_longlong, ctypes.c_double, i, vp,
                                    ctypes.POINTER(SolveResult), ctypes.POINTER(RkStats)], i),
            "asat_rk_set_one_workgroup": ([vp, i], i),
            "asat_rk_uses_one_workgroup": ([vp, ctypes.POINTER(i)], i),
            "asat_rk_solve": ([vp, vp, vp, i, ctypes.c_longlong, ctypes.c_double, i, ctypes.POINTER(SolveResult)], i),
            "asat_rk_reset": ([vp], i),
            "asat_rk_stats": ([vp, ctypes.POINTER(RkStats)], i),
            "asat_rk_last_step_ms": ([vp, ctypes.POINTER(ctypes.c_float)], i),
            "asat_rk_profile": ([vp, vp, vp, i, vp], i),
            "asat_rk_set_kernel_timer": ([vp, i], i),
            "asat_rk_kernel_timer": ([vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_longlong)], i),
            "asat_util_copy_bench": ([ctypes.c_longlong, i, ctypes.POINTER(ctypes.c_float)], i),
            "asat_util_gpu_name": ([ctypes.c_char_p, i], i),
            "asat_util_stream_bench": ([ctypes.c_longlong, i, ctypes.POINTER(ctypes.c_float),
                                        ctypes.POINTER(ctypes.c_float)], i),
            "asat_rk_set_step_log": ([vp, vp, ctypes.c_longlong], i),
            "asat_adj_create": ([pvp], i),
            "asat_adj_destroy": ([vp], i),
            "asat_adj_set_problem": ([vp, vp, vp, i], i),
            "asat_adj_set_params": ([vp, ctypes.c_double, i, i], i),
            "asat_adj_forward": ([vp, i, vp, vp], i),
            "asat_adj_get_steps": ([vp, vp], i),
            "asat_adj_backward": ([vp, vp], i),
            "asat_adj_replay": ([vp, i, vp, vp, vp, vp], i),
            "asat_adj_replay_f64": ([vp, i, vp, vp, vp], i),
            "asat_adj_timing": ([vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)], i),
            "asat_adj_attempts": ([vp, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong)], i),
            "asat_rng_create": ([ctypes.c_ulonglong, pvp], i),
            "asat_rng_destroy": ([vp], i),
            "asat_rng_uniform": ([vp, vp, i, i], i),
            "asat_batch_create": ([pvp], i),
            "asat_batch_destroy": ([vp], i),
            "asat_batch_add": ([vp, vp, vp, i, vp, ip], i),
            "asat_batch_set_params": ([vp, ctypes.c_double, ctypes.c_longlong, ctypes.c_double, i, i], i),
            "asat_batch_set_tiling": ([vp, i, i, i], i),
            "asat_batch_set_chunk": ([vp, i], i),
            "asat_batch_prepare": ([vp, i], i),
            "asat_batch_run": ([vp, i, i], i),
            "asat_batch_reset": ([vp], i),
            "asat_batch_size": ([vp, ip], i),
            "asat_batch_result": ([vp, i, ctypes.POINTER(BatchResult)], i),
            "asat_batch_get_state": ([vp, i, vp], i),
            "asat_batch_stats": ([vp, ctypes.POINTER(ctypes.c_double), ip, ip], i),
            "asat_batch_bucket_info": ([vp, i, ctypes.POINTER(BatchBucket)], i),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def copy_bench(nbytes, reps=10):
    """Microseconds per plain device copy of nbytes (HBM counter calibration)."""
    us = ctypes.c_float()
    _check(lib().asat_util_copy_bench(int(nbytes), int(reps), ctypes.byref(us)))
    return us.value


def stream_bench(nbytes=1 << 30, reps=5):
    """Practical HBM rates (GB/s) of a well-formed copy and read-only pass."""
    uc, ur = ctypes.c_float(), ctypes.c_float()
    _check(lib().asat_util_stream_bench(int(nbytes), int(reps), ctypes.byref(uc), ctypes.byref(ur)))
    return {"copy_GBps": 2 * nbytes / (uc.value * 1e-6) / 1e9, "read_GBps": nbytes / (ur.value * 1e-6) / 1e9}


def GetGpuName():
    """Name of the current device (AnalogSAT/util/gpuidtool.cpp:28-40)."""
    buf = ctypes.create_string_buffer(256)
    _check(lib().asat_util_gpu_name(buf, len(buf)))
    return buf.value.decode(errors="replace")


def _check(rc):
    if rc == ASAT_OK:
        return rc
    msg = lib().asat_last_error().decode()
    if rc == ASAT_ERR_INVALID_ARGUMENT:
        raise ValueError(msg)
    if rc == ASAT_UNDERFLOW:
        return rc
    raise RuntimeError(msg)


def _prec(precision):
    if precision in ("f64", "double", np.float64, ASAT_F64):
        return ASAT_F64, np.float64
    if precision in ("f32", "float", np.float32, ASAT_F32):
        return ASAT_F32, np.float32
    raise ValueError("precision must be 'f64' or 'f32'")


# ----------------------------------------------------------------------------
# problems
# ----------------------------------------------------------------------------
class SatProblem:
    """Flat clause list (AnalogSAT/problem/SatProblem.h:32-73).

    Holds the clauses as given; zero literals and empty clauses are dropped by
    the library exactly as ``SatProblem::AddClauses`` does.
    """

    def __init__(self, literals=None, lengths=None):
        self.literals = np.zeros(0, np.int32) if literals is None else np.ascontiguousarray(literals, np.int32)
        self.lengths = np.zeros(0, np.int32) if lengths is None else np.ascontiguousarray(lengths, np.int32)

    def AddClause(self, clause):
        self.AddClauses(list(clause), [len(clause)])

    def AddClauses(self, literals, lengths):
        self.literals = np.concatenate([self.literals, np.asarray(literals, np.int32)])
        self.lengths = np.concatenate([self.lengths, np.asarray(lengths, np.int32)])

    def Clear(self):
        self.literals = np.zeros(0, np.int32)
        self.lengths = np.zeros(0, np.int32)

    def _nonzero(self):
        keep_l = self.literals != 0
        clause_id = np.repeat(np.arange(len(self.lengths)), self.lengths)
        nz_len = np.bincount(clause_id[keep_l], minlength=len(self.lengths))
        return keep_l, nz_len

    def Get_N(self):
        return int(np.abs(self.literals).max()) if len(self.literals) else 0

    def Get_M(self):
        return int((self._nonzero()[1] > 0).sum()) if len(self.lengths) else 0

    def Get_K(self):
        return int(self._nonzero()[1].max()) if len(self.lengths) else 0

    @classmethod
    def from_clauses(cls, clauses):
        return cls([l for c in clauses for l in c], [len(c) for c in clauses])

    @classmethod
    def read_cnf(cls, path):
        """DIMACS reader with CnfReader semantics (AnalogSAT/io/CnfReader.cpp:48-141)."""
        with open(path) as f:
            text = f.read()
        n = m = -1
        toks = []
        for line in text.splitlines():
            t = line.strip()
            if not t or t[0] in "c%":
                continue
            if t[0] == "p":
                parts = t.split()
                if len(parts) < 4 or parts[1] != "cnf":
                    raise ValueError("expected: p cnf <variables> <clauses>")
                n, m = int(parts[2]), int(parts[3])
                continue
            toks.extend(int(v) for v in t.split())
        if n < 0 or m < 0:
            raise ValueError("could not find a SAT declaration")
        lits, lens, cur = [], [], 0
        for v in toks:
            if len(lens) == m:
                break
            if v == 0:
                lens.append(cur)
                cur = 0
            else:
                lits.append(v)
                cur += 1
        if len(lens) < m:
            raise ValueError("file is missing %d clauses" % (m - len(lens)))
        return cls(lits, lens)

    def write_cnf(self, path):
        """CnfWriter semantics (AnalogSAT/io/CnfWriter.cpp:49-74)."""
        keep_l, nz = self._nonzero()
        lits = self.literals[keep_l]
        with open(path, "w") as f:
            f.write("p cnf %d %d\n" % (self.Get_N(), int((nz > 0).sum())))
            off = 0
            for L in nz:
                if L == 0:
                    continue
                f.write("".join("%d " % v for v in lits[off:off + L]) + "0\n")
                off += L


class FastSatMaker:
    """AnalogSAT/problem/FastSatMaker.h:29-61 -- same seed, same instance stream.

    Successive MakeSatProblem calls continue one generator stream, like the
    reference's object.
    """

    def __init__(self, seed=None):
        seed = int(_time.time_ns() & 0xFFFFFFFF) if seed is None else int(seed)
        self._engine = ctypes.c_ulonglong()
        _check(lib().asat_problem_seed_engine(seed, ctypes.byref(self._engine)))
        self.k, self.n, self.m = 3, 100, 325   # FastSatMaker::SetDefaults

    def Set_K(self, k): self.k = int(k)
    def Set_N(self, n): self.n = int(n)
    def Set_M(self, m): self.m = int(m)
    def Get_K(self): return self.k
    def Get_N(self): return self.n
    def Get_M(self): return self.m

    def MakeSatProblem(self, problem=None):
        out = np.zeros(self.m * self.k, np.int32)
        _check(lib().asat_problem_fastsat_stream(self.n, self.m, self.k, ctypes.byref(self._engine),
                                                 out.ctypes.data))
        p = SatProblem(out, np.full(self.m, self.k, np.int32))
        if problem is not None:
            problem.literals, problem.lengths = p.literals, p.lengths
        return p


class RandomSatMaker(FastSatMaker):
    """AnalogSAT/problem/RandomSatMaker.h:29-63 -- FastSatMaker's draw with
    repeated clauses drawn again; same seed, same instance stream."""

    def MakeSatProblem(self, problem=None):
        out = np.zeros(self.m * self.k, np.int32)
        _check(lib().asat_problem_randomsat_stream(self.n, self.m, self.k, ctypes.byref(self._engine),
                                                   out.ctypes.data))
        p = SatProblem(out, np.full(self.m, self.k, np.int32))
        if problem is not None:
            problem.literals, problem.lengths = p.literals, p.lengths
        return p


def randomsat(n, m, k, seed):
    """One RandomSatMaker(seed) instance."""
    out = np.zeros(m * k, np.int32)
    _check(lib().asat_problem_randomsat(n, m, k, seed, out.ctypes.data))
    return SatProblem(out, np.full(m, k, np.int32))


def fastsat(n, m, k, seed):
    """One FastSatMaker(seed) instance."""
    out = np.zeros(m * k, np.int32)
    _check(lib().asat_problem_fastsat(n, m, k, seed, out.ctypes.data))
    return SatProblem(out, np.full(m, k, np.int32))


# ----------------------------------------------------------------------------
# system, states, integrator, rng
# ----------------------------------------------------------------------------
class HipState:
    """ISatState over a device buffer (AnalogSAT/solver/ISatState.h:31-161)."""

    def __init__(self, sat):
        self.sat = sat
        self._h = ctypes.c_void_p()
        _check(lib().asat_state_create(sat._h, ctypes.byref(self._h)))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and h.value and _lib is not None:
            _lib.asat_state_destroy(h)
            self._h = None

    def Get_N(self): return self.sat.Get_N()
    def Get_M(self): return self.sat.Get_M()

    def CopyTo(self, target=None):
        out = np.zeros(self.sat.Get_N() + self.sat.Get_M(), self.sat.dtype)
        _check(lib().asat_state_get_public(self._h, out.ctypes.data))
        if target is not None:
            target[:] = out
        return out

    def CopyFrom(self, source):
        if isinstance(source, HipState):
            _check(lib().asat_state_copy(self._h, source._h))
            return
        src = np.ascontiguousarray(source, dtype=self.sat.dtype)
        if src.size < self.sat.Get_N() + self.sat.Get_M():
            raise ValueError("source vector is too short")
        _check(lib().asat_state_set_public(self._h, src.ctypes.data))

    def CopyBooleanTo(self):
        out = np.zeros(self.sat.Get_N(), np.uint8)
        _check(lib().asat_state_get_boolean(self._h, out.ctypes.data))
        return out.astype(bool)

    def SetZero(self):
        _check(lib().asat_state_set_zero(self._h))

    def GetAllocSize(self):
        v = ctypes.c_int()
        _check(lib().asat_state_alloc_size(self._h, ctypes.byref(v)))
        return v.value

    def GetState(self):
        """(device pointer, aux offset) of the internal layout."""
        p, off = ctypes.c_void_p(), ctypes.c_int()
        _check(lib().asat_state_device_ptr(self._h, ctypes.byref(p), ctypes.byref(off)))
        return p.value, off.value


class HipRandom:
    """ISatRandom (AnalogSAT/solver/ISatRandom.h:25-33) on a Philox4x32-10 stream."""

    def __init__(self, seed=None):
        self._h = ctypes.c_void_p()
        seed = int(_time.time_ns()) & 0xFFFFFFFFFFFFFFFF if seed is None else int(seed)
        _check(lib().asat_rng_create(seed, ctypes.byref(self._h)))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and h.value and _lib is not None:
            _lib.asat_rng_destroy(h)

    def GenerateUniform(self, dev_ptr, length, precision="f64"):
        _check(lib().asat_rng_uniform(self._h, ctypes.c_void_p(dev_ptr), int(length), _prec(precision)[0]))


class HipSat:
    """ISat on the GPU (AnalogSAT/solver/ISat.h:31-87); precision 'f64' or 'f32'."""

    def __init__(self, precision="f64"):
        self.precision, self.dtype = _prec(precision)
        self._h = ctypes.c_void_p()
        _check(lib().asat_sys_create(self.precision, ctypes.byref(self._h)))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and h.value and _lib is not None:
            _lib.asat_sys_destroy(h)

    def SetProblem(self, problem):
        lits = np.ascontiguousarray(problem.literals, np.int32)
        lens = np.ascontiguousarray(problem.lengths, np.int32)
        _check(lib().asat_sys_set_problem(self._h, lits.ctypes.data, lens.ctypes.data, len(lens)))

    def _nmk(self):
        n, m, k = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(lib().asat_sys_get_nmk(self._h, ctypes.byref(n), ctypes.byref(m), ctypes.byref(k)))
        return n.value, m.value, k.value

    def Get_N(self): return self._nmk()[0]
    def Get_M(self): return self._nmk()[1]
    def Get_K(self): return self._nmk()[2]

    def Get_B(self):
        b = ctypes.c_double()
        _check(lib().asat_sys_get_b(self._h, ctypes.byref(b)))
        return b.value

    def Set_B(self, b):
        _check(lib().asat_sys_set_b(self._h, float(b)))

    def MakeState(self):
        return HipState(self)

    def MakeRandomState(self, random):
        st = self.MakeState()
        self.SetRandomState(st, random)
        return st

    def SetRandomState(self, state, random):
        _check(lib().asat_sys_set_random_state(self._h, state._h, random._h))

    def GetDerivatives(self, dxdt, state, time=0.0):
        _check(lib().asat_sys_rhs(self._h, dxdt._h, state._h, float(time)))

    def GetClauseViolationCount(self, state):
        if not isinstance(state, HipState):
            tmp = self.MakeState()
            tmp.CopyFrom(state)
            state = tmp
        v = ctypes.c_int()
        _check(lib().asat_sys_violations(self._h, state._h, ctypes.byref(v)))
        return v.value

    def Get_Q(self):
        q = ctypes.c_double()
        _check(lib().asat_sys_get_q(self._h, ctypes.byref(q)))
        return q.value

    def Set_Q(self, q):
        _check(lib().asat_sys_set_q(self._h, float(q)))

    def ClauseOrder(self):
        """Clause index of each position of a state's device aux block."""
        out = np.zeros(self.Get_M(), np.int32)
        _check(lib().asat_sys_clause_order(self._h, out.ctypes.data))
        return out

    def Layout(self):
        """The passes and launch shapes chosen for the current problem (dict):
        clause_pass 0 tiled / 1 generic, var_pass 0 LDS regions / 1 gather /
        2 one workgroup per variable, groups, group_vars, tiles, grids."""
        out = LayoutInfo()
        _check(lib().asat_sys_layout(self._h, ctypes.byref(out)))
        return out.as_dict()


class HipSatTanh(HipSat):
    """CpuSatTanh / CudaSatTanh3 on the GPU (AnalogSAT/cpu_sat/CpuSatTanh.cpp:35-117):
    a slot adds knorm a |K_mi| tanh(q f_mi) c_mi; q defaults to 1.1."""

    def __init__(self, precision="f64", q=1.1):
        super().__init__(precision)
        _check(lib().asat_sys_set_family(self._h, 1))
        self.Set_Q(q)


class HipRungeKutta:
    """IODEInt (AnalogSAT/solver/IODEInt.h:28-60) -- Cash-Karp RK45, device-side step control."""

    def __init__(self):
        self._h = ctypes.c_void_p()
        _check(lib().asat_rk_create(ctypes.byref(self._h)))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and h.value and _lib is not None:
            _lib.asat_rk_destroy(h)

    def SetEpsilon(self, eps):
        _check(lib().asat_rk_set_eps(self._h, float(eps)))

    def SetErrorFloor(self, floor):
        """xscal = |x| + |h dxdt| + floor (reference: floor = TINY = 1e-20, the default)."""
        _check(lib().asat_rk_set_xscal_floor(self._h, float(floor)))

    def GetEpsilon(self):
        e = ctypes.c_double()
        _check(lib().asat_rk_get_eps(self._h, ctypes.byref(e)))
        return e.value

    def Configure(self, sat):
        self._sat = sat
        _check(lib().asat_rk_configure(self._h, sat._h))

    def Step(self, old_state, new_state, count):
        return _check(lib().asat_rk_step(self._h, old_state._h, new_state._h, int(count))) == ASAT_OK

    def Reset(self):
        _check(lib().asat_rk_reset(self._h))

    def SetOneWorkgroup(self, enable):
        """Allow (True, the default) or forbid the one-workgroup persistent
        path for problems that fit in one workgroup's LDS."""
        _check(lib().asat_rk_set_one_workgroup(self._h, 1 if enable else 0))

    def UsesOneWorkgroup(self):
        v = ctypes.c_int()
        _check(lib().asat_rk_uses_one_workgroup(self._h, ctypes.byref(v)))
        return bool(v.value)

    def SolveBatches(self, old_state, new_state, batch, max_steps, max_time, time_float=False):
        """SatSolver's loop without a callback (asat_rk_solve): batches of
        `batch` accepted steps, SAT checked after each; on the one-workgroup
        path the whole loop runs on the device.  Returns the SolveResult."""
        out = SolveResult()
        _check(lib().asat_rk_solve(self._h, old_state._h, new_state._h, int(batch), int(max_steps), float(max_time),
                                   1 if time_float else 0, ctypes.byref(out)))
        return out

    def stats(self):
        s = RkStats()
        _check(lib().asat_rk_stats(self._h, ctypes.byref(s)))
        return s

    def GetTime(self): return self.stats().time
    def GetLastStepSize(self): return self.stats().hdid
    def GetRHSCount(self): return int(self.stats().rhscount)
    def GetStepCount(self): return int(self.stats().steps)
    def GetAttemptCount(self): return int(self.stats().attempts)

    def profile_kernels(self, old_state, new_state, reps=20):
        """Average microseconds per launch: clause J=1..6, var J=1..6, control."""
        us = np.zeros(13, np.float32)
        _check(lib().asat_rk_profile(self._h, old_state._h, new_state._h, int(reps), us.ctypes.data))
        return us

    def set_kernel_timer(self, kind):
        """Time one kernel kind (0..12, the profile_kernels order; -1: off)
        inside every attempt of the following Step calls."""
        _check(lib().asat_rk_set_kernel_timer(self._h, int(kind)))

    def kernel_timer(self):
        """(total ms, launches) of the kernel timer since it was set."""
        ms, n = ctypes.c_double(), ctypes.c_longlong()
        _check(lib().asat_rk_kernel_timer(self._h, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def last_step_ms(self):
        v = ctypes.c_float()
        _check(lib().asat_rk_last_step_ms(self._h, ctypes.byref(v)))
        return v.value


# ----------------------------------------------------------------------------
# SatSolver (AnalogSAT/solver/SatSolver.h:41-356)
# ----------------------------------------------------------------------------
class SatSolver:
    """Batch loop around an ISat and an IODEInt; same semantics as the reference.

    SAT is checked and limits are compared only between batches of
    ``batchSize`` accepted steps (SatSolver.h:295-331).
    """

    def __init__(self, sat, odeint):
        self.sat = sat
        self.odeint = odeint
        self.maxTime = 100.0
        self.maxSteps = 100000
        self.batchSize = 20
        self.callback = None
        self.running = False
        self.oldState = self.newState = None
        self._Reset()

    def _Reset(self):
        self.rhsCount = 0
        self.walltime = 0.0
        self.cputime = 0.0
        self.steps = 0
        self.time = 0.0
        self.lastStepSize = 0.0
        self.lastViolationCount = 0

    def GetStepCount(self): return self.steps
    def GetRhsCount(self): return self.rhsCount
    def GetElapsedTime(self): return self.time
    def GetLastStepSize(self): return self.lastStepSize
    def GetLastClauseViolationCount(self): return self.lastViolationCount

    def GetElapsedWallTime(self):
        if self.running:
            self.walltime = _time.perf_counter() - self._wallstart
        return self.walltime

    def GetElapsedCpuTime(self):
        """Process CPU seconds since the start of the last Solve (SatSolver.h:107-121)."""
        if self.running:
            self.cputime = _time.process_time() - self._cpustart
        return self.cputime

    def SetCallback(self, fn): self.callback = fn
    def ClearCallback(self): self.callback = None
    def GetBatchSize(self): return self.batchSize

    def SetBatchSize(self, b):
        if b < 1:
            raise ValueError("Batch size must be positive")
        self.batchSize = int(b)

    def GetMaxTime(self): return self.maxTime

    def SetMaxTime(self, t):
        if t <= 0:
            raise ValueError("Max time must be positive")
        self.maxTime = float(t)

    def GetMaxSteps(self): return self.maxSteps

    def SetMaxSteps(self, s):
        if s < 1:
            raise ValueError("Max iterations must be positive")
        self.maxSteps = int(s)

    def SetProblem(self, problem):
        self.sat.SetProblem(problem)
        self._Reset()
        self.oldState = self.sat.MakeState()
        self.newState = self.sat.MakeState()
        self.lastViolationCount = self.sat.GetClauseViolationCount(self.oldState)

    def GetSatState(self):
        return self.oldState

    def SetSatState(self, state):
        self.oldState.CopyFrom(state)
        self.lastViolationCount = self.sat.GetClauseViolationCount(self.oldState)

    def SetRandomInitialState(self, random):
        self.sat.SetRandomState(self.oldState, random)
        self.lastViolationCount = self.sat.GetClauseViolationCount(self.oldState)

    def Solve(self):
        self.running = True
        self._Reset()
        self.odeint.Configure(self.sat)
        self._wallstart = _time.perf_counter()
        self._cpustart = _time.process_time()
        result = SatResult.SAT_UNKNOWN
        if self.callback is None and hasattr(self.odeint, "SolveBatches"):
            # no callback: the integrator runs the same loop itself
            r = self.odeint.SolveBatches(self.oldState, self.newState, self.batchSize, self.maxSteps, self.maxTime)
            self.steps = int(r.steps)
            self.lastViolationCount = int(r.last_violations)
            result = SatResult(r.result)
            self.walltime = _time.perf_counter() - self._wallstart
            self.cputime = _time.process_time() - self._cpustart
            st = self.odeint.stats()
            self.time = st.time
            self.rhsCount = int(st.rhscount)
            self.lastStepSize = st.hdid
            self.running = False
            return result
        while self.time < self.maxTime and self.steps < self.maxSteps:
            before = self.odeint.GetStepCount()
            if not self.odeint.Step(self.oldState, self.newState, self.batchSize):
                result = SatResult.SAT_UNDERFLOW
                break
            st = self.odeint.stats()
            self.steps += int(st.steps) - before
            self.time = st.time
            self.lastViolationCount = self.sat.GetClauseViolationCount(self.oldState)
            self.rhsCount = int(st.rhscount)
            self.lastStepSize = st.hdid
            if self.callback is not None and not self.callback(self):
                result = SatResult.SAT_ODE_INTERRUPTED
                break
            if self.lastViolationCount == 0:
                result = SatResult.SAT_SOLUTION_FOUND
                break
        if result == SatResult.SAT_UNKNOWN:
            if self.time >= self.maxTime:
                result = SatResult.SAT_MAXTIME_REACHED
            elif self.steps >= self.maxSteps:
                result = SatResult.SAT_MAXITER_REACHED
        self.walltime = _time.perf_counter() - self._wallstart
        self.cputime = _time.process_time() - self._cpustart
        st = self.odeint.stats()
        self.time = st.time
        self.rhsCount = int(st.rhscount)
        self.lastStepSize = st.hdid
        self.running = False
        return result


# ----------------------------------------------------------------------------
# batched ragged solver (no reference counterpart)
# ----------------------------------------------------------------------------
BATCH_STATUS = {0: "running", 1: "solved", 2: "max steps", 3: "underflow", 4: "max time"}


class HipBatch:
    """Many independent trajectories in persistent launches (asat_batch_*).

    Every trajectory runs the FP32-state / FP64-control Cash-Karp integrator
    of an ``HipSat('f32')`` system; SAT is checked every ``check_every``
    accepted steps, the minimum violated-clause count is tracked.
    """

    def __init__(self):
        self._h = ctypes.c_void_p()
        _check(lib().asat_batch_create(ctypes.byref(self._h)))
        self._sizes = []

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and h.value and _lib is not None:
            _lib.asat_batch_destroy(h)

    def add(self, problem, pub0):
        lits = np.ascontiguousarray(problem.literals, np.int32)
        lens = np.ascontiguousarray(problem.lengths, np.int32)
        x0 = np.ascontiguousarray(pub0, np.float32)
        idx = ctypes.c_int()
        _check(lib().asat_batch_add(self._h, lits.ctypes.data, lens.ctypes.data, len(lens), x0.ctypes.data,
                                    ctypes.byref(idx)))
        self._sizes.append(len(x0))
        return idx.value

    def set_params(self, eps=1e-6, max_steps=10000, max_time=1e300, chunk_steps=256, check_every=1):
        _check(lib().asat_batch_set_params(self._h, float(eps), int(max_steps), float(max_time), int(chunk_steps),
                                           int(check_every)))

    def set_error_floor(self, floor):
        _check(lib().asat_batch_set_xscal_floor(self._h, float(floor)))

    def set_tiling(self, block=0, lds_mode=-1, nbuckets=4):
        _check(lib().asat_batch_set_tiling(self._h, int(block), int(lds_mode), int(nbuckets)))

    def set_chunk(self, clauses=0):
        """Clause-chunked kernel (lds_mode 3): at most this many clauses per
        chunk (0: as many as the LDS holds)."""
        _check(lib().asat_batch_set_chunk(self._h, int(clauses)))

    def prepare(self, autotune=True):
        _check(lib().asat_batch_prepare(self._h, 1 if autotune else 0))

    def run(self, max_chunks=1 << 30, graph_chunks=4):
        _check(lib().asat_batch_run(self._h, int(max_chunks), int(graph_chunks)))

    def reset(self):
        _check(lib().asat_batch_reset(self._h))

    def __len__(self):
        return len(self._sizes)

    def result(self, i):
        r = BatchResult()
        _check(lib().asat_batch_result(self._h, int(i), ctypes.byref(r)))
        return {f: getattr(r, f) for f, _ in BatchResult._fields_}

    def results(self):
        return [self.result(i) for i in range(len(self))]

    def state(self, i):
        out = np.zeros(self._sizes[i], np.float32)
        _check(lib().asat_batch_get_state(self._h, int(i), out.ctypes.data))
        return out

    def stats(self):
        ms, ch, nb = ctypes.c_double(), ctypes.c_int(), ctypes.c_int()
        _check(lib().asat_batch_stats(self._h, ctypes.byref(ms), ctypes.byref(ch), ctypes.byref(nb)))
        return {"run_ms": ms.value, "chunks": ch.value, "buckets": nb.value}

    def buckets(self):
        out = []
        for b in range(self.stats()["buckets"]):
            r = BatchBucket()
            _check(lib().asat_batch_bucket_info(self._h, b, ctypes.byref(r)))
            out.append({f: getattr(r, f) for f, _ in BatchBucket._fields_})
        return out


# ----------------------------------------------------------------------------
# discrete adjoint (no reference counterpart)
# ----------------------------------------------------------------------------
class HipAdjoint:
    """Gradient of V(T) = sum_m a_m K_m^2 after T accepted FP32 steps with
    respect to [s(0), a(0)] for a batch of trajectories of one instance
    (asat_adj_*).  Forward adaptive with logged steps and checkpoints every
    ``checkpoint_every`` steps; backward in FP64."""

    def __init__(self, problem, eps=1e-6, steps=500, checkpoint_every=32):
        self._h = ctypes.c_void_p()
        _check(lib().asat_adj_create(ctypes.byref(self._h)))
        lits = np.ascontiguousarray(problem.literals, np.int32)
        lens = np.ascontiguousarray(problem.lengths, np.int32)
        _check(lib().asat_adj_set_problem(self._h, lits.ctypes.data, lens.ctypes.data, len(lens)))
        _check(lib().asat_adj_set_params(self._h, float(eps), int(steps), int(checkpoint_every)))
        self.T = int(steps)
        self.nvar = problem.Get_N() + problem.Get_M()
        self.ntraj = 0

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and h.value and _lib is not None:
            _lib.asat_adj_destroy(h)

    def forward(self, pub0):
        """pub0: (ntraj, n+m) start states; returns V(T) per trajectory (float64)."""
        x = np.ascontiguousarray(pub0, np.float32).reshape(-1, self.nvar)
        self.ntraj = x.shape[0]
        V = np.zeros(self.ntraj)
        _check(lib().asat_adj_forward(self._h, self.ntraj, x.ctypes.data, V.ctypes.data))
        return V

    def steps(self):
        h = np.zeros((self.ntraj, self.T))
        _check(lib().asat_adj_get_steps(self._h, h.ctypes.data))
        return h

    def backward(self):
        """dV(T)/d[s(0), a(0)] of the last forward run, shape (ntraj, n+m), float64."""
        g = np.zeros((self.ntraj, self.nvar))
        _check(lib().asat_adj_backward(self._h, g.ctypes.data))
        return g

    def replay(self, pub0, h):
        """V(T) and x(T) of T FP32 steps with the given step sizes (ntraj, T)."""
        x = np.ascontiguousarray(pub0, np.float32).reshape(-1, self.nvar)
        h = np.ascontiguousarray(h, np.float64).reshape(x.shape[0], self.T)
        V = np.zeros(x.shape[0])
        xT = np.zeros_like(x)
        _check(lib().asat_adj_replay(self._h, x.shape[0], x.ctypes.data, h.ctypes.data, V.ctypes.data, xT.ctypes.data))
        self.ntraj = 0
        return V, xT

    def replay_f64(self, pub0, h):
        """V(T) of T steps with the given step sizes and FP64 state."""
        x = np.ascontiguousarray(pub0, np.float64).reshape(-1, self.nvar)
        h = np.ascontiguousarray(h, np.float64).reshape(x.shape[0], self.T)
        V = np.zeros(x.shape[0])
        _check(lib().asat_adj_replay_f64(self._h, x.shape[0], x.ctypes.data, h.ctypes.data, V.ctypes.data))
        self.ntraj = 0
        return V

    def timing(self):
        f, b = ctypes.c_double(), ctypes.c_double()
        _check(lib().asat_adj_timing(self._h, ctypes.byref(f), ctypes.byref(b)))
        return f.value, b.value

    def attempts(self):
        """(attempts summed over the trajectories, lockstep attempts) of the last forward run."""
        t, l = ctypes.c_longlong(), ctypes.c_longlong()
        _check(lib().asat_adj_attempts(self._h, ctypes.byref(t), ctypes.byref(l)))
        return t.value, l.value


def energy_function():
    """torch.autograd.Function computing V(T) per trajectory from (s0, a0).

    Usage: V = energy_function().apply(s0, a0, adjoint) with s0 (ntraj, n),
    a0 (ntraj, m) tensors on any device; backward returns FP64 gradients
    produced by asat_adj_backward (cast to the inputs' dtype/device).
    """
    import torch

    class CTDSEnergy(torch.autograd.Function):
        @staticmethod
        def forward(ctx, s0, a0, adjoint):
            pub = torch.cat([s0, a0], dim=1).detach().to("cpu", torch.float32).numpy()
            V = adjoint.forward(pub)
            ctx.adjoint = adjoint
            ctx.n = s0.shape[1]
            ctx.meta = (s0.dtype, s0.device, a0.dtype, a0.device)
            return torch.from_numpy(V).to(s0.device)

        @staticmethod
        def backward(ctx, gV):
            g = torch.from_numpy(ctx.adjoint.backward())
            g = g * gV.detach().to("cpu", torch.float64)[:, None]
            sd, sdev, ad, adev = ctx.meta
            return g[:, :ctx.n].to(sdev, sd), g[:, ctx.n:].to(adev, ad), None

    return CTDSEnergy
