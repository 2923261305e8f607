// One-workgroup persistent integrator for problems that fit in LDS.
//
// The plain solver's attempt is 13 kernels; for a small instance (such as
// N = 50, M = 213) every one of them is a few microseconds of dispatch for
// well under a microsecond of work.  Here one workgroup holds the whole
// integrator in LDS -- state x, x_new, k1..k5, the stage spin input, the slot
// contributions (in var-major order: each variable's terms are one run, in
// the reference's accumulation order), the literals and each slot's
// var-major position -- and runs
// attempt after attempt with workgroup barriers between the clause and var
// phases of each stage.  Every thread evaluates the FP64 step controller
// itself from the block-wide error maximum, so no thread waits for a
// broadcast.  Optionally the SatSolver loop runs on the device too: after
// every `batch` accepted steps the violated-clause count of the thresholded
// state is taken (by the next step's first clause phase, which reads those
// spins anyway, with one ballot per wave) and the run ends on SAT, on the
// step or time limit, or on step-size underflow
// (AnalogSAT/solver/SatSolver.h:313-331, the branch without a callback).
//
// Arithmetic follows the multi-kernel path operation for operation (the
// generic clause pass k_clause_gen and the var passes' var_stage), so the
// two paths give bit-identical states, step sizes and counters:
//   clause j:  f_q = 1 -+ s (padding literal: s = -1), K_m = knorm prod f_q
//              left to right, aux stage input from x and k_1..k_{J-1},
//              k_J(a) = a K^2, slot terms (KmiProd order, stop at the first
//              padding literal for K <= 10; CalculateAll order above)
//   var i:     contributions added in the reference's accumulation order
//              (clauses in input order, literals in clause order), then the
//              stage combination (AnalogSAT/cpu_sat/CpuRungeKutta.cpp:180-215)
//   control:   AnalogSAT/cpu_sat/CpuRungeKutta.cpp:217-267 in FP64
// The bias term (b > 0) is not covered: such systems take the multi-kernel
// path.
//
// The family is a template parameter (FAM): the whole attempt loop is inlined,
// and a tanh branch compiled into every stage of the CTDS kernel would more
// than double its code (about 10 700 instructions for the FP64 K = 3 kernel
// against 3 800 without), past what the instruction cache holds, with one
// wave per SIMD to hide the misses.  The tanh family is built in the generic
// (KT = 0) form only.
#pragma once

#include "asat_common.h"

namespace asat {

constexpr int SMALL_LDS_MAX = 150 * 1024;   // dynamic LDS of the small path
constexpr long long SMALL_ATTEMPT_CAP = 8192;   // attempts per launch before the kernel pauses

// SatSolver loop state on the device, kept across launches
struct SmallSolve {
    long long solver_steps;   // SatSolver::steps: accepted steps of completed batches
    int bleft;                // accepted steps left in the current batch (0: at the top of the loop)
    int result;               // -1 while running, else a SatResult code
    int last_viol;            // SatSolver::lastViolationCount
    int check;                // a batch just ended: count the violated clauses of x before going on
};

// SatResult codes (AnalogSAT/solver/SatResult.h:24-32)
constexpr int SMALL_SAT = 0, SMALL_UNDERFLOW = 1, SMALL_MAXTIME = 3, SMALL_MAXITER = 4;

struct SmallArgs {
    const int* lit;       // m*K literals in processing order (zero padded)
    const int* vptr;      // n+2: variable i's terms are var-major positions vptr[i] .. vptr[i+1]-1
    const int* spos;      // m*K: each slot's var-major position (-1: padding)
    void* x;              // state buffer: read at entry, newest state written at exit
    Ctrl* ctrl;
    SmallSolve* solve;    // null: Step mode (run until ctrl->steps reaches target)
    long long target;     // Step mode: accepted-step count to reach
    long long max_steps;  // solve mode: SatSolver::maxSteps
    double max_time;      // solve mode: SatSolver::maxTime
    int time_float;       // solve mode: compare the time as a float (TFloat = float)
    int batch;            // solve mode: SatSolver::batchSize
    int n, m, k, aoff, alloc, nlist;
    int runlen;           // > 0: every variable's run is runlen long (uniform runs), 0: ragged runs
    int family;           // selects the kernel instantiation (FAM)
    double q;
    double knorm;
};

// offset of the term lists after the stage spins (elements): 16-byte
// aligned for the var phase's vector loads
__host__ __device__ inline int small_cb_offset(int n) { return (n + 1 + 3) / 4 * 4; }

// LDS bytes of the small path for a problem
inline long long small_lds_bytes(int n, int m, int k, long long alloc, long long nlist, int esize)
{
    const long long mk = (long long)m * k;
    return (long long)esize * (7 * alloc + small_cb_offset(n) + nlist + 1) + 4 * (2 * mk + (n + 2));
}

// Uniform runs: every variable's run of terms has the same length D, the
// largest degree rounded up to SMALL_RUN_CHUNK, with its terms first and -0.0
// after them.  x + (-0.0) == x for every x under round to nearest, so the
// padding leaves each sum bit for bit as it was; the var phase then adds
// whole chunks read with 16-byte loads, with no per-lane bounds.  Chosen when
// the padded lists are at most SMALL_RUN_WASTE times the terms (plus a chunk
// per variable) and fit in LDS.
constexpr int SMALL_RUN_CHUNK = 8;
constexpr int SMALL_RUN_WASTE = 2;

// Var-major term lists of the one-workgroup kernel.  lit: m rows of k
// literals in processing order (zero padded); visit: the rows in the order
// the reference accumulates them (visit[c] = row of its clause c; null: row
// order).  Writes vptr (n + 2: variable i's run is vptr[i] .. vptr[i+1]-1)
// and spos (m * k: each slot's list position; a padding slot gets nlist, a
// dummy entry after the lists that the clause phase may write and nothing
// reads).  Returns the run length (0: ragged runs, each exactly the
// variable's degree) and the list length in nlist.  Allocates nothing.
inline int small_build_lists(const int* lit, const int* visit, int n, int m, int k, long long alloc, int esize,
                             int* vptr, int* spos, long long* nlist)
{
    for (int v = 0; v < n + 2; v++) vptr[v] = 0;
    long long nl = 0;
    for (long long s = 0; s < (long long)m * k; s++)
        if (lit[s]) { vptr[(lit[s] < 0 ? -lit[s] : lit[s]) + 1]++; nl++; }
    int dmax = 0;
    for (int v = 1; v <= n; v++) dmax = vptr[v + 1] > dmax ? vptr[v + 1] : dmax;
    const long long d = (long long)(dmax + SMALL_RUN_CHUNK - 1) / SMALL_RUN_CHUNK * SMALL_RUN_CHUNK;
    const bool uniform = d > 0 && (long long)n * d <= SMALL_RUN_WASTE * nl + (long long)SMALL_RUN_CHUNK * n &&
                         (long long)n * d < (1LL << 30) &&
                         small_lds_bytes(n, m, k, alloc, (long long)n * d, esize) <= SMALL_LDS_MAX;
    // run starts
    if (uniform) {
        for (int v = 1; v <= n + 1; v++) vptr[v] = (int)((v - 1) * d);
    } else {
        for (int v = 1; v <= n; v++) vptr[v + 1] += vptr[v];
    }
    // fill: vptr[v] is the cursor of variable v while filling, restored
    // after; padding slots get the dummy entry after the lists
    const int dummy = (int)(uniform ? (long long)n * d : nl);
    for (int c = 0; c < m; c++) {
        const long long r = visit ? visit[c] : c;
        for (int q = 0; q < k; q++) {
            const int l = lit[r * k + q];
            spos[r * k + q] = l ? vptr[l < 0 ? -l : l]++ : dummy;
        }
    }
    if (uniform) {
        for (int v = 1; v <= n; v++) vptr[v] = (int)((v - 1) * d);
    } else {
        for (int v = n; v >= 1; v--) vptr[v + 1] = vptr[v];
        vptr[1] = 0;
    }
    *nlist = uniform ? (long long)n * d : nl;
    return uniform ? (int)d : 0;
}

// sum of one uniform run (length a multiple of SMALL_RUN_CHUNK, 16-byte
// aligned), left to right from 0; the loads of up to four chunks are issued
// before the first add
template <typename T>
__device__ __forceinline__ T sum_uniform(const T* run, int len)
{
    constexpr int V = 16 / sizeof(T);
    typedef T vec __attribute__((ext_vector_type(V)));
    constexpr int C = SMALL_RUN_CHUNK, B = 4;   // up to B chunks in flight (len is wave-uniform)
    T s = (T)0;
    for (int c = 0; c < len; c += B * C) {
        T t[B * C];
#pragma unroll
        for (int b = 0; b < B; b++)
            if (b == 0 || c + b * C < len)
#pragma unroll
                for (int q = 0; q < C; q += V) {
                    const vec w = *(const vec*)(run + c + b * C + q);
#pragma unroll
                    for (int u = 0; u < V; u++) t[b * C + q + u] = w[u];
                }
#pragma unroll
        for (int b = 0; b < B; b++)
            if (b == 0 || c + b * C < len)
#pragma unroll
                for (int q = 0; q < C; q++) s += t[b * C + q];
    }
    return s;
}

template <typename T>
struct SmallLds {
    T *x, *xn, *k1, *k2, *k3, *k4, *k5, *sp, *cb;
    int *lit, *vptr, *spos;
};

// Register mode (REG): 256 threads, clauses of KT = 3 or 4 slots, at most
// 512 clauses and variables: each thread keeps its (at most two) clauses'
// literals and var-major positions and its variables' term ranges in
// registers for the whole launch.
constexpr int SMALL_RPT = 2;
template <int KT, int BS>
struct SmallReg {
    static constexpr bool on = KT > 0 && BS == 256;
    int lit[SMALL_RPT][KT > 0 ? KT : 1];
    int pos[SMALL_RPT][KT > 0 ? KT : 1];
    int e0[SMALL_RPT], e1[SMALL_RPT];
};

// one clause of the clause phase of stage J; lm / pm: its literals and its
// slots' var-major positions (LDS pointers or register arrays)
template <typename T, int KT, int J, int FAM, typename LitT, typename PosT>
__device__ __forceinline__ bool small_clause_one(const SmallArgs& a, const SmallLds<T>& L, const T* sp_in,
                                                 const LitT& lm, const PosT& pm, int j, T hs, double h,
                                                 double hnext, double floor, T knorm, double& err, bool count)
{
    constexpr int KM = KT ? KT : GEN_KMAX;
    const int K = KT ? KT : a.k;
    // J = 1 with count: is the clause violated by the thresholded spins
    // (k_violations semantics: literal true when its spin is > 0, resp. not
    // > 0 when negated; a padding literal is FALSE; it reads spin slot 0)
    bool violated = false;
    if constexpr (J == 1) {
        if (count) {
            bool sat = false;
            for (int q = 0; q < K; q++) {
                const int l = lm[q];
                const bool t = sp_in[l < 0 ? -l : l] > (T)0;
                sat = sat || (l > 0 && t) || (l < 0 && !t);
            }
            violated = !sat;
        }
    }
    {
        T f[KM];
        T km = knorm;
#pragma unroll 4
        for (int q = 0; q < K; q++) {
            // spin_factor without its branch: slot 0 of x and of the stage
            // spins holds -1, the padding literal's spin
            const int l = lm[q];
            const T sv = sp_in[l < 0 ? -l : l];
            const T fq = l < 0 ? (T)1.0 + sv : (T)1.0 - sv;
            km *= fq;
            f[q] = fq;
        }
        const int ai = a.aoff + j;
        const T xa = L.x[ai];
        T am, k1 = 0, k3 = 0, k4 = 0, k5 = 0;
        if constexpr (J == 1) {
            am = xa;
        } else {
            k1 = L.k1[ai];
            const T k2 = J > 2 ? L.k2[ai] : (T)0;
            k3 = J > 3 ? L.k3[ai] : (T)0;
            k4 = J > 4 ? L.k4[ai] : (T)0;
            k5 = J > 5 ? L.k5[ai] : (T)0;
            am = stage_input<T, J>(xa, k1, k2, k3, k4, k5, hs);
        }
        const T kj = am * km * km;
        if constexpr (J == 1) L.k1[ai] = kj;
        else if constexpr (J == 2) L.k2[ai] = kj;
        else if constexpr (J == 3) L.k3[ai] = kj;
        else if constexpr (J == 4) L.k4[ai] = kj;
        else if constexpr (J == 5) L.k5[ai] = kj;
        else {
            T xn;
            double e;
            final_combination<T>(xa, k1, k3, k4, k5, kj, hs, h, hnext, floor, xn, e);
            err = fmax(err, e);
            L.xn[ai] = xn;
        }
        T* out = L.cb;
        if constexpr (FAM == ASAT_FAMILY_TANH) {
            const T r = knorm * am;
            const bool ex = K >= 3 && K <= 10;
#pragma unroll 4
            for (int q = 0; q < K; q++) {
                const int l = lm[q];
                if (l == 0) break;
                T p = (T)1.0;
                bool first = true;
#pragma unroll 4
                for (int qq = 0; qq < K; qq++) {
                    if (qq == q) continue;
                    p = first ? f[qq] : p * f[qq];
                    first = false;
                }
                out[pm[q]] = tanh_term<T>(r, p, f[q], (T)a.q, l < 0, ex);
            }
        } else if (K <= 10) {
            // padding literals (zero padded at the end of the row) store
            // into the dummy list entry, which nothing reads
            const T r = km * knorm * am;
#pragma unroll 4
            for (int q = 0; q < K; q++) {
                const int l = lm[q];
                T p = (T)1.0;
                bool first = true;
#pragma unroll 4
                for (int qq = 0; qq < K; qq++) {
                    if (qq == q) continue;
                    p = first ? f[qq] : p * f[qq];
                    first = false;
                }
                T rr = r * p;
                rr *= l < 0 ? (T)-2.0 : (T)2.0;
                out[pm[q]] = rr;
            }
        } else {
            for (int q = 0; q < K; q++) {
                const int l = lm[q];
                if (l == 0) continue;
                T temp = km;
                for (int qq = 0; qq < K; qq++) temp *= (qq == q) ? (T)1.0 : f[qq];
                out[pm[q]] = (T)2.0 * temp * knorm * am * (l < 0 ? (T)-1.0 : (T)1.0);
            }
        }
    }
    return violated;
}

// clause phase of stage J: thread per clause; sp_in holds the stage's spins
// (x itself for J = 1).  count (J = 1): also count the violated clauses of x;
// the wave's count, in lane 0 (lane 0 takes part in every pass of its wave
// over the clauses), is returned.
template <typename T, int KT, int BS, int J, int FAM>
__device__ __forceinline__ int small_clause(const SmallArgs& a, const SmallLds<T>& L, const SmallReg<KT, BS>& R,
                                            const T* sp_in, T hs, double h, double hnext, double floor, T knorm,
                                            double& err, bool count = false)
{
    int cnt = 0;
    if constexpr (SmallReg<KT, BS>::on) {
#pragma unroll
        for (int u = 0; u < SMALL_RPT; u++) {
            const int j = (int)threadIdx.x + u * BS;
            if (j < a.m) {
                const bool v = small_clause_one<T, KT, J, FAM>(a, L, sp_in, R.lit[u], R.pos[u], j, hs, h, hnext,
                                                               floor, knorm, err, count);
                if (count) cnt += (int)__popcll(__ballot(v));
            }
        }
    } else {
        const int K = KT ? KT : a.k;
        for (int j = (int)threadIdx.x; j < a.m; j += BS) {
            const bool v = small_clause_one<T, KT, J, FAM>(a, L, sp_in, L.lit + j * K, L.spos + j * K, j, hs, h,
                                                           hnext, floor, knorm, err, count);
            if (count) cnt += (int)__popcll(__ballot(v));
        }
    }
    return cnt;
}

// var phase of stage J: thread per variable; fresh = false only on a retry's
// J = 1 (k1 kept from the step's first attempt)
template <typename T, int J>
__device__ __forceinline__ void small_var_one(const SmallLds<T>& L, int i, int e0, int e1, int runlen, bool fresh, T hs,
                                              double h, double hnext, double floor, double& err)
{
    // the stage combination's inputs, loaded ahead of the sum so that one
    // LDS latency covers both
    const T xi = L.x[i];
    const T c1 = L.k1[i];
    const T c2 = J > 2 ? L.k2[i] : (T)0;
    const T c3 = J > 3 ? L.k3[i] : (T)0;
    const T c4 = J > 4 ? L.k4[i] : (T)0;
    const T c5 = J > 5 ? L.k5[i] : (T)0;
    T kj = (T)0;
    if (fresh) kj = runlen ? sum_uniform(L.cb + e0, runlen) : sum_run((T)0, L.cb, e0, e1);
    if constexpr (J == 1) {
        if (fresh) L.k1[i] = kj;
        else kj = c1;
        L.sp[i] = stage_input<T, 2>(xi, kj, 0, 0, 0, 0, hs);
    } else if constexpr (J == 2) {
        L.k2[i] = kj;
        L.sp[i] = stage_input<T, 3>(xi, c1, kj, 0, 0, 0, hs);
    } else if constexpr (J == 3) {
        L.k3[i] = kj;
        L.sp[i] = stage_input<T, 4>(xi, c1, c2, kj, 0, 0, hs);
    } else if constexpr (J == 4) {
        L.k4[i] = kj;
        L.sp[i] = stage_input<T, 5>(xi, c1, c2, c3, kj, 0, hs);
    } else if constexpr (J == 5) {
        L.k5[i] = kj;
        L.sp[i] = stage_input<T, 6>(xi, c1, c2, c3, c4, kj, hs);
    } else {
        T xn;
        double e;
        final_combination<T>(xi, c1, c3, c4, c5, kj, hs, h, hnext, floor, xn, e);
        err = fmax(err, e);
        L.xn[i] = xn;
    }
}

template <typename T, int KT, int BS, int J>
__device__ __forceinline__ void small_var(const SmallArgs& a, const SmallLds<T>& L, const SmallReg<KT, BS>& R,
                                          bool fresh, T hs, double h, double hnext, double floor, double& err)
{
    if constexpr (SmallReg<KT, BS>::on) {
#pragma unroll
        for (int u = 0; u < SMALL_RPT; u++) {
            const int i = (int)threadIdx.x + 1 + u * BS;
            if (i <= a.n) small_var_one<T, J>(L, i, R.e0[u], R.e1[u], a.runlen, fresh, hs, h, hnext, floor, err);
        }
    } else {
        for (int i = (int)threadIdx.x + 1; i <= a.n; i += BS)
            small_var_one<T, J>(L, i, L.vptr[i], L.vptr[i + 1], a.runlen, fresh, hs, h, hnext, floor, err);
    }
}

// block maximum in every thread, the wave maxima taken by DPP
template <int BS>
__device__ __forceinline__ double small_block_max(double v, double* sh)
{
    v = wave_max_dpp(v);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = sh[0];
#pragma unroll
    for (int w = 1; w < BS / 64; w++) r = fmax(r, sh[w]);
    __syncthreads();
    return r;
}

template <typename T, int KT, int BS, int FAM>
__global__ __launch_bounds__(BS) void k_small(SmallArgs a)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char small_lds[];
    __shared__ double shd[BS / 64];
    __shared__ int shi[BS / 64];
    const int alloc = a.alloc, n = a.n, mk = a.m * (KT ? KT : a.k);
    SmallLds<T> L;
    T* base = (T*)small_lds;
    L.x = base;
    L.xn = base + alloc;
    L.k1 = base + 2 * alloc; L.k2 = base + 3 * alloc; L.k3 = base + 4 * alloc; L.k4 = base + 5 * alloc;
    L.k5 = base + 6 * alloc;
    L.sp = base + 7 * alloc;
    L.cb = L.sp + small_cb_offset(n);
    L.lit = (int*)(L.cb + a.nlist + 1);   // after the lists and their dummy entry
    L.spos = L.lit + mk;
    L.vptr = L.spos + mk;
    {
        const T* gx = (const T*)a.x;
        for (int i = (int)threadIdx.x; i < alloc; i += BS) {
            const T v = i == 0 ? (T)-1.0 : gx[i];   // slot 0: the padding literal's spin
            L.x[i] = v;
            L.xn[i] = v;
            L.k1[i] = L.k2[i] = L.k3[i] = L.k4[i] = L.k5[i] = (T)0;
        }
        for (int i = (int)threadIdx.x; i <= n; i += BS) L.sp[i] = i == 0 ? (T)-1.0 : (T)0;   // slot 0: padding spin
        for (int i = (int)threadIdx.x; i < mk; i += BS) { L.lit[i] = a.lit[i]; L.spos[i] = a.spos[i]; }
        for (int i = (int)threadIdx.x; i <= a.nlist; i += BS) L.cb[i] = (T)-0.0;   // uniform runs' padding
        for (int i = (int)threadIdx.x; i < n + 2; i += BS) L.vptr[i] = a.vptr[i];
    }
    SmallReg<KT, BS> R;
    if constexpr (SmallReg<KT, BS>::on) {
#pragma unroll
        for (int u = 0; u < SMALL_RPT; u++) {
            const int j = (int)threadIdx.x + u * BS;
#pragma unroll
            for (int q = 0; q < (KT > 0 ? KT : 1); q++) {
                R.lit[u][q] = j < a.m ? a.lit[j * KT + q] : 0;
                R.pos[u][q] = j < a.m ? a.spos[j * KT + q] : -1;
            }
            const int i = (int)threadIdx.x + 1 + u * BS;
            R.e0[u] = i <= n ? a.vptr[i] : 0;
            R.e1[u] = i <= n ? a.vptr[i + 1] : 0;
        }
    }
    // the control block, identical in every thread
    const Ctrl* c = a.ctrl;
    double h = c->h, hnext = c->hnext, time = c->time, hdid = c->hdid, errmax = c->errmax;
    const double eps = c->eps, floor = c->floor;
    long long steps = c->steps, attempts = c->attempts, rhscount = c->rhscount;
    int step_start = c->step_start, underflow = c->underflow;
    double* hlog = c->hlog;
    const long long hlog_len = c->hlog_len;
    SmallSolve ss = {0, 0, -1, 0, 0};
    if (a.solve) ss = *a.solve;
    __syncthreads();
    const T knorm = (T)a.knorm;
    long long launched = 0;
    bool done = underflow || (a.solve ? ss.result >= 0 : steps >= a.target);
    while (!done) {
        const T hs = (T)h;
        double err = 0.0;
        // When the previous accepted step ended a batch, this step's first
        // clause phase (which reads every clause's spins of x) also counts
        // the violated clauses of x; SAT is decided on that count before the
        // loop's limits are checked, as in SatSolver (SatSolver.h:313-331).
        // An attempt that ends the loop here leaves x, the clock and the
        // counters as they were.
        const bool check = a.solve && ss.check;
        int cnt = 0;
        if (step_start) cnt = small_clause<T, KT, BS, 1, FAM>(a, L, R, L.x, hs, h, hnext, floor, knorm, err, check);
        if (check && (threadIdx.x & 63) == 0) shi[threadIdx.x >> 6] = cnt;
        __syncthreads();
        if (check) {
            int v = 0;
#pragma unroll
            for (int w = 0; w < BS / 64; w++) v += shi[w];
            ss.check = 0;
            ss.last_viol = v;
            if (v == 0) {
                ss.result = SMALL_SAT;
                break;
            }
        }
        if (a.solve && ss.bleft == 0) {
            // top of the SatSolver loop: while (time < maxTime && steps < maxSteps)
            const double tc = a.time_float ? (double)(float)time : time;
            if (!(tc < a.max_time && ss.solver_steps < a.max_steps)) {
                ss.result = tc >= a.max_time ? SMALL_MAXTIME : SMALL_MAXITER;
                break;
            }
            ss.bleft = a.batch;
        }
        // pause between steps (k1 of the current x is not carried over)
        if (step_start && launched >= SMALL_ATTEMPT_CAP) break;
        small_var<T, KT, BS, 1>(a, L, R, step_start != 0, hs, h, hnext, floor, err);
        __syncthreads();
        small_clause<T, KT, BS, 2, FAM>(a, L, R, L.sp, hs, h, hnext, floor, knorm, err);
        __syncthreads();
        small_var<T, KT, BS, 2>(a, L, R, true, hs, h, hnext, floor, err);
        __syncthreads();
        small_clause<T, KT, BS, 3, FAM>(a, L, R, L.sp, hs, h, hnext, floor, knorm, err);
        __syncthreads();
        small_var<T, KT, BS, 3>(a, L, R, true, hs, h, hnext, floor, err);
        __syncthreads();
        small_clause<T, KT, BS, 4, FAM>(a, L, R, L.sp, hs, h, hnext, floor, knorm, err);
        __syncthreads();
        small_var<T, KT, BS, 4>(a, L, R, true, hs, h, hnext, floor, err);
        __syncthreads();
        small_clause<T, KT, BS, 5, FAM>(a, L, R, L.sp, hs, h, hnext, floor, knorm, err);
        __syncthreads();
        small_var<T, KT, BS, 5>(a, L, R, true, hs, h, hnext, floor, err);
        __syncthreads();
        small_clause<T, KT, BS, 6, FAM>(a, L, R, L.sp, hs, h, hnext, floor, knorm, err);
        __syncthreads();
        small_var<T, KT, BS, 6>(a, L, R, true, hs, h, hnext, floor, err);
        // block maximum (its barriers also order the x_new writes before any
        // read of the next attempt)
        const double emax = small_block_max<BS>(err, shd) / eps;
        launched++;
        errmax = emax;
        attempts += 1;
        rhscount += step_start ? 6 : 5;
        if (emax <= 1.0) {
            const double hn = emax > TabD::ERRCON ? TabD::SAFETY * h * pow_ctrl(emax, 5) : 5.0 * h;
            hdid = h;
            time += h;
            if (hlog && steps < hlog_len && threadIdx.x == 0) hlog[steps] = h;
            steps += 1;
            T* t = L.x; L.x = L.xn; L.xn = t;
            step_start = 1;
            hnext = hn;
            h = hn;
            if (a.solve) {
                if (--ss.bleft == 0) {
                    ss.solver_steps += a.batch;
                    ss.check = 1;
                }
            } else if (steps >= a.target) {
                done = true;
            }
        } else {
            const double htemp = TabD::SAFETY * h * pow_ctrl(emax, 4);
            const double hn = h >= 0.0 ? fmax(htemp, TabD::HFACTOR * h) : fmin(htemp, TabD::HFACTOR * h);
            h = hn;
            step_start = 0;
            if (time + hn == time) {
                underflow = 1;
                if (a.solve) ss.result = SMALL_UNDERFLOW;
                done = true;
            }
        }
    }
    __syncthreads();
    {
        T* gx = (T*)a.x;
        for (int i = (int)threadIdx.x; i < alloc; i += BS) gx[i] = L.x[i];
    }
    if (threadIdx.x == 0) {
        Ctrl* cw = a.ctrl;
        cw->h = h; cw->hnext = hnext; cw->time = time; cw->hdid = hdid; cw->errmax = errmax;
        cw->steps = steps; cw->attempts = attempts; cw->rhscount = rhscount;
        cw->step_start = step_start; cw->underflow = underflow; cw->active = 0;
        if (a.solve) *a.solve = ss;
    }
}

}  // namespace asat
