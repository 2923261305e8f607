// Batched ragged CTDS solver for MI355X: thousands of independent trajectories
// (each with its own instance, N, M, k, step size and early exit) in one
// persistent launch.  No reference counterpart: the reference runs one
// trajectory per process (Frontend/src/bench_runner.cpp:257-348 launches
// CudaSat*+CudaRungeKutta once per sample).
//
// Mapping: one workgroup owns one trajectory at a time and runs the whole
// Cash-Karp loop for it; RK stages are separated by workgroup barriers, not
// kernel launches.  Per stage:
//   clause phase: each thread takes clauses j = tid, tid+BS, ...; gathers the
//     stage spins from LDS, forms K_m, the aux stage input (x_a, k1..k5 from
//     the workgroup's scratch), k_J(a_m) = a K^2, and writes each literal's
//     contribution 2 c a K K_mi straight into var-major order (LDS, or the
//     workgroup's global scratch for instances too big for LDS)
//   var phase:    each thread takes variables i = tid, tid+BS, ...; sums its
//     contiguous run of contributions in slot order (the reference's
//     accumulation order), writes k_J(s_i) and the next stage's spin input
// For instances whose contributions do not fit in LDS, the clause-chunked
// mode keeps no contribution array: per chunk of clauses the clause phase
// leaves r_m in LDS, and each variable forms its terms of that chunk from
// 8-byte var-major records (the clause, the sign, the other literals) and
// adds them in the same order.
// Stage 6 forms x_new and |xerr/xscal|, a block FP64 max feeds the step
// controller (thread 0, FP64, correctly rounded pow), which accepts (parity
// flip) or rejects (retry with smaller h).  After each accepted step the
// violated-clause count of the thresholded state is taken; SAT ends the
// trajectory, and the minimum count is tracked (MAXSAT).
//
// Work queue: workgroups pull trajectory ids from an atomic counter (ids are
// ordered largest first); each launch advances every live trajectory by at
// most chunk_steps accepted steps, so a launch is bounded and a sequence of
// launches is captured once in a hipGraph.
//
// Per-slot packing (uint32): bits 19..31 variable (1-based, 0 = padding),
// bit 18 negated, bits 0..17 position of the slot's contribution in the
// instance's var-major order.  Hence N <= 8191 and L = K*M <= 262143 per
// instance; bigger instances belong to the plain solver.
#include "asat_common.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <vector>

namespace asat {

extern thread_local std::string g_last_error;

namespace batch {

constexpr int KMAX = 4;
constexpr int MAXN = 8191;
constexpr int MAXL = 262143;

enum Status { RUNNING = 0, SOLVED = 1, BUDGET = 2, UNDERFLOW_ = 3, MAXTIME = 4 };

struct Desc {                 // one instance / trajectory (device)
    long long code_off;       // into code[]
    long long row_off;        // into rowptr[]
    long long state_off;      // into xa[]/xb[] (N spins then M aux)
    long long vcode_off;      // into vcode[] (clause-chunked mode; -1: not built)
    int n, m, k, nslots;      // nslots = number of contributions (non-padding slots)
    float knorm;
    int pad_;
};

// Clause-chunked mode: one 64-bit record per contribution, in var-major
// order: bits 0..17 the clause, bit 18 the literal's sign, then the k - 1
// other literals of the clause in clause order, 14 bits each (variable << 1
// | negated; variable 0 = padding).
constexpr int VC_OTHER_SHIFT = 19;

struct TrajCtrl {             // persistent per-trajectory control (device)
    double t, h, hnext, hdid, errmax, sat_time;
    long long steps, attempts, rhscount;
    int parity, status, min_viol, viol, sat_step, pad_;
};

struct Params {
    const Desc* desc;
    const uint32_t* code;
    const int* rowptr;
    float* xa;
    float* xb;
    TrajCtrl* ctrl;
    float* scratch;          // per-workgroup: 5 * stride (k1..k5) [+ maxL contributions]
    long long scratch_stride;  // floats per k vector in scratch (>= max n+m)
    long long scratch_per_wg;  // floats per workgroup
    const int* order;        // trajectory ids, largest first
    int ntraj;
    int* queue;              // atomic work counter (reset by the host before a launch)
    double eps;
    double floor;            // constant term of xscal (1e-20 = reference TINY)
    double max_time;
    long long max_steps;
    int chunk_steps;
    int check_every;         // take the violation count every this many accepted steps
    int sp_stride;           // floats reserved for the stage spins at the start of LDS
    const unsigned long long* vcode;   // clause-chunked mode: var-major contribution records
    int chunk_clauses;       // clause-chunked mode: clauses per chunk (LDS for their r_m)
};

// ---------------------------------------------------------------------------
// MODE 0: contributions in the workgroup's global scratch; 1: in LDS; 3:
// clause-chunked, no contribution array: per chunk of clauses, the clause
// phase leaves r_m = K_m knorm a_m in LDS and every variable walks on
// through its contribution records of that chunk, forming each term
// r_m * (product of the other factors) * (+-2) and adding it in the same
// order as the other modes (LDS: spins, running sums, record cursors, r_m).
template <int BS, int MODE>
__global__ __launch_bounds__(BS) void k_batch(Params P)
{
    extern __shared__ float lds[];          // [stage spins: sp_stride] then contributions
    __shared__ double red[BS / 64];
    __shared__ int redi[BS / 64];
    __shared__ int s_traj;

    float* sp = lds;                        // stage spin input, index i-1 for variable i
    float* scratch = P.scratch + (long long)blockIdx.x * P.scratch_per_wg;
    float* kbuf1 = scratch;
    float* kbuf2 = scratch + P.scratch_stride;
    float* kbuf3 = scratch + 2 * P.scratch_stride;
    float* kbuf4 = scratch + 3 * P.scratch_stride;
    float* kbuf5 = scratch + 4 * P.scratch_stride;
    const int tid = threadIdx.x;

    for (;;) {
        if (tid == 0) s_traj = atomicAdd(P.queue, 1);
        __syncthreads();
        // the trajectory id is the same in every lane: reading it as a
        // scalar lets the descriptor and the instance's base pointers live
        // in scalar registers, leaving the vector registers to the state
        const int slot = __builtin_amdgcn_readfirstlane(s_traj);
        __syncthreads();
        if (slot >= P.ntraj) break;
        const int tr = __builtin_amdgcn_readfirstlane(P.order[slot]);
        const Desc d = P.desc[tr];
        TrajCtrl* cp = P.ctrl + tr;
        if (cp->status != RUNNING) continue;    // uniform: every thread reads the same word

        const int n = d.n, m = d.m, k = d.k;
        const uint32_t* code = P.code + d.code_off;
        const int* rowptr = P.rowptr + d.row_off;
        float* contrib = MODE == 1 ? (lds + P.sp_stride) : (scratch + 5 * P.scratch_stride);
        float* vsum = lds + P.sp_stride;                          // MODE 3
        int* vcur = (int*)(lds + 2 * P.sp_stride);                // MODE 3
        float* rl = lds + 3 * P.sp_stride;                        // MODE 3
        const unsigned long long* vcode = MODE == 3 ? P.vcode + d.vcode_off : nullptr;
        const int chunk = MODE == 3 ? P.chunk_clauses : m;
        const float knorm = d.knorm;
        // scratch layout per k vector: spins at [0, n), aux at [n, n+m)
        double t = cp->t, h = cp->h, hnext = cp->hnext, hdid = cp->hdid, errmax_last = cp->errmax;
        long long steps = cp->steps, attempts = cp->attempts, rhscount = cp->rhscount;
        int parity = cp->parity, min_viol = cp->min_viol, viol = cp->viol;
        int status = RUNNING;
        int sat_step = cp->sat_step;
        double sat_time = cp->sat_time;
        int done_in_chunk = 0;

        while (status == RUNNING && done_in_chunk < P.chunk_steps) {
            float* x = (parity ? P.xb : P.xa) + d.state_off;
            float* xn = (parity ? P.xa : P.xb) + d.state_off;
            // ---- stage 1 at x: spins into LDS
            for (int i = tid; i < n; i += BS) sp[i] = x[i];
            __syncthreads();
            bool accepted = false;
            bool first_attempt = true;
            while (!accepted) {
                const float hs = (float)h;
                const bool had_stage1 = first_attempt;
                double err_local = 0.0;
#pragma unroll 1
                for (int J = first_attempt ? 1 : 2; J <= 6; J++) {
                  for (int c0 = 0; c0 < m; c0 += chunk) {
                    const int c1 = min(m, c0 + chunk);
                    // ---------------- clause phase (MODE 3: one chunk of clauses)
                    for (int j = c0 + tid; j < c1; j += BS) {
                        const uint32_t* cj = code + (long long)j * k;
                        float f[KMAX];
                        uint32_t cc[KMAX];
                        float km = knorm;
#pragma unroll
                        for (int q = 0; q < KMAX; q++) {
                            f[q] = 1.0f;
                            cc[q] = 0;
                            if (q < k) {
                                uint32_t c = cj[q];
                                cc[q] = c;
                                int v = (int)(c >> 19);
                                float s = v ? sp[v - 1] : -1.0f;
                                f[q] = (c >> 18) & 1u ? 1.0f + s : 1.0f - s;
                                km *= f[q];
                            }
                        }
                        const int ai = n + j;
                        const float xa = x[ai];
                        float am, k1 = 0.f, k3 = 0.f, k4 = 0.f, k5 = 0.f;
                        if (J == 1) {
                            am = xa;
                        } else {
                            k1 = kbuf1[ai];
                            float k2 = J > 2 ? kbuf2[ai] : 0.f;
                            k3 = J > 3 ? kbuf3[ai] : 0.f;
                            k4 = J > 4 ? kbuf4[ai] : 0.f;
                            k5 = J > 5 ? kbuf5[ai] : 0.f;
                            switch (J) {
                            case 2: am = stage_input<float, 2>(xa, k1, k2, k3, k4, k5, hs); break;
                            case 3: am = stage_input<float, 3>(xa, k1, k2, k3, k4, k5, hs); break;
                            case 4: am = stage_input<float, 4>(xa, k1, k2, k3, k4, k5, hs); break;
                            case 5: am = stage_input<float, 5>(xa, k1, k2, k3, k4, k5, hs); break;
                            default: am = stage_input<float, 6>(xa, k1, k2, k3, k4, k5, hs); break;
                            }
                        }
                        const float kj = am * km * km;
                        switch (J) {
                        case 1: kbuf1[ai] = kj; break;
                        case 2: kbuf2[ai] = kj; break;
                        case 3: kbuf3[ai] = kj; break;
                        case 4: kbuf4[ai] = kj; break;
                        case 5: kbuf5[ai] = kj; break;
                        default: {
                            float xv;
                            double e;
                            final_combination<float>(xa, k1, k3, k4, k5, kj, hs, h, hnext, P.floor, xv, e);
                            xn[ai] = xv;
                            err_local = fmax(err_local, e);
                        }
                        }
                        const float r = km * knorm * am;
                        if constexpr (MODE == 3) {
                            rl[j - c0] = r;
                        } else {
#pragma unroll
                            for (int q = 0; q < KMAX; q++) {
                                if (q < k && (cc[q] >> 19) != 0) {
                                    float p = 1.0f;
                                    bool first = true;
#pragma unroll
                                    for (int qq = 0; qq < KMAX; qq++) {
                                        if (qq == q || qq >= k) continue;
                                        p = first ? f[qq] : p * f[qq];
                                        first = false;
                                    }
                                    float rr = r * p;
                                    rr *= (cc[q] >> 18) & 1u ? -2.0f : 2.0f;
                                    contrib[cc[q] & 0x3FFFFu] = rr;
                                }
                            }
                        }
                    }
                    __syncthreads();
                    if constexpr (MODE == 3) {
                        // each variable's terms of this chunk, in record
                        // (clause) order: the product of the other factors
                        // as the clause phase forms it (KmiProd order)
                        for (int i = tid; i < n; i += BS) {
                            int e = c0 == 0 ? rowptr[i] : vcur[i];
                            const int e1 = rowptr[i + 1];
                            float sum = c0 == 0 ? 0.0f : vsum[i];
                            bool more = e < e1;
                            while (more) {
                                // four records read ahead of the chunk-end test
                                unsigned long long vb[4];
#pragma unroll
                                for (int u = 0; u < 4; u++) vb[u] = e + u < e1 ? vcode[e + u] : ~0ull;
#pragma unroll
                                for (int u = 0; u < 4; u++) {
                                    if (!more) break;
                                    const unsigned long long vc = vb[u];
                                    const int j = (int)(vc & 0x3FFFFull);
                                    if (e >= e1 || j >= c1) { more = false; break; }
                                    float p = 1.0f;
                                    bool first = true;
                                    for (int t = 0; t < k - 1; t++) {
                                        const unsigned o = (unsigned)(vc >> (VC_OTHER_SHIFT + 14 * t)) & 0x3FFFu;
                                        const int v = (int)(o >> 1);
                                        const float sv = v ? sp[v - 1] : -1.0f;
                                        const float fo = (o & 1u) ? 1.0f + sv : 1.0f - sv;
                                        p = first ? fo : p * fo;
                                        first = false;
                                    }
                                    float rr = rl[j - c0] * p;
                                    rr *= (vc >> 18) & 1ull ? -2.0f : 2.0f;
                                    sum += rr;
                                    e++;
                                }
                                more = more && e < e1;
                            }
                            vsum[i] = sum;
                            vcur[i] = e;
                        }
                        __syncthreads();
                    }
                  }
                    // ---------------- var phase
                    for (int i = tid; i < n; i += BS) {
                        float sum = 0.0f;
                        if constexpr (MODE == 3) {
                            sum = vsum[i];
                        } else {
                            const int e0 = rowptr[i], e1 = rowptr[i + 1];
                            sum = sum_run(sum, contrib, e0, e1);
                        }
                        const float kj = sum;
                        const float xi = x[i];
                        switch (J) {
                        case 1:
                            kbuf1[i] = kj;
                            sp[i] = stage_input<float, 2>(xi, kj, 0.f, 0.f, 0.f, 0.f, hs);
                            break;
                        case 2:
                            kbuf2[i] = kj;
                            sp[i] = stage_input<float, 3>(xi, kbuf1[i], kj, 0.f, 0.f, 0.f, hs);
                            break;
                        case 3:
                            kbuf3[i] = kj;
                            sp[i] = stage_input<float, 4>(xi, kbuf1[i], kbuf2[i], kj, 0.f, 0.f, hs);
                            break;
                        case 4:
                            kbuf4[i] = kj;
                            sp[i] = stage_input<float, 5>(xi, kbuf1[i], kbuf2[i], kbuf3[i], kj, 0.f, hs);
                            break;
                        case 5:
                            kbuf5[i] = kj;
                            sp[i] = stage_input<float, 6>(xi, kbuf1[i], kbuf2[i], kbuf3[i], kbuf4[i], kj, hs);
                            break;
                        default: {
                            float xv;
                            double e;
                            final_combination<float>(xi, kbuf1[i], kbuf3[i], kbuf4[i], kbuf5[i], kj, hs, h, hnext,
                                                     P.floor, xv, e);
                            xn[i] = xv;
                            err_local = fmax(err_local, e);
                        }
                        }
                    }
                    __syncthreads();
                    if (J == 1) {
                        // a retry restarts at stage 2: its spin input needs the new h
                        first_attempt = false;
                    }
                }
                // ---------------- error norm + control (FP64), evaluated
                // identically by every thread so h, hnext, t stay uniform
                const double errmax = block_max_all<BS>(err_local, red) / P.eps;
                errmax_last = errmax;
                attempts += 1;
                rhscount += had_stage1 ? 6 : 5;
                if (errmax <= 1.0) {
                    double hn = errmax > TabD::ERRCON ? TabD::SAFETY * h * pow_ctrl(errmax, 5) : 5.0 * h;
                    hdid = h;
                    t += h;
                    steps += 1;
                    hnext = hn;
                    h = hn;
                    accepted = true;
                } else {
                    double htemp = TabD::SAFETY * h * pow_ctrl(errmax, 4);
                    double hn = h >= 0.0 ? fmax(htemp, TabD::HFACTOR * h) : fmin(htemp, TabD::HFACTOR * h);
                    if (t + hn == t) {
                        h = hn;
                        status = UNDERFLOW_;
                        break;
                    }
                    h = hn;
                    // retry: stage-2 spin input from x, k1 and the reduced h
                    const float hs2 = (float)h;
                    for (int i = tid; i < n; i += BS)
                        sp[i] = stage_input<float, 2>(x[i], kbuf1[i], 0.f, 0.f, 0.f, 0.f, hs2);
                    __syncthreads();
                }
            }
            if (status != RUNNING) break;
            parity ^= 1;
            done_in_chunk += 1;
            // ---------------- violated-clause count on the accepted state
            if (steps % P.check_every == 0 || steps >= P.max_steps) {
                const float* xs = (parity ? P.xb : P.xa) + d.state_off;
                int cnt = 0;
                for (int j = tid; j < m; j += BS) {
                    const uint32_t* cj = code + (long long)j * k;
                    int sat = 0;
                    for (int q = 0; q < k && !sat; q++) {
                        uint32_t c = cj[q];
                        int v = (int)(c >> 19);
                        if (!v) continue;
                        int tv = xs[v - 1] > 0.0f;
                        sat = (c >> 18) & 1u ? !tv : tv;
                    }
                    cnt += !sat;
                }
                viol = block_sum_all<BS>(cnt, redi);
                if (viol < min_viol) min_viol = viol;
                if (viol == 0) {
                    status = SOLVED;
                    sat_step = (int)steps;
                    sat_time = t;
                }
            }
            if (status == RUNNING && steps >= P.max_steps) status = BUDGET;
            if (status == RUNNING && t >= P.max_time) status = MAXTIME;
            __syncthreads();
        }
        if (tid == 0) {
            cp->t = t; cp->h = h; cp->hnext = hnext; cp->hdid = hdid; cp->errmax = errmax_last;
            cp->steps = steps; cp->attempts = attempts; cp->rhscount = rhscount;
            cp->parity = parity; cp->status = status; cp->min_viol = min_viol; cp->viol = viol;
            cp->sat_step = sat_step; cp->sat_time = sat_time;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Register-resident variant: thread tid owns clauses j = tid + u*BS (u < CPT)
// and variables i = tid + v*BS (v < VPT) for the whole trajectory; their
// state x and the stage derivatives k1..k5 live in registers, so a step
// touches global memory only for the literal codes and row pointers (a few
// hundred KB per trajectory, re-read every stage from L2).  The arithmetic
// and its order are those of k_batch; x_new of the final stage goes to the
// k2 registers, which the final combination does not read.
template <int CPT, int VPT>
struct RegState {
    float xa[CPT], a1[CPT], a2[CPT], a3[CPT], a4[CPT], a5[CPT];
    float xv[VPT], v1[VPT], v2[VPT], v3[VPT], v4[VPT], v5[VPT];
    int e0[VPT], e1[VPT];
};

// KK = 3 or 4: every instance of the launch has clauses of KK literal slots
// (the code is compiled for that length); KK = 0: the length k is read at
// run time (at most KMAX)
template <int BS, int CPT, int VPT, int J, int KK>
__device__ __forceinline__ void reg_stage(RegState<CPT, VPT>& R, const uint32_t* code, int n, int m, int k,
                                          float knorm, float* sp, float* contrib, float hs, double h, double hnext,
                                          double floor, double& err)
{
    constexpr int KQ = KK ? KK : KMAX;
    if constexpr (KK != 0) k = KK;
    const int tid = threadIdx.x;
    // ---------------- clause phase
    // the next clause's codes are loaded while this one is processed
    uint32_t nx[KQ];
#pragma unroll
    for (int q = 0; q < KQ; q++) nx[q] = (tid < m && (KK != 0 || q < k)) ? code[(long long)tid * k + q] : 0u;
#pragma unroll
    for (int u = 0; u < CPT; u++) {
        const int j = tid + u * BS;
        uint32_t cur[KQ];
#pragma unroll
        for (int q = 0; q < KQ; q++) cur[q] = nx[q];
        if (u + 1 < CPT) {
            const int jn = j + BS;
#pragma unroll
            for (int q = 0; q < KQ; q++) nx[q] = (jn < m && (KK != 0 || q < k)) ? code[(long long)jn * k + q] : 0u;
        }
        if (j < m) {
            float f[KQ];
            uint32_t cc[KQ];
            float km = knorm;
#pragma unroll
            for (int q = 0; q < KQ; q++) {
                f[q] = 1.0f;
                cc[q] = 0;
                if (KK != 0 || q < k) {
                    uint32_t c = cur[q];
                    cc[q] = c;
                    int v = (int)(c >> 19);
                    float s = v ? sp[v - 1] : -1.0f;
                    f[q] = (c >> 18) & 1u ? 1.0f + s : 1.0f - s;
                    km *= f[q];
                }
            }
            float am;
            if constexpr (J == 1) am = R.xa[u];
            else am = stage_input<float, J>(R.xa[u], R.a1[u], R.a2[u], R.a3[u], R.a4[u], R.a5[u], hs);
            const float kj = am * km * km;
            if constexpr (J == 1) R.a1[u] = kj;
            else if constexpr (J == 2) R.a2[u] = kj;
            else if constexpr (J == 3) R.a3[u] = kj;
            else if constexpr (J == 4) R.a4[u] = kj;
            else if constexpr (J == 5) R.a5[u] = kj;
            else {
                float xv;
                double e;
                final_combination<float>(R.xa[u], R.a1[u], R.a3[u], R.a4[u], R.a5[u], kj, hs, h, hnext, floor, xv, e);
                R.a2[u] = xv;
                err = fmax(err, e);
            }
            const float r = km * knorm * am;
#pragma unroll
            for (int q = 0; q < KQ; q++) {
                if ((KK != 0 || q < k) && (cc[q] >> 19) != 0) {
                    float p = 1.0f;
                    bool first = true;
#pragma unroll
                    for (int qq = 0; qq < KQ; qq++) {
                        if (qq == q || (KK == 0 && qq >= k)) continue;
                        p = first ? f[qq] : p * f[qq];
                        first = false;
                    }
                    float rr = r * p;
                    rr *= (cc[q] >> 18) & 1u ? -2.0f : 2.0f;
                    contrib[cc[q] & 0x3FFFFu] = rr;
                }
            }
        }
        // one clause at a time: keeps the live registers to the state plus
        // one clause's temporaries
        __builtin_amdgcn_sched_barrier(0);
    }
    __syncthreads();
    // ---------------- var phase
#pragma unroll
    for (int v = 0; v < VPT; v++) {
        const int i = tid + v * BS;
        if (i < n) {
            float sum = 0.0f;
            sum = sum_run(sum, contrib, R.e0[v], R.e1[v]);
            const float kj = sum;
            const float xi = R.xv[v];
            if constexpr (J == 1) {
                R.v1[v] = kj;
                sp[i] = stage_input<float, 2>(xi, kj, 0.f, 0.f, 0.f, 0.f, hs);
            } else if constexpr (J == 2) {
                R.v2[v] = kj;
                sp[i] = stage_input<float, 3>(xi, R.v1[v], kj, 0.f, 0.f, 0.f, hs);
            } else if constexpr (J == 3) {
                R.v3[v] = kj;
                sp[i] = stage_input<float, 4>(xi, R.v1[v], R.v2[v], kj, 0.f, 0.f, hs);
            } else if constexpr (J == 4) {
                R.v4[v] = kj;
                sp[i] = stage_input<float, 5>(xi, R.v1[v], R.v2[v], R.v3[v], kj, 0.f, hs);
            } else if constexpr (J == 5) {
                R.v5[v] = kj;
                sp[i] = stage_input<float, 6>(xi, R.v1[v], R.v2[v], R.v3[v], R.v4[v], kj, hs);
            } else {
                float xn;
                double e;
                final_combination<float>(xi, R.v1[v], R.v3[v], R.v4[v], R.v5[v], kj, hs, h, hnext, floor, xn, e);
                R.v2[v] = xn;
                err = fmax(err, e);
            }
        }
    }
    __syncthreads();
}

template <int BS, int CPT, int VPT, int KK>
__global__ __launch_bounds__(BS) void k_batch_reg(Params P)
{
    extern __shared__ float lds[];          // [stage spins: sp_stride] then contributions
    __shared__ double red[BS / 64];
    __shared__ int redi[BS / 64];
    __shared__ int s_traj;
    float* sp = lds;
    float* contrib = lds + P.sp_stride;
    const int tid = threadIdx.x;

    for (;;) {
        if (tid == 0) s_traj = atomicAdd(P.queue, 1);
        __syncthreads();
        // the trajectory id is the same in every lane: reading it as a
        // scalar lets the descriptor and the instance's base pointers live
        // in scalar registers, leaving the vector registers to the state
        const int slot = __builtin_amdgcn_readfirstlane(s_traj);
        __syncthreads();
        if (slot >= P.ntraj) break;
        const int tr = __builtin_amdgcn_readfirstlane(P.order[slot]);
        const Desc d = P.desc[tr];
        TrajCtrl* cp = P.ctrl + tr;
        if (cp->status != RUNNING) continue;    // uniform: every thread reads the same word

        const int n = d.n, m = d.m, k = d.k;
        const uint32_t* code = P.code + d.code_off;
        const int* rowptr = P.rowptr + d.row_off;
        const float knorm = d.knorm;
        double t = cp->t, h = cp->h, hnext = cp->hnext, hdid = cp->hdid, errmax_last = cp->errmax;
        long long steps = cp->steps, attempts = cp->attempts, rhscount = cp->rhscount;
        int parity = cp->parity, min_viol = cp->min_viol, viol = cp->viol;
        int status = RUNNING;
        int sat_step = cp->sat_step;
        double sat_time = cp->sat_time;
        int done_in_chunk = 0;

        RegState<CPT, VPT> R;
        {
            const float* x = (parity ? P.xb : P.xa) + d.state_off;
#pragma unroll
            for (int u = 0; u < CPT; u++) {
                const int j = tid + u * BS;
                R.xa[u] = j < m ? x[n + j] : 0.f;
                R.a1[u] = R.a2[u] = R.a3[u] = R.a4[u] = R.a5[u] = 0.f;
            }
#pragma unroll
            for (int v = 0; v < VPT; v++) {
                const int i = tid + v * BS;
                R.xv[v] = i < n ? x[i] : 0.f;
                R.e0[v] = i < n ? rowptr[i] : 0;
                R.e1[v] = i < n ? rowptr[i + 1] : 0;
                R.v1[v] = R.v2[v] = R.v3[v] = R.v4[v] = R.v5[v] = 0.f;
            }
        }

        while (status == RUNNING && done_in_chunk < P.chunk_steps) {
            // ---- stage 1 at x: spins into LDS
#pragma unroll
            for (int v = 0; v < VPT; v++) {
                const int i = tid + v * BS;
                if (i < n) sp[i] = R.xv[v];
            }
            __syncthreads();
            bool accepted = false;
            bool first_attempt = true;
            while (!accepted) {
                const float hs = (float)h;
                const bool had_stage1 = first_attempt;
                double err_local = 0.0;
                if (first_attempt)
                    reg_stage<BS, CPT, VPT, 1, KK>(R, code, n, m, k, knorm, sp, contrib, hs, h, hnext, P.floor, err_local);
                first_attempt = false;
                reg_stage<BS, CPT, VPT, 2, KK>(R, code, n, m, k, knorm, sp, contrib, hs, h, hnext, P.floor, err_local);
                reg_stage<BS, CPT, VPT, 3, KK>(R, code, n, m, k, knorm, sp, contrib, hs, h, hnext, P.floor, err_local);
                reg_stage<BS, CPT, VPT, 4, KK>(R, code, n, m, k, knorm, sp, contrib, hs, h, hnext, P.floor, err_local);
                reg_stage<BS, CPT, VPT, 5, KK>(R, code, n, m, k, knorm, sp, contrib, hs, h, hnext, P.floor, err_local);
                reg_stage<BS, CPT, VPT, 6, KK>(R, code, n, m, k, knorm, sp, contrib, hs, h, hnext, P.floor, err_local);
                // ---------------- error norm + control (FP64), uniform in every thread
                const double errmax = block_max_all<BS>(err_local, red) / P.eps;
                errmax_last = errmax;
                attempts += 1;
                rhscount += had_stage1 ? 6 : 5;
                if (errmax <= 1.0) {
                    double hn = errmax > TabD::ERRCON ? TabD::SAFETY * h * pow_ctrl(errmax, 5) : 5.0 * h;
                    hdid = h;
                    t += h;
                    steps += 1;
                    hnext = hn;
                    h = hn;
                    accepted = true;
                } else {
                    double htemp = TabD::SAFETY * h * pow_ctrl(errmax, 4);
                    double hn = h >= 0.0 ? fmax(htemp, TabD::HFACTOR * h) : fmin(htemp, TabD::HFACTOR * h);
                    if (t + hn == t) {
                        h = hn;
                        status = UNDERFLOW_;
                        break;
                    }
                    h = hn;
                    // retry: stage-2 spin input from x, k1 and the reduced h
                    const float hs2 = (float)h;
#pragma unroll
                    for (int v = 0; v < VPT; v++) {
                        const int i = tid + v * BS;
                        if (i < n) sp[i] = stage_input<float, 2>(R.xv[v], R.v1[v], 0.f, 0.f, 0.f, 0.f, hs2);
                    }
                    __syncthreads();
                }
            }
            if (status != RUNNING) break;
            // accept: x <- x_new (held in the k2 registers)
#pragma unroll
            for (int u = 0; u < CPT; u++) R.xa[u] = R.a2[u];
#pragma unroll
            for (int v = 0; v < VPT; v++) R.xv[v] = R.v2[v];
            parity ^= 1;
            done_in_chunk += 1;
            // ---------------- violated-clause count on the accepted state
            if (steps % P.check_every == 0 || steps >= P.max_steps) {
#pragma unroll
                for (int v = 0; v < VPT; v++) {
                    const int i = tid + v * BS;
                    if (i < n) sp[i] = R.xv[v];
                }
                __syncthreads();
                int cnt = 0;
#pragma unroll
                for (int u = 0; u < CPT; u++) {
                    const int j = tid + u * BS;
                    if (j >= m) continue;
                    const uint32_t* cj = code + (long long)j * k;
                    int sat = 0;
                    for (int q = 0; q < k && !sat; q++) {
                        uint32_t c = cj[q];
                        int v = (int)(c >> 19);
                        if (!v) continue;
                        int tv = sp[v - 1] > 0.0f;
                        sat = (c >> 18) & 1u ? !tv : tv;
                    }
                    cnt += !sat;
                }
                viol = block_sum_all<BS>(cnt, redi);
                if (viol < min_viol) min_viol = viol;
                if (viol == 0) {
                    status = SOLVED;
                    sat_step = (int)steps;
                    sat_time = t;
                }
            }
            if (status == RUNNING && steps >= P.max_steps) status = BUDGET;
            if (status == RUNNING && t >= P.max_time) status = MAXTIME;
            __syncthreads();
        }
        // the state after the last accepted step goes to the parity buffer it names
        {
            float* x = (parity ? P.xb : P.xa) + d.state_off;
#pragma unroll
            for (int u = 0; u < CPT; u++) {
                const int j = tid + u * BS;
                if (j < m) x[n + j] = R.xa[u];
            }
#pragma unroll
            for (int v = 0; v < VPT; v++) {
                const int i = tid + v * BS;
                if (i < n) x[i] = R.xv[v];
            }
        }
        if (tid == 0) {
            cp->t = t; cp->h = h; cp->hnext = hnext; cp->hdid = hdid; cp->errmax = errmax_last;
            cp->steps = steps; cp->attempts = attempts; cp->rhscount = rhscount;
            cp->parity = parity; cp->status = status; cp->min_viol = min_viol; cp->viol = viol;
            cp->sat_step = sat_step; cp->sat_time = sat_time;
        }
        __syncthreads();
    }
}

__global__ void k_reset_queue(int* q)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *q = 0;
}

static inline long long round_up_ll(long long a, long long b) { return (a + b - 1) / b * b; }

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct Instance {
    int n = 0, m = 0, k = 0, nslots = 0;
    std::vector<uint32_t> code;   // m*k
    std::vector<int> rowptr;      // n+1
    std::vector<float> pub0;      // n+m
};

// contribution records of the clause-chunked mode (see VC_OTHER_SHIFT), in
// var-major (contribution) order
static void build_vcode(const Instance& I, unsigned long long* out)
{
    const int k = I.k;
    for (int j = 0; j < I.m; j++) {
        const uint32_t* cj = I.code.data() + (size_t)j * k;
        for (int q = 0; q < k; q++) {
            const uint32_t c = cj[q];
            if ((c >> 19) == 0) continue;
            unsigned long long vc = (unsigned long long)j | ((unsigned long long)((c >> 18) & 1u) << 18);
            int t = 0;
            for (int qq = 0; qq < k; qq++) {
                if (qq == q) continue;
                const uint32_t o = ((cj[qq] >> 19) << 1) | ((cj[qq] >> 18) & 1u);
                vc |= (unsigned long long)o << (VC_OTHER_SHIFT + 14 * t);
                t++;
            }
            out[c & 0x3FFFFu] = vc;
        }
    }
}

struct Bucket {
    std::vector<int> ids;          // trajectory ids, largest first
    int block = 0;                 // workgroup size (256/512/1024), 0 = autotune
    int lds = -1;                  // contributions in LDS (1) or global scratch (0); 2: register-
                                   // resident kernel (contributions in LDS); 3: clause-chunked
                                   // (no contribution array); -1 = auto
    int chunk = 0;                 // clause-chunked mode: clauses per chunk
    bool vcode = false;            // clause-chunked records built for this bucket's instances
    int cpt = 0, vpt = 0;          // register kernel: clauses / variables per thread
    int kk = 0;                    // register kernel: 3 when every instance has 3-literal clauses, else 0
    int grid = 0;
    int maxn = 0, maxm = 0, maxl = 0;
    int sp_stride = 0;
    long long scratch_stride = 0, scratch_per_wg = 0;
    size_t lds_bytes = 0;
    bool lds_fits = false;
    int* d_order = nullptr;
    int* d_queue = nullptr;
    float* d_scratch = nullptr;
    double tuned_us = -1.0;        // autotune: microseconds per trajectory-attempt
    void free_dev()
    {
        (void)hipFree(d_order); (void)hipFree(d_queue); (void)hipFree(d_scratch);
        d_order = nullptr; d_queue = nullptr; d_scratch = nullptr;
    }
};

struct Batch {
    std::vector<Instance> inst;
    double eps = 1e-6;
    double floor = 1e-20;
    double max_time = 1e300;
    long long max_steps = 10000;
    int chunk_steps = 256;
    int check_every = 1;
    int force_block = 0;           // 0: autotune per bucket
    int force_lds = -1;            // -1: auto
    int force_chunk = 0;           // clause-chunked mode: clauses per chunk, 0 = as large as LDS allows
    int nbuckets_req = 4;
    bool prepared = false;
    std::vector<Bucket> buckets;
    // device
    Desc* d_desc = nullptr;
    uint32_t* d_code = nullptr;
    unsigned long long* d_vcode = nullptr;
    int* d_rowptr = nullptr;
    float* d_xa = nullptr;
    float* d_xb = nullptr;
    TrajCtrl* d_ctrl = nullptr;
    std::vector<TrajCtrl> h_ctrl, h_ctrl0;
    std::vector<float> x0;
    std::vector<long long> state_off;
    hipStream_t stream = nullptr;
    int cus = 0;
    size_t lds_limit = 0;
    // results
    std::vector<double> solved_wall_ms;
    double run_ms = 0.0;
    int chunks_run = 0;

    void free_dev()
    {
        (void)hipFree(d_desc); (void)hipFree(d_code); (void)hipFree(d_rowptr); (void)hipFree(d_xa);
        (void)hipFree(d_xb); (void)hipFree(d_ctrl); (void)hipFree(d_vcode);
        d_desc = nullptr; d_code = nullptr; d_rowptr = nullptr; d_xa = d_xb = nullptr; d_ctrl = nullptr;
        d_vcode = nullptr;
        for (auto& b : buckets) b.free_dev();
    }
    ~Batch()
    {
        free_dev();
        if (stream) (void)hipStreamDestroy(stream);
    }
};

// build one instance: SatProblem::AddClauses semantics, clause matrix padded to
// K, contributions numbered in var-major slot order
static void build_instance(const int* lits, const int* lens, int nclauses, Instance& I)
{
    require(nclauses > 0 && lits && lens, ASAT_ERR_INVALID_ARGUMENT, "empty instance");
    int n = 0, m = 0, k = 0;
    size_t off = 0;
    for (int j = 0; j < nclauses; j++) {
        require(lens[j] >= 0, ASAT_ERR_INVALID_ARGUMENT, "negative clause length");
        int len1 = 0;
        for (int i = 0; i < lens[j]; i++) {
            int l = lits[off + i];
            if (!l) continue;
            require(l != INT32_MIN, ASAT_ERR_INVALID_ARGUMENT, "literal out of range");
            n = std::max(n, std::abs(l));
            len1++;
        }
        if (len1) { m++; k = std::max(k, len1); }
        off += (size_t)lens[j];
    }
    require(m > 0, ASAT_ERR_INVALID_ARGUMENT, "empty instance");
    require(k <= KMAX, ASAT_ERR_UNSUPPORTED, "batched solver supports clauses of up to 4 literals");
    require(n <= MAXN, ASAT_ERR_UNSUPPORTED, "batched solver supports up to 8191 variables per instance");
    require((long long)m * k <= MAXL, ASAT_ERR_UNSUPPORTED, "batched solver supports up to 262143 literal slots");
    I.n = n; I.m = m; I.k = k;
    std::vector<int> C((size_t)m * k, 0);
    off = 0;
    int row = 0;
    for (int j = 0; j < nclauses; j++) {
        int q = 0;
        for (int i = 0; i < lens[j]; i++) {
            int l = lits[off + i];
            if (l) C[(size_t)row * k + q++] = l;
        }
        if (q) row++;
        off += (size_t)lens[j];
    }
    // rowptr[i-1] .. rowptr[i]: contributions of variable i, in slot order
    std::vector<int> cnt((size_t)n + 1, 0);
    for (int l : C)
        if (l) cnt[std::abs(l)]++;
    I.rowptr.assign((size_t)n + 1, 0);
    for (int i = 1; i <= n; i++) I.rowptr[i] = I.rowptr[i - 1] + cnt[i];
    I.nslots = I.rowptr[n];
    std::vector<int> fill(I.rowptr.begin(), I.rowptr.end() - 1);
    I.code.assign((size_t)m * k, 0u);
    for (size_t s = 0; s < C.size(); s++) {
        int l = C[s];
        if (!l) continue;
        int v = std::abs(l);
        uint32_t p = (uint32_t)fill[v - 1]++;
        I.code[s] = ((uint32_t)v << 19) | ((l < 0 ? 1u : 0u) << 18) | p;
    }
}

static size_t bucket_lds(const Bucket& b, bool lds)
{
    return sizeof(float) * ((size_t)b.sp_stride + (lds ? (size_t)round_up_ll(b.maxl, 4) : 0));
}

// clause-chunked mode: spins, running sums and record cursors per variable,
// then r_m of one chunk of clauses
static size_t chunk_lds(const Bucket& b, int chunk)
{
    return sizeof(float) * (3 * (size_t)b.sp_stride + (size_t)round_up_ll(chunk, 4));
}

// largest chunk (a multiple of 256 clauses, at most the bucket's largest m)
// with which `wgs` workgroups fit in the LDS of one CU; 0 if none does
static int chunk_for(const Batch& B, const Bucket& b, int wgs);

// register-resident variants (workgroup size, clauses and variables per
// thread); a bucket uses the first one its largest instance fits
struct RegVariant {
    int bs, cpt, vpt;
};
static const RegVariant REG_VARIANTS[] = {{256, 1, 1}, {256, 4, 2}, {512, 8, 4}, {768, 12, 3}, {1024, 9, 2}, {512, 17, 4}};

// first variant the bucket's largest instance fits (of workgroup size bs, or any)
static const RegVariant* reg_variant(const Bucket& b, int bs = 0)
{
    for (const RegVariant& r : REG_VARIANTS)
        if ((!bs || r.bs == bs) && b.maxm <= r.bs * r.cpt && b.maxn <= r.bs * r.vpt) return &r;
    return nullptr;
}

#define ASAT_REG_VARIANTS(X) X(256, 1, 1) X(256, 4, 2) X(512, 8, 4) X(768, 12, 3) X(1024, 9, 2) X(512, 17, 4)

static const void* reg_kernel(int bs, int cpt, int vpt, int kk)
{
#define ASAT_REG_PTR(B, C, V)                                                          \
    if (bs == B && cpt == C && vpt == V)                                               \
        return kk == 3 ? (const void*)k_batch_reg<B, C, V, 3> : (const void*)k_batch_reg<B, C, V, 0>;
    ASAT_REG_VARIANTS(ASAT_REG_PTR)
#undef ASAT_REG_PTR
    return nullptr;
}

static void launch_reg(hipStream_t st, const Params& P, int grid, int bs, int cpt, int vpt, int kk, size_t shmem)
{
#define ASAT_REG_LAUNCH(B, C, V)                                                               \
    if (bs == B && cpt == C && vpt == V) {                                                     \
        if (kk == 3) hipLaunchKernelGGL((k_batch_reg<B, C, V, 3>), dim3(grid), dim3(B), shmem, st, P); \
        else hipLaunchKernelGGL((k_batch_reg<B, C, V, 0>), dim3(grid), dim3(B), shmem, st, P); \
        return;                                                                                \
    }
    ASAT_REG_VARIANTS(ASAT_REG_LAUNCH)
#undef ASAT_REG_LAUNCH
    throw Error(ASAT_ERR_HIP, "no register-resident kernel for this tiling");
}

template <int MODE>
static void launch_mode(hipStream_t st, const Params& P, int grid, int bs, size_t shmem)
{
    if (bs == 256) hipLaunchKernelGGL((k_batch<256, MODE>), dim3(grid), dim3(256), shmem, st, P);
    else if (bs == 512) hipLaunchKernelGGL((k_batch<512, MODE>), dim3(grid), dim3(512), shmem, st, P);
    else hipLaunchKernelGGL((k_batch<1024, MODE>), dim3(grid), dim3(1024), shmem, st, P);
}

template <int MODE>
static const void* kernel_of(int bs)
{
    return bs == 256 ? (const void*)k_batch<256, MODE> : bs == 512 ? (const void*)k_batch<512, MODE>
                                                                   : (const void*)k_batch<1024, MODE>;
}

static void set_lds_attr(int bs, int mode, size_t bytes)
{
    const void* f = mode == 1 ? kernel_of<1>(bs) : mode == 3 ? kernel_of<3>(bs) : kernel_of<0>(bs);
    HIP_CHECK(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
}

static void dispatch(hipStream_t st, const Params& P, const Bucket& b)
{
    const int grid = b.grid, bs = b.block;
    const size_t shmem = b.lds_bytes;
    if (b.lds == 2) launch_reg(st, P, grid, bs, b.cpt, b.vpt, b.kk, shmem);
    else if (b.lds == 1) launch_mode<1>(st, P, grid, bs, shmem);
    else if (b.lds == 3) launch_mode<3>(st, P, grid, bs, shmem);
    else launch_mode<0>(st, P, grid, bs, shmem);
}

static Params params_of(const Batch& B, const Bucket& b)
{
    Params P;
    P.desc = B.d_desc; P.code = B.d_code; P.rowptr = B.d_rowptr; P.xa = B.d_xa; P.xb = B.d_xb;
    P.ctrl = B.d_ctrl; P.scratch = b.d_scratch; P.scratch_stride = b.scratch_stride;
    P.scratch_per_wg = b.scratch_per_wg; P.order = b.d_order; P.ntraj = (int)b.ids.size();
    P.queue = b.d_queue; P.eps = B.eps; P.floor = B.floor; P.max_time = B.max_time; P.max_steps = B.max_steps;
    P.chunk_steps = B.chunk_steps; P.check_every = std::max(1, B.check_every);
    P.sp_stride = b.sp_stride;
    P.vcode = B.d_vcode;
    P.chunk_clauses = b.lds == 3 ? b.chunk : 0;
    return P;
}

// resident workgroups per CU for a configuration with `lds_bytes` of LDS each
static int per_cu(const Batch& B, int bs, size_t lds_bytes)
{
    int by_lds = std::max(1, (int)(B.lds_limit / std::max<size_t>(lds_bytes, 1)));
    int by_threads = std::max(1, 2048 / bs);
    return std::max(1, std::min(by_lds, by_threads));
}

static int chunk_for(const Batch& B, const Bucket& b, int wgs)
{
    // 1 KB left for the kernel's static shared arrays (block reductions)
    const long long avail = (long long)(B.lds_limit / wgs) - (long long)chunk_lds(b, 0) - 1024;
    long long c = std::min<long long>(avail / (long long)sizeof(float), round_up_ll(b.maxm, 256));
    c = c / 256 * 256;
    return c >= 256 ? (int)c : 0;
}

// mode: 0 contributions in global scratch, 1 in LDS, 2 register-resident
// kernel (contributions in LDS; bs comes from its variant).  grid_cap > 0
// limits the resident workgroups (global-scratch mode: the scattered
// contribution writes of fewer workgroups stay in the Infinity Cache).
static void configure_bucket(Batch& B, Bucket& b, int bs, int mode, int grid_cap = 0, int chunk = 0)
{
    const bool lds = mode == 1 || mode == 2;
    if (mode == 2) {
        const RegVariant* r = reg_variant(b, bs);
        if (!r) r = reg_variant(b);
        require(r != nullptr, ASAT_ERR_INVALID_ARGUMENT, "instances too large for the register-resident kernel");
        bs = r->bs;
        b.cpt = r->cpt;
        b.vpt = r->vpt;
    }
    if (mode == 3) {
        require(b.vcode, ASAT_ERR_INVALID_ARGUMENT, "clause-chunked records were not built for this bucket");
        if (chunk <= 0) chunk = chunk_for(B, b, 1);
        if (B.force_chunk > 0) chunk = std::min(chunk, B.force_chunk);
        require(chunk > 0, ASAT_ERR_UNSUPPORTED, "instances too large for the clause-chunked kernel");
    }
    b.block = bs;
    b.lds = mode;
    b.chunk = mode == 3 ? chunk : 0;
    b.lds_bytes = mode == 3 ? chunk_lds(b, chunk) : bucket_lds(b, lds);
    b.grid = std::max(1, std::min((int)b.ids.size(), B.cus * per_cu(B, bs, mode == 0 ? 0 : b.lds_bytes)));
    if (grid_cap > 0) b.grid = std::min(b.grid, grid_cap);
    b.scratch_stride = round_up_ll((long long)b.maxn + b.maxm, 64);
    b.scratch_per_wg = mode == 2 ? 0 : 5 * b.scratch_stride + (mode == 0 ? round_up_ll(b.maxl, 64) : 0);
    (void)hipFree(b.d_scratch);
    b.d_scratch = nullptr;
    const size_t sbytes = sizeof(float) * std::max<size_t>((size_t)b.scratch_per_wg * b.grid, 1);
    HIP_CHECK(hipMalloc((void**)&b.d_scratch, sbytes));
    HIP_CHECK(hipMemset(b.d_scratch, 0, sbytes));
    if (mode == 2)
        HIP_CHECK(hipFuncSetAttribute(reg_kernel(bs, b.cpt, b.vpt, b.kk), hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)b.lds_bytes));
    else
        set_lds_attr(bs, mode, b.lds_bytes);
}

static void reset_state(Batch& B)
{
    const size_t T = B.inst.size();
    HIP_CHECK(hipMemcpy(B.d_xa, B.x0.data(), sizeof(float) * B.x0.size(), hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(B.d_xb, B.x0.data(), sizeof(float) * B.x0.size(), hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(B.d_ctrl, B.h_ctrl0.data(), sizeof(TrajCtrl) * T, hipMemcpyHostToDevice));
    B.h_ctrl = B.h_ctrl0;
    B.solved_wall_ms.assign(T, -1.0);
}

// Autotuner: for each bucket, time one short chunk (2 accepted steps of every
// trajectory in the bucket) for each workgroup size and contribution placement
// that fits (for global scratch also with the resident workgroups capped so
// that their contribution scratch totals 64, 128 or 256 MB), keep the
// fastest per trajectory-attempt, then restore the start.
static void autotune(Batch& B)
{
    const int saved_chunk = B.chunk_steps;
    B.chunk_steps = 2;
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    for (auto& b : B.buckets) {
        struct Cand { int bs, mode, cap, chunk; };
        std::vector<Cand> cand;
        for (int bs : {256, 512, 768, 1024}) {
            const RegVariant* rv = reg_variant(b, bs);
            if (rv && b.lds_fits && (B.force_lds == -1 || B.force_lds == 2) && (!B.force_block || B.force_block == bs))
                cand.push_back({bs, 2, 0, 0});
        }
        const long long contrib_bytes = 4 * round_up_ll(b.maxl, 64);
        for (int bs : {256, 512, 1024}) {
            if (B.force_block && bs != B.force_block) continue;
            if (b.lds_fits && (B.force_lds == -1 || B.force_lds == 1)) cand.push_back({bs, 1, 0, 0});
            if (b.vcode && (B.force_lds == -1 || B.force_lds == 3))
                for (int wgs : {1, 2}) {
                    const int ch = chunk_for(B, b, wgs);
                    if (ch > 0) cand.push_back({bs, 3, 0, ch});
                }
            if (B.force_lds == -1 || B.force_lds == 0) {
                cand.push_back({bs, 0, 0, 0});
                const int full = std::max(1, std::min((int)b.ids.size(), B.cus * per_cu(B, bs, 0)));
                for (long long mb : {256LL, 128LL, 64LL}) {
                    const int cap = (int)std::max<long long>(8, (mb << 20) / std::max<long long>(contrib_bytes, 1));
                    if (cap < full) cand.push_back({bs, 0, cap, 0});
                }
            }
        }
        require(!cand.empty(), ASAT_ERR_INVALID_ARGUMENT, "no kernel configuration fits this bucket");
        double best = 1e300;
        Cand pick = cand.front();
        if (cand.size() > 1) {
            for (auto& c : cand) {
                configure_bucket(B, b, c.bs, c.mode, c.cap, c.chunk);
                reset_state(B);
                Params P = params_of(B, b);
                hipLaunchKernelGGL(k_reset_queue, dim3(1), dim3(64), 0, B.stream, b.d_queue);
                HIP_CHECK(hipEventRecord(e0, B.stream));
                dispatch(B.stream, P, b);
                HIP_CHECK(hipEventRecord(e1, B.stream));
                HIP_CHECK(hipEventSynchronize(e1));
                HIP_CHECK(hipGetLastError());
                float ms = 0.f;
                HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
                std::vector<TrajCtrl> hc(B.inst.size());
                HIP_CHECK(hipMemcpy(hc.data(), B.d_ctrl, sizeof(TrajCtrl) * hc.size(), hipMemcpyDeviceToHost));
                long long att = 0;
                for (int id : b.ids) att += hc[id].attempts;
                double us = ms * 1e3 / std::max<long long>(att, 1);
                if (us < best) { best = us; pick = c; }
            }
        }
        configure_bucket(B, b, pick.bs, pick.mode, pick.cap, pick.chunk);
        b.tuned_us = best < 1e300 ? best : -1.0;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    B.chunk_steps = saved_chunk;
    reset_state(B);
}

static void prepare(Batch& B, bool tune)
{
    require(!B.inst.empty(), ASAT_ERR_INVALID_ARGUMENT, "no trajectories added");
    if (!B.stream) HIP_CHECK(hipStreamCreateWithFlags(&B.stream, hipStreamNonBlocking));
    HIP_CHECK(hipStreamSynchronize(B.stream));
    B.free_dev();
    const int T = (int)B.inst.size();
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    HIP_CHECK(hipDeviceGetAttribute(&B.cus, hipDeviceAttributeMultiprocessorCount, dev));
    {
        int v = 0, v2 = 0;
        HIP_CHECK(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        B.lds_limit = (size_t)v;
        if (hipDeviceGetAttribute(&v2, hipDeviceAttributeSharedMemPerBlockOptin, dev) == hipSuccess && v2 > v)
            B.lds_limit = (size_t)v2;
    }
    std::vector<Desc> desc(T);
    long long code_total = 0, row_total = 0, state_total = 0;
    B.state_off.assign(T, 0);
    for (int i = 0; i < T; i++) {
        const Instance& I = B.inst[i];
        desc[i].code_off = code_total;
        desc[i].row_off = row_total;
        desc[i].state_off = state_total;
        B.state_off[i] = state_total;
        desc[i].n = I.n; desc[i].m = I.m; desc[i].k = I.k; desc[i].nslots = I.nslots;
        desc[i].vcode_off = -1;
        desc[i].knorm = (float)std::pow(2.0, -I.k);
        code_total += (long long)I.code.size();
        row_total += (long long)I.rowptr.size();
        state_total += round_up_ll((long long)I.n + I.m, 64);
    }
    std::vector<uint32_t> code(code_total);
    std::vector<int> rows(row_total);
    B.x0.assign(state_total, 0.f);
    for (int i = 0; i < T; i++) {
        const Instance& I = B.inst[i];
        std::copy(I.code.begin(), I.code.end(), code.begin() + desc[i].code_off);
        std::copy(I.rowptr.begin(), I.rowptr.end(), rows.begin() + desc[i].row_off);
        std::copy(I.pub0.begin(), I.pub0.end(), B.x0.begin() + desc[i].state_off);
    }
    // size buckets: trajectories sorted by work (slots + clauses), largest
    // first, split into up to nbuckets_req groups of equal count
    std::vector<int> order(T);
    std::iota(order.begin(), order.end(), 0);
    auto work = [&](int i) { return (long long)B.inst[i].nslots + B.inst[i].m; };
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return work(a) > work(b); });
    int nb = std::max(1, std::min(B.nbuckets_req, T));
    if (work(order.front()) == work(order.back())) nb = 1;   // uniform batch: one launch
    B.buckets.assign(nb, Bucket());
    for (int q = 0; q < nb; q++) {
        Bucket& b = B.buckets[q];
        int lo = (int)((long long)T * q / nb), hi = (int)((long long)T * (q + 1) / nb);
        b.ids.assign(order.begin() + lo, order.begin() + hi);
        bool allk3 = true;
        for (int id : b.ids) {
            b.maxn = std::max(b.maxn, B.inst[id].n);
            b.maxm = std::max(b.maxm, B.inst[id].m);
            b.maxl = std::max(b.maxl, B.inst[id].nslots);
            allk3 = allk3 && B.inst[id].k == 3;
        }
        b.kk = allk3 ? 3 : 0;
        b.sp_stride = (int)round_up_ll((long long)b.maxn + 1, 4);
        b.lds_fits = bucket_lds(b, true) <= B.lds_limit;
        HIP_CHECK(hipMalloc((void**)&b.d_order, sizeof(int) * b.ids.size()));
        HIP_CHECK(hipMalloc((void**)&b.d_queue, sizeof(int)));
        HIP_CHECK(hipMemcpy(b.d_order, b.ids.data(), sizeof(int) * b.ids.size(), hipMemcpyHostToDevice));
    }
    // clause-chunked records for the buckets whose contributions do not fit in LDS
    long long vc_total = 0;
    for (auto& b : B.buckets) {
        b.vcode = (!b.lds_fits || B.force_lds == 3) && chunk_for(B, b, 1) > 0;
        if (!b.vcode) continue;
        for (int id : b.ids) {
            desc[id].vcode_off = vc_total;
            vc_total += B.inst[id].nslots;
        }
    }
    if (vc_total > 0) {
        std::vector<unsigned long long> vc((size_t)vc_total);
        for (int i = 0; i < T; i++)
            if (desc[i].vcode_off >= 0) build_vcode(B.inst[i], vc.data() + desc[i].vcode_off);
        HIP_CHECK(hipMalloc((void**)&B.d_vcode, sizeof(unsigned long long) * (size_t)vc_total));
        HIP_CHECK(hipMemcpy(B.d_vcode, vc.data(), sizeof(unsigned long long) * (size_t)vc_total,
                            hipMemcpyHostToDevice));
    }
    HIP_CHECK(hipMalloc((void**)&B.d_desc, sizeof(Desc) * T));
    HIP_CHECK(hipMalloc((void**)&B.d_code, sizeof(uint32_t) * std::max<long long>(code_total, 1)));
    HIP_CHECK(hipMalloc((void**)&B.d_rowptr, sizeof(int) * std::max<long long>(row_total, 1)));
    HIP_CHECK(hipMalloc((void**)&B.d_xa, sizeof(float) * std::max<long long>(state_total, 1)));
    HIP_CHECK(hipMalloc((void**)&B.d_xb, sizeof(float) * std::max<long long>(state_total, 1)));
    HIP_CHECK(hipMalloc((void**)&B.d_ctrl, sizeof(TrajCtrl) * T));
    HIP_CHECK(hipMemcpy(B.d_desc, desc.data(), sizeof(Desc) * T, hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(B.d_code, code.data(), sizeof(uint32_t) * code_total, hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(B.d_rowptr, rows.data(), sizeof(int) * row_total, hipMemcpyHostToDevice));
    B.h_ctrl0.assign(T, TrajCtrl());
    for (int i = 0; i < T; i++) {
        TrajCtrl& c = B.h_ctrl0[i];
        memset(&c, 0, sizeof(c));
        c.h = 1e-9; c.hnext = 1e-9;
        c.min_viol = INT32_MAX; c.viol = -1; c.sat_step = -1; c.sat_time = -1.0;
        c.status = RUNNING;
    }
    for (auto& b : B.buckets) {
        int bs = B.force_block ? B.force_block : 1024;
        const RegVariant* rv = reg_variant(b);
        const bool reg_ok = rv && b.lds_fits && (!B.force_block || B.force_block == rv->bs);
        int mode = B.force_lds == -1 ? (reg_ok ? 2 : (b.lds_fits ? 1 : (b.vcode ? 3 : 0))) : B.force_lds;
        require(mode == 0 || mode == 3 || b.lds_fits, ASAT_ERR_UNSUPPORTED, "instances too large for LDS contributions");
        require(mode != 3 || b.vcode, ASAT_ERR_UNSUPPORTED, "instances too large for the clause-chunked kernel");
        require(mode != 2 || reg_variant(b), ASAT_ERR_UNSUPPORTED,
                "instances too large for the register-resident kernel");
        configure_bucket(B, b, bs, mode);
    }
    reset_state(B);
    if (tune) autotune(B);
    B.prepared = true;
}

// run chunks until every trajectory has finished (or max_chunks chunks)
static void run(Batch& B, int max_chunks, int graph_chunks)
{
    require(B.prepared, ASAT_ERR_NOT_CONFIGURED, "batch has not been prepared");
    require(max_chunks > 0 && graph_chunks > 0, ASAT_ERR_INVALID_ARGUMENT, "chunk counts must be positive");
    const int T = (int)B.inst.size();
    // events, graph and executable are released on every exit, and a capture
    // left open by an error is ended, so the stream stays usable
    struct Owned {
        hipStream_t st;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        bool capturing = false;
        ~Owned()
        {
            if (capturing) {
                hipGraph_t g = nullptr;
                if (hipStreamEndCapture(st, &g) == hipSuccess && g) (void)hipGraphDestroy(g);
            }
            if (exec) (void)hipGraphExecDestroy(exec);
            if (graph) (void)hipGraphDestroy(graph);
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
        }
    } o{B.stream};
    HIP_CHECK(hipEventCreate(&o.e0));
    HIP_CHECK(hipEventCreate(&o.e1));
    hipEvent_t e0 = o.e0, e1 = o.e1;
    // one graph = graph_chunks x (per bucket: reset its queue, persistent launch)
    HIP_CHECK(hipStreamBeginCapture(B.stream, hipStreamCaptureModeThreadLocal));
    o.capturing = true;
    for (int c = 0; c < graph_chunks; c++) {
        for (auto& b : B.buckets) {
            Params P = params_of(B, b);
            hipLaunchKernelGGL(k_reset_queue, dim3(1), dim3(64), 0, B.stream, b.d_queue);
            dispatch(B.stream, P, b);
        }
    }
    o.capturing = false;
    HIP_CHECK(hipStreamEndCapture(B.stream, &o.graph));
    HIP_CHECK(hipGraphInstantiate(&o.exec, o.graph, nullptr, nullptr, 0));
    hipGraphExec_t exec = o.exec;
    auto t0 = std::chrono::steady_clock::now();
    HIP_CHECK(hipEventRecord(e0, B.stream));
    int launched = 0;
    std::vector<TrajCtrl>& hc = B.h_ctrl;
    while (launched < max_chunks) {
        HIP_CHECK(hipGraphLaunch(exec, B.stream));
        launched += graph_chunks;
        HIP_CHECK(hipMemcpyAsync(hc.data(), B.d_ctrl, sizeof(TrajCtrl) * T, hipMemcpyDeviceToHost, B.stream));
        HIP_CHECK(hipStreamSynchronize(B.stream));
        double now = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        int live = 0;
        for (int i = 0; i < T; i++) {
            if (hc[i].status == RUNNING) live++;
            else if (hc[i].status == SOLVED && B.solved_wall_ms[i] < 0) B.solved_wall_ms[i] = now;
        }
        if (!live) break;
    }
    HIP_CHECK(hipEventRecord(e1, B.stream));
    HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    B.run_ms += ms;
    B.chunks_run += launched;
}

}  // namespace batch
}  // namespace asat

// ===========================================================================
// C ABI
// ===========================================================================
using namespace asat;
using namespace asat::batch;

struct asat_batch { Batch b; };

#define BAPI_BEGIN try {
#define BAPI_END                                                 \
    }                                                            \
    catch (const asat::Error& e) {                               \
        asat::g_last_error = e.what();                           \
        return e.code;                                           \
    }                                                            \
    catch (const std::exception& e) {                            \
        asat::g_last_error = e.what();                           \
        return ASAT_ERR_HIP;                                     \
    }                                                            \
    return ASAT_OK;

extern "C" {

int asat_batch_create(asat_batch_t** out)
{
    BAPI_BEGIN
    require(out != nullptr, ASAT_ERR_INVALID_ARGUMENT, "null argument: out");
    *out = new asat_batch();
    BAPI_END
}

int asat_batch_destroy(asat_batch_t* b)
{
    BAPI_BEGIN
    delete b;
    BAPI_END
}

int asat_batch_add(asat_batch_t* b, const int* literals, const int* lengths, int nclauses, const float* pub0,
                   int* index_out)
{
    BAPI_BEGIN
    require(b && pub0, ASAT_ERR_INVALID_ARGUMENT, "null argument");
    Instance I;
    build_instance(literals, lengths, nclauses, I);
    I.pub0.assign(pub0, pub0 + (size_t)I.n + I.m);
    b->b.inst.push_back(std::move(I));
    b->b.prepared = false;
    if (index_out) *index_out = (int)b->b.inst.size() - 1;
    BAPI_END
}

int asat_batch_set_params(asat_batch_t* b, double eps, long long max_steps, double max_time, int chunk_steps,
                          int check_every)
{
    BAPI_BEGIN
    require(b != nullptr, ASAT_ERR_INVALID_ARGUMENT, "null batch");
    require(!(eps < 1e-16 || eps > 1.0), ASAT_ERR_INVALID_ARGUMENT, "epsilon must be between 1e-16 and 1");
    require(max_steps >= 1, ASAT_ERR_INVALID_ARGUMENT, "Max iterations must be positive");
    require(max_time > 0.0, ASAT_ERR_INVALID_ARGUMENT, "Max time must be positive");
    require(chunk_steps >= 1 && check_every >= 1, ASAT_ERR_INVALID_ARGUMENT, "chunk/check sizes must be positive");
    b->b.eps = eps; b->b.max_steps = max_steps; b->b.max_time = max_time;
    b->b.chunk_steps = chunk_steps; b->b.check_every = check_every;
    BAPI_END
}

int asat_batch_set_xscal_floor(asat_batch_t* b, double floor)
{
    BAPI_BEGIN
    require(b != nullptr, ASAT_ERR_INVALID_ARGUMENT, "null batch");
    require(floor > 0.0 && floor <= 1.0, ASAT_ERR_INVALID_ARGUMENT, "xscal floor must be in (0, 1]");
    b->b.floor = floor;
    BAPI_END
}

int asat_batch_set_chunk(asat_batch_t* b, int clauses)
{
    BAPI_BEGIN
    require(b != nullptr, ASAT_ERR_INVALID_ARGUMENT, "null batch");
    require(clauses >= 0, ASAT_ERR_INVALID_ARGUMENT, "chunk size must be >= 0");
    b->b.force_chunk = clauses;
    b->b.prepared = false;
    BAPI_END
}

int asat_batch_set_tiling(asat_batch_t* b, int block, int lds_mode, int nbuckets)
{
    BAPI_BEGIN
    require(b != nullptr, ASAT_ERR_INVALID_ARGUMENT, "null batch");
    require(block == 0 || block == 256 || block == 512 || block == 1024, ASAT_ERR_INVALID_ARGUMENT,
            "block must be 0 (autotune), 256, 512 or 1024");
    require(lds_mode >= -1 && lds_mode <= 3, ASAT_ERR_INVALID_ARGUMENT, "lds_mode must be -1, 0, 1, 2 or 3");
    require(nbuckets >= 1 && nbuckets <= 64, ASAT_ERR_INVALID_ARGUMENT, "nbuckets must be in [1, 64]");
    b->b.force_block = block; b->b.force_lds = lds_mode; b->b.nbuckets_req = nbuckets;
    b->b.prepared = false;
    BAPI_END
}

int asat_batch_prepare(asat_batch_t* b, int autotune)
{
    BAPI_BEGIN
    require(b != nullptr, ASAT_ERR_INVALID_ARGUMENT, "null batch");
    prepare(b->b, autotune != 0);
    BAPI_END
}

int asat_batch_run(asat_batch_t* b, int max_chunks, int graph_chunks)
{
    BAPI_BEGIN
    require(b != nullptr, ASAT_ERR_INVALID_ARGUMENT, "null batch");
    run(b->b, max_chunks, graph_chunks);
    BAPI_END
}

int asat_batch_reset(asat_batch_t* b)
{
    BAPI_BEGIN
    require(b != nullptr && b->b.prepared, ASAT_ERR_NOT_CONFIGURED, "batch has not been prepared");
    reset_state(b->b);
    b->b.run_ms = 0.0;
    b->b.chunks_run = 0;
    BAPI_END
}

int asat_batch_size(const asat_batch_t* b, int* count)
{
    BAPI_BEGIN
    require(b && count, ASAT_ERR_INVALID_ARGUMENT, "null argument");
    *count = (int)b->b.inst.size();
    BAPI_END
}

int asat_batch_result(const asat_batch_t* b, int index, asat_batch_result_t* out)
{
    BAPI_BEGIN
    require(b && out, ASAT_ERR_INVALID_ARGUMENT, "null argument");
    require(b->b.prepared, ASAT_ERR_NOT_CONFIGURED, "batch has not been prepared");
    require(index >= 0 && index < (int)b->b.inst.size(), ASAT_ERR_INVALID_ARGUMENT, "index out of range");
    const TrajCtrl& c = b->b.h_ctrl[index];
    out->status = c.status;
    out->steps = c.steps;
    out->attempts = c.attempts;
    out->rhscount = c.rhscount;
    out->time = c.t;
    out->hdid = c.hdid;
    out->hnext = c.hnext;
    out->min_violations = c.min_viol == INT32_MAX ? -1 : c.min_viol;
    out->violations = c.viol;
    out->sat_step = c.sat_step;
    out->sat_time = c.sat_time;
    out->solved_wall_ms = b->b.solved_wall_ms[index];
    BAPI_END
}

int asat_batch_get_state(const asat_batch_t* b, int index, float* pub)
{
    BAPI_BEGIN
    require(b && pub, ASAT_ERR_INVALID_ARGUMENT, "null argument");
    require(b->b.prepared, ASAT_ERR_NOT_CONFIGURED, "batch has not been prepared");
    require(index >= 0 && index < (int)b->b.inst.size(), ASAT_ERR_INVALID_ARGUMENT, "index out of range");
    const Instance& I = b->b.inst[index];
    const float* src = (b->b.h_ctrl[index].parity ? b->b.d_xb : b->b.d_xa) + b->b.state_off[index];
    HIP_CHECK(hipStreamSynchronize(b->b.stream));
    HIP_CHECK(hipMemcpy(pub, src, sizeof(float) * ((size_t)I.n + I.m), hipMemcpyDeviceToHost));
    BAPI_END
}

int asat_batch_stats(const asat_batch_t* b, double* run_ms, int* chunks, int* nbuckets)
{
    BAPI_BEGIN
    require(b != nullptr, ASAT_ERR_INVALID_ARGUMENT, "null batch");
    if (run_ms) *run_ms = b->b.run_ms;
    if (chunks) *chunks = b->b.chunks_run;
    if (nbuckets) *nbuckets = (int)b->b.buckets.size();
    BAPI_END
}

int asat_batch_bucket_info(const asat_batch_t* b, int bucket, asat_batch_bucket_t* out)
{
    BAPI_BEGIN
    require(b && out, ASAT_ERR_INVALID_ARGUMENT, "null argument");
    require(bucket >= 0 && bucket < (int)b->b.buckets.size(), ASAT_ERR_INVALID_ARGUMENT, "bucket out of range");
    const Bucket& k = b->b.buckets[bucket];
    out->trajectories = (int)k.ids.size();
    out->max_n = k.maxn;
    out->max_m = k.maxm;
    out->max_slots = k.maxl;
    out->block = k.block;
    out->lds = k.lds;
    out->grid = k.grid;
    out->lds_bytes = (long long)k.lds_bytes;
    out->tuned_us_per_attempt = k.tuned_us;
    BAPI_END
}

}  // extern "C"
