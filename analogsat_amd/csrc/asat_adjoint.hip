// Discrete adjoint of the FP32 Cash-Karp integrator on MI355X.  No reference
// counterpart: the reference only integrates forward.
//
// Gradient of V(T) = sum_m a_m K_m(x_T)^2 after T accepted steps, with
// respect to the start state, for a batch of trajectories of one instance.
// Forward: every trajectory runs the adaptive Cash-Karp integrator (FP32
// state, FP64 control, the plain solver's arithmetic) with its own step
// size, all trajectories in lockstep (one per blockIdx.y); every accepted
// step size is logged on the device and a checkpoint is kept every C steps.
// Reverse, again in lockstep, segment by segment from the last checkpoint:
//   recompute the segment's states with the logged (frozen) step sizes, then
//   for each step backwards recompute its six stage inputs Y_l and run
//     mu_l = h (c_l lam' + sum_{j>l} b_jl nu_j),  nu_l = J(Y_l)^T mu_l,
//     lam  = lam' + sum_l nu_l
// The recomputation uses the forward's FP32 arithmetic (same operation
// order), so it reproduces the realised forward states exactly; everything
// adjoint (mu, nu, lam, the VJP) is FP64.  VJP of one clause (f_q = 1 - c_q
// s_q, K = 2^-k prod f, K_q / K_qp the products without q / without q and p):
//   u_a    = w_a K^2 + sum_q w_q 2 c_q K K_q
//   u_s[p] = w_a (-2 a c_p K K_p) + w_p (-2 a K_p^2) + sum_{q!=p} w_q (-4 a c_q c_p K K_qp)
// Per-variable sums run over each variable's slots in slot order
// (deterministic).  FP32 terms are written at their var-major positions (a
// variable's terms are one contiguous run; 510 KB per trajectory at
// N = 1e4, M = 42 500, L2-resident); FP64 terms stay in clause order and are gathered through the
// var-major slot list (scattered FP64 writes cost more than the gathers).
// Internal layout of every vector: [0 | s_1..s_N | a_1..a_M], stride D.
#include "asat_common.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

namespace asat {
extern thread_local std::string g_last_error;

namespace adj {

constexpr int BLK = 256;

struct Dev {
    const int* lit;       // m*k (zero padded)
    const int* csr_ptr;   // n+2, var-major over slots
    const int* csr_slot;  // slot indices, ascending per variable
    const int* slot_pos;  // each slot's position in that var-major order, -1: padding
    int n, m, k;
    long long D;          // vector stride (>= n+m+1)
    long long L;          // m*k
    int ntraj;            // trajectories (grids are padded to a multiple of 8)
};

// Trajectory-batched grids are (blocks per trajectory, ntraj rounded up to
// 8).  Workgroups are dealt to the 8 XCDs round-robin by linear id; the id is
// remapped so that every block of trajectory tr runs on XCD tr % 8, and each
// XCD keeps its own trajectories' gather targets (spins, contributions) in
// its L2.  Blocks of the padding trajectories (tr >= ntraj) return at once.
struct TB {
    long long x, tr;
};
static inline int ntraj8(int ntraj) { return (ntraj + 7) / 8 * 8; }

__device__ __forceinline__ TB traj_block()
{
    const long long per = 8LL * gridDim.x;
    const long long L = (long long)blockIdx.y * gridDim.x + blockIdx.x;
    const long long r = L % per;
    return {r >> 3, (L / per) * 8 + (r & 7)};
}

// ---- recompute (forward arithmetic; T = float, or double for the FP64 replay)
template <typename T, int K>
__global__ __launch_bounds__(BLK) void k_rhs_clause(Dev d, const T* Y, T* out, T* contrib)
{
    const TB tb = traj_block();
    const int m = tb.x * BLK + threadIdx.x;
    const long long tr = tb.tr;
    if (tr >= d.ntraj) return;
    if (m >= d.m) return;
    const T* y = Y + tr * d.D;
    int l[K];
    T f[K];
#pragma unroll
    for (int q = 0; q < K; q++) l[q] = d.lit[(size_t)m * K + q];
    T km = (T)(1.0 / (1 << K));
#pragma unroll
    for (int q = 0; q < K; q++) {
        f[q] = spin_factor<T>(l[q], y);
        km *= f[q];
    }
    const long long ai = d.n + 1 + m;
    const T am = y[ai];
    out[tr * d.D + ai] = am * km * km;
    const T r = km * (T)(1.0 / (1 << K)) * am;
    T o[K];
#pragma unroll
    for (int q = 0; q < K; q++) o[q] = (T)0;
    write_contribs<T, K, 0>(l, f, r, o);
    if constexpr (sizeof(T) == 4) {
        T* c = contrib + tr * d.L;
#pragma unroll
        for (int q = 0; q < K; q++)
            if (l[q] != 0) c[d.slot_pos[(size_t)m * K + q]] = o[q];
    } else {
        T* c = contrib + tr * d.L + (long long)m * K;
#pragma unroll
        for (int q = 0; q < K; q++)
            if (l[q] != 0) c[q] = o[q];
    }
}

template <typename T>
__global__ __launch_bounds__(BLK) void k_rhs_var(Dev d, const T* contrib, T* out)
{
    const TB tb = traj_block();
    const int i = tb.x * BLK + threadIdx.x + 1;
    const long long tr = tb.tr;
    if (tr >= d.ntraj) return;
    if (i > d.n) return;
    const T* c = contrib + tr * d.L;
    T sum = (T)0;
    if constexpr (sizeof(T) == 4)
        sum = sum_run(sum, c, d.csr_ptr[i], d.csr_ptr[i + 1]);
    else
        sum = sum_run_idx(sum, c, d.csr_slot, d.csr_ptr[i], d.csr_ptr[i + 1]);
    out[tr * d.D + i] = sum;
}

// stage input J of every element (spins and aux) for every trajectory
template <typename T, int J>
__global__ __launch_bounds__(BLK) void k_stage(Dev d, const T* X, const T* k1, const T* k2, const T* k3, const T* k4,
                                                const T* k5, const double* h, T* Y)
{
    const TB tb = traj_block();
    const long long i = (long long)tb.x * BLK + threadIdx.x + 1;
    const long long tr = tb.tr;
    if (tr >= d.ntraj) return;
    if (i > (long long)d.n + d.m) return;
    const long long o = tr * d.D + i;
    const T hs = (T)h[tr];
    Y[o] = stage_input<T, J>(X[o], k1[o], J > 2 ? k2[o] : (T)0, J > 3 ? k3[o] : (T)0, J > 4 ? k4[o] : (T)0,
                             J > 5 ? k5[o] : (T)0, hs);
}

template <typename T>
__global__ __launch_bounds__(BLK) void k_final(Dev d, const T* X, const T* k1, const T* k3, const T* k4, const T* k6,
                                               const double* h, T* Xn)
{
    const TB tb = traj_block();
    const long long i = (long long)tb.x * BLK + threadIdx.x;
    const long long tr = tb.tr;
    if (tr >= d.ntraj) return;
    if (i > (long long)d.n + d.m) return;
    const long long o = tr * d.D + i;
    if (i == 0) {
        Xn[o] = (T)-1.0;
        return;
    }
    using B = Tab<T>;
    const T hs = (T)h[tr];
    Xn[o] = X[o] + hs * (B::c1 * k1[o] + B::c3 * k3[o] + B::c4 * k4[o] + B::c6 * k6[o]);
}

// ---- FP64 adjoint ------------------------------------------------------------
template <int K>
__device__ __forceinline__ double prod_skip(const double (&fd)[K], int s1, int s2, double kn)
{
    double p = kn;
#pragma unroll
    for (int q = 0; q < K; q++)
        if (q != s1 && q != s2) p = p * fd[q];
    return p;
}

// u_a directly, per-slot spin contributions into contrib64
template <int K>
__global__ __launch_bounds__(BLK) void k_vjp_clause(Dev d, const float* Y, const double* w, double* u,
                                                    double* contrib)
{
    const TB tb = traj_block();
    const int m = tb.x * BLK + threadIdx.x;
    const long long tr = tb.tr;
    if (tr >= d.ntraj) return;
    if (m >= d.m) return;
    const float* y = Y + tr * d.D;
    const double* wt = w + tr * d.D;
    const double kn = 1.0 / (1 << K);
    int v[K];
    double fd[K], cs[K], wv[K];
#pragma unroll
    for (int q = 0; q < K; q++) {
        const int l = d.lit[(size_t)m * K + q];
        v[q] = l < 0 ? -l : l;
        cs[q] = l < 0 ? -1.0 : 1.0;
        fd[q] = (double)spin_factor<float>(l, y);
        wv[q] = v[q] ? wt[v[q]] : 0.0;
    }
    const double Kc = prod_skip<K>(fd, -1, -1, kn);
    const long long ai = d.n + 1 + m;
    const double a = (double)y[ai];
    const double wa = wt[ai];
    double ua = wa * Kc * Kc;
#pragma unroll
    for (int q = 0; q < K; q++)
        if (v[q]) ua += wv[q] * (2.0 * cs[q] * Kc * prod_skip<K>(fd, q, -1, kn));
    u[tr * d.D + ai] = ua;
    double* c = contrib + tr * d.L + (long long)m * K;
#pragma unroll
    for (int p = 0; p < K; p++) {
        if (!v[p]) continue;
        const double Kp = prod_skip<K>(fd, p, -1, kn);
        double t = wa * (-2.0 * a * cs[p] * Kc * Kp) + wv[p] * (-2.0 * a * Kp * Kp);
#pragma unroll
        for (int q = 0; q < K; q++) {
            if (q == p || !v[q]) continue;
            t += wv[q] * (-4.0 * a * cs[q] * cs[p] * Kc * prod_skip<K>(fd, q, p, kn));
        }
        c[p] = t;
    }
}

// grad V: aux K^2, spin contributions -2 a c K K_p; per-block partial sums of V
template <typename TX, int K>
__global__ __launch_bounds__(BLK) void k_gradv_clause(Dev d, const TX* X, double* g, double* contrib,
                                                      double* vpart, int nblk)
{
    const TB tb = traj_block();
    __shared__ double sh[BLK / 64];
    const int m = tb.x * BLK + threadIdx.x;
    const long long tr = tb.tr;
    if (tr >= d.ntraj) return;
    double vloc = 0.0;
    if (m < d.m) {
        const TX* x = X + tr * d.D;
        const double kn = 1.0 / (1 << K);
        int v[K];
        double fd[K], cs[K];
#pragma unroll
        for (int q = 0; q < K; q++) {
            const int l = d.lit[(size_t)m * K + q];
            v[q] = l < 0 ? -l : l;
            cs[q] = l < 0 ? -1.0 : 1.0;
            fd[q] = (double)spin_factor<TX>(l, x);
        }
        const double Kc = prod_skip<K>(fd, -1, -1, kn);
        const long long ai = d.n + 1 + m;
        const double a = (double)x[ai];
        vloc = a * Kc * Kc;
        g[tr * d.D + ai] = Kc * Kc;
        double* c = contrib + tr * d.L + (long long)m * K;
#pragma unroll
        for (int p = 0; p < K; p++)
            if (v[p]) c[p] = -2.0 * a * cs[p] * Kc * prod_skip<K>(fd, p, -1, kn);
    }
    // deterministic block sum (fixed tree), summed in block order on the host side
    double s = vloc;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < BLK / 64; i++) t += sh[i];
        vpart[tr * nblk + tb.x] = t;
    }
}

__global__ void k_sum_parts(const double* vpart, int nblk, double* V)
{
    const int tr = blockIdx.x;
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int b = 0; b < nblk; b++) t += vpart[(long long)tr * nblk + b];
        V[tr] = t;
    }
}

// mu = h * (Cw_l lam + sum_{j>l} B_jl nu_j)   (j descending order of terms as written)
__global__ __launch_bounds__(BLK) void k_mu(Dev d, int l, const double* lam, const double* nu, const double* hh,
                                            double* mu)
{
    const TB tb = traj_block();
    const double Bj[7][6] = {
        {0}, {0},
        {(double)Tab<float>::b21},
        {(double)Tab<float>::b31, (double)Tab<float>::b32},
        {(double)Tab<float>::b41, (double)Tab<float>::b42, (double)Tab<float>::b43},
        {(double)Tab<float>::b51, (double)Tab<float>::b52, (double)Tab<float>::b53, (double)Tab<float>::b54},
        {(double)Tab<float>::b61, (double)Tab<float>::b62, (double)Tab<float>::b63, (double)Tab<float>::b64,
         (double)Tab<float>::b65}};
    const double Cw[7] = {0, (double)Tab<float>::c1, 0, (double)Tab<float>::c3, (double)Tab<float>::c4, 0,
                          (double)Tab<float>::c6};
    const long long i = (long long)tb.x * BLK + threadIdx.x;
    const long long tr = tb.tr;
    if (tr >= d.ntraj) return;
    if (i > (long long)d.n + d.m) return;
    const long long o = tr * d.D + i;
    if (i == 0) {
        mu[o] = 0.0;
        return;
    }
    const long long vs = (long long)d.ntraj * d.D;   // stride between nu_j vectors
    double acc = Cw[l] * lam[o];
    for (int j = l + 1; j <= 6; j++) acc += Bj[j][l - 1] * nu[(long long)(j - 1) * vs + o];
    mu[o] = hh[tr] * acc;
}

__global__ __launch_bounds__(BLK) void k_lam(Dev d, double* lam, const double* nu)
{
    const TB tb = traj_block();
    const long long i = (long long)tb.x * BLK + threadIdx.x + 1;
    const long long tr = tb.tr;
    if (tr >= d.ntraj) return;
    if (i > (long long)d.n + d.m) return;
    const long long o = tr * d.D + i;
    const long long vs = (long long)d.ntraj * d.D;
    lam[o] = lam[o] + nu[o] + nu[vs + o] + nu[2 * vs + o] + nu[3 * vs + o] + nu[4 * vs + o] + nu[5 * vs + o];
}

__global__ void k_zero_slot0(double* v, long long D, int ntraj)
{
    const int tr = blockIdx.x * blockDim.x + threadIdx.x;
    if (tr < ntraj) v[(long long)tr * D] = 0.0;
}

// plain-solver layout (spins 1..N, aux at aoff) -> internal layout (aux at N+1)
// ---- lockstep adaptive forward (FP32 state, FP64 control) -------------------
// Every trajectory runs CpuRungeKutta's Step/rkqs (AnalogSAT/cpu_sat/
// CpuRungeKutta.cpp:65-100,217-267) with its own h; all trajectories take
// their attempts together.  The arithmetic is that of the plain FP32
// integrator, so each trajectory's accepted steps are the ones that solver
// takes alone.
struct TCtl {
    double h, hnext, time, errmax;
    int steps, attempts, active, underflow, accept, ckpt;   // ckpt: checkpoint slot to write on accept, or -1
};

// per-trajectory controller (the plain solver's k_control); one block each
__global__ __launch_bounds__(BLK) void k_ctl_batch(TCtl* ctl, const double* parts, int nblk, double eps, int T,
                                                   int C, double* hlog, double* hcur)
{
    __shared__ double sh[BLK / 64];
    const int tr = blockIdx.x;
    TCtl& c = ctl[tr];
    if (!c.active) {
        // finished: nothing more is accepted or checkpointed for it
        if (threadIdx.x == 0) { c.accept = 0; c.ckpt = -1; }
        return;
    }
    double v = 0.0;
    for (int i = threadIdx.x; i < nblk; i += BLK) v = fmax(v, parts[(long long)tr * nblk + i]);
    v = block_max<BLK>(v, sh);
    if (threadIdx.x != 0) return;
    const double errmax = v / eps;
    const double h = c.h;
    c.errmax = errmax;
    c.attempts += 1;
    c.accept = 0;
    c.ckpt = -1;
    if (errmax <= 1.0) {
        const double hn = errmax > TabD::ERRCON ? TabD::SAFETY * h * pow_ctrl(errmax, 5) : 5.0 * h;
        hlog[(long long)tr * T + c.steps] = h;
        c.time += h;
        c.steps += 1;
        c.accept = 1;
        if (c.steps % C == 0 || c.steps == T) c.ckpt = (c.steps + C - 1) / C;
        c.hnext = hn;
        c.h = hn;
        if (c.steps >= T) c.active = 0;
    } else {
        const double htemp = TabD::SAFETY * h * pow_ctrl(errmax, 4);
        const double hn = h >= 0.0 ? fmax(htemp, TabD::HFACTOR * h) : fmin(htemp, TabD::HFACTOR * h);
        c.h = hn;
        if (c.time + hn == c.time) {
            c.underflow = 1;
            c.active = 0;
        }
    }
    hcur[tr] = c.h;
}

// accepted trajectories: x <- x_new, and the checkpoint when one is due
__global__ __launch_bounds__(BLK) void k_accept(Dev d, const TCtl* ctl, const float* Xn, float* X, float* ckpt,
                                                long long ntraj)
{
    const TB tb = traj_block();
    const long long i = (long long)tb.x * BLK + threadIdx.x;
    const long long tr = tb.tr;
    if (tr >= d.ntraj) return;
    if (i > (long long)d.n + d.m || !ctl[tr].accept) return;
    const long long o = tr * d.D + i;
    const float v = Xn[o];
    X[o] = v;
    const int c = ctl[tr].ckpt;
    if (c >= 0) ckpt[((long long)c * ntraj + tr) * d.D + i] = v;
}

__global__ void k_ctl_init(TCtl* ctl, double* hcur, int ntraj)
{
    const int tr = blockIdx.x * blockDim.x + threadIdx.x;
    if (tr >= ntraj) return;
    TCtl c;
    c.h = 1e-9; c.hnext = 1e-9; c.time = 0.0; c.errmax = 0.0;   // CpuRungeKutta::Reset
    c.steps = 0; c.attempts = 0; c.active = 1; c.underflow = 0; c.accept = 0; c.ckpt = -1;
    ctl[tr] = c;
    hcur[tr] = 1e-9;
}

// ---- fused stages (FP32 state): the plain solver's stage fusion ------------
// Stage J of every trajectory in two passes, with the stage inputs formed
// inside them (the plain solver's k_clause_gen / var_stage order), so there
// are no stage-input, final-combination or error passes of their own:
//   clause pass J: K_m from the stage spins, the aux stage input from x and
//     k_1..k_{J-1}, k_J(a) = a K^2 (J = 6: x_new of the aux element, and in
//     the adaptive run |xerr/xscal|), slot terms at their var-major positions
//   var pass J:    each variable's terms added in slot order, k_J(s), then
//     stage J+1's spin input (J = 6: x_new, and |xerr/xscal|)
// FZ = false: the adaptive lockstep run (h from the control block, inactive
// trajectories skipped, spin inputs in one buffer S, the J = 6 passes leave
// one error maximum per block in parts[tr]: clause blocks, then var blocks).
// FZ = true: a frozen step with the logged h of every trajectory; the six
// stage inputs Y_1..Y_6 of every element are kept (Y_1 = x) for the
// adjoint, and the spins of Y_J are stage J's spin input.
struct FArgs {
    Dev d;
    const float* X;
    float* Xn;            // null: no x_new (frozen step whose successor is not needed)
    float* k[6];          // k1..k6 (k6 is not stored)
    float* S;             // adaptive: spin input of the next stage
    float* Y;             // frozen: [6][ntraj][D] stage inputs
    float* c;             // [ntraj][L] slot terms, var-major
    long long V;          // ntraj * D
    const double* hcol;   // frozen: step size of every trajectory
    const TCtl* ctl;      // adaptive: control blocks
    double* parts;
    int nblk_c, nblk_v;
};

template <int K, int J, bool FZ>
__global__ __launch_bounds__(BLK) void k_fclause(FArgs a)
{
    __shared__ double sh[BLK / 64];
    const TB tb = traj_block();
    const long long tr = tb.tr;
    if (tr >= a.d.ntraj) return;
    if (!FZ && !a.ctl[tr].active) return;
    const int m = tb.x * BLK + threadIdx.x;
    const double h = FZ ? a.hcol[tr] : a.ctl[tr].h;
    const double hnext = FZ ? h : a.ctl[tr].hnext;
    const float hs = (float)h;
    const long long row = tr * a.d.D;
    const float* x = a.X + row;
    const float* sp = J == 1 ? x : (FZ ? a.Y + (J - 1) * a.V : a.S) + row;
    double err = 0.0;
    if (m < a.d.m) {
        int l[K];
        float f[K];
#pragma unroll
        for (int q = 0; q < K; q++) l[q] = a.d.lit[(size_t)m * K + q];
        float km = (float)(1.0 / (1 << K));
#pragma unroll
        for (int q = 0; q < K; q++) {
            f[q] = spin_factor<float>(l[q], sp);
            km *= f[q];
        }
        const long long ai = a.d.n + 1 + m;
        float am, k1 = 0.f, k3 = 0.f, k4 = 0.f, k5 = 0.f;
        if constexpr (J == 1) {
            am = x[ai];
        } else {
            k1 = a.k[0][row + ai];
            const float k2 = J > 2 ? a.k[1][row + ai] : 0.f;
            k3 = J > 3 ? a.k[2][row + ai] : 0.f;
            k4 = J > 4 ? a.k[3][row + ai] : 0.f;
            k5 = J > 5 ? a.k[4][row + ai] : 0.f;
            am = stage_input<float, J>(x[ai], k1, k2, k3, k4, k5, hs);
        }
        if constexpr (FZ) a.Y[(J - 1) * a.V + row + ai] = am;
        const float kj = am * km * km;
        if constexpr (J < 6) {
            a.k[J - 1][row + ai] = kj;
        } else if (a.Xn) {
            float xn;
            final_combination<float>(x[ai], k1, k3, k4, k5, kj, hs, h, hnext, TabD::TINY, xn, err);
            a.Xn[row + ai] = xn;
            if (m == 0) a.Xn[row] = -1.0f;
        }
        const float r = km * (float)(1.0 / (1 << K)) * am;
        float o[K];
#pragma unroll
        for (int q = 0; q < K; q++) o[q] = 0.f;
        write_contribs<float, K, 0>(l, f, r, o);
        float* cr = a.c + tr * a.d.L;
#pragma unroll
        for (int q = 0; q < K; q++)
            if (l[q] != 0) cr[a.d.slot_pos[(size_t)m * K + q]] = o[q];
    }
    if constexpr (J == 6 && !FZ) {
        const double bm = block_max<BLK>(err, sh);
        if (threadIdx.x == 0) a.parts[tr * (a.nblk_c + a.nblk_v) + tb.x] = bm;
    }
}

template <int J, bool FZ>
__global__ __launch_bounds__(BLK) void k_fvar(FArgs a)
{
    __shared__ double sh[BLK / 64];
    const TB tb = traj_block();
    const long long tr = tb.tr;
    if (tr >= a.d.ntraj) return;
    if (!FZ && !a.ctl[tr].active) return;
    const int i = tb.x * BLK + threadIdx.x + 1;
    const double h = FZ ? a.hcol[tr] : a.ctl[tr].h;
    const double hnext = FZ ? h : a.ctl[tr].hnext;
    const float hs = (float)h;
    const long long row = tr * a.d.D;
    double err = 0.0;
    if (i <= a.d.n) {
        const float* cr = a.c + tr * a.d.L;
        float sum = 0.f;
        sum = sum_run(sum, cr, a.d.csr_ptr[i], a.d.csr_ptr[i + 1]);
        const float kj = sum;
        const long long o = row + i;
        const float xi = a.X[o];
        // stage J+1's spin input goes to S, or to Y_{J+1} in a frozen step
        float* nxt = FZ ? a.Y + J * a.V : a.S;
        if constexpr (FZ && J == 1) a.Y[o] = xi;
        if constexpr (J == 1) {
            a.k[0][o] = kj;
            nxt[o] = stage_input<float, 2>(xi, kj, 0.f, 0.f, 0.f, 0.f, hs);
        } else if constexpr (J == 2) {
            a.k[1][o] = kj;
            nxt[o] = stage_input<float, 3>(xi, a.k[0][o], kj, 0.f, 0.f, 0.f, hs);
        } else if constexpr (J == 3) {
            a.k[2][o] = kj;
            nxt[o] = stage_input<float, 4>(xi, a.k[0][o], a.k[1][o], kj, 0.f, 0.f, hs);
        } else if constexpr (J == 4) {
            a.k[3][o] = kj;
            nxt[o] = stage_input<float, 5>(xi, a.k[0][o], a.k[1][o], a.k[2][o], kj, 0.f, hs);
        } else if constexpr (J == 5) {
            a.k[4][o] = kj;
            nxt[o] = stage_input<float, 6>(xi, a.k[0][o], a.k[1][o], a.k[2][o], a.k[3][o], kj, hs);
        } else if (a.Xn) {
            float xn;
            final_combination<float>(xi, a.k[0][o], a.k[2][o], a.k[3][o], a.k[4][o], kj, hs, h, hnext, TabD::TINY,
                                     xn, err);
            a.Xn[o] = xn;
        }
    }
    if constexpr (J == 6 && !FZ) {
        const double bm = block_max<BLK>(err, sh);
        if (threadIdx.x == 0) a.parts[tr * (a.nblk_c + a.nblk_v) + a.nblk_c + tb.x] = bm;
    }
}

// ---------------------------------------------------------------------------
struct Adj {
    int n = 0, m = 0, k = 0;
    long long D = 0, L = 0;
    double eps = 1e-6;
    int T = 0, C = 32;
    int ntraj = 0;
    bool have_problem = false, have_forward = false;
    std::vector<int> lits_in, lens_in;
    hipStream_t stream = nullptr;
    int *d_lit = nullptr, *d_csr_ptr = nullptr, *d_csr_slot = nullptr, *d_slot_pos = nullptr;
    // per batch
    float* ckpt = nullptr;       // [ncheck][ntraj][D]
    float* seg = nullptr;        // [C+1][ntraj][D]
    float* Y = nullptr;          // [6][ntraj][D]
    float* Yseg = nullptr;       // [C][6][ntraj][D] stage inputs of a segment's steps (backward)
    float* kk = nullptr;         // [6][ntraj][D]
    float* c32 = nullptr;        // [ntraj][L]
    double* c64 = nullptr;       // [ntraj][L]
    double* lam = nullptr;       // [ntraj][D]
    double* mu = nullptr;        // [ntraj][D]
    double* nu = nullptr;        // [6][ntraj][D]
    double* hlog = nullptr;      // [ntraj][T]
    double* hcol = nullptr;      // [T][ntraj] step sizes by step
    double* hcolr = nullptr;     // [T][ntraj] FP32-rounded step sizes (adjoint)
    const double* hcur = nullptr;   // current step's column
    const double* hcurr = nullptr;
    double* vpart = nullptr;
    double* dV = nullptr;
    TCtl* ctl = nullptr;         // [ntraj] lockstep forward control
    double* hstep = nullptr;     // [ntraj] current attempt's step size
    double* eparts = nullptr;    // [ntraj][nblk] error-norm block maxima
    int ncheck = 0;
    double fwd_ms = 0, bwd_ms = 0;
    long long fwd_attempts = 0;   // forward: attempts summed over the trajectories
    long long fwd_lockstep = 0;   // forward: lockstep attempts launched

    void free_batch()
    {
        for (void* p : {(void*)ckpt, (void*)seg, (void*)Y, (void*)Yseg, (void*)kk, (void*)c32, (void*)c64, (void*)lam,
                        (void*)mu, (void*)nu, (void*)hlog, (void*)hcol, (void*)hcolr, (void*)vpart, (void*)dV,
                        (void*)ctl, (void*)hstep, (void*)eparts})
            (void)hipFree(p);
        ckpt = seg = Y = Yseg = kk = c32 = nullptr;
        c64 = lam = mu = nu = hlog = hcol = hcolr = vpart = dV = hstep = eparts = nullptr;
        ctl = nullptr;
    }
    ~Adj()
    {
        free_batch();
        (void)hipFree(d_lit); (void)hipFree(d_csr_ptr); (void)hipFree(d_csr_slot); (void)hipFree(d_slot_pos);
        if (stream) (void)hipStreamDestroy(stream);
    }
    Dev dev() const { return Dev{d_lit, d_csr_ptr, d_csr_slot, d_slot_pos, n, m, k, D, L, ntraj}; }
};

static void set_problem(Adj& A, const int* lits, const int* lens, int nclauses)
{
    require(nclauses > 0 && lits && lens, ASAT_ERR_INVALID_ARGUMENT, "empty problem");
    int n = 0, m = 0, k = 0;
    size_t off = 0;
    for (int j = 0; j < nclauses; j++) {
        int len1 = 0;
        for (int i = 0; i < lens[j]; i++)
            if (lits[off + i]) { n = std::max(n, std::abs(lits[off + i])); len1++; }
        if (len1) { m++; k = std::max(k, len1); }
        off += (size_t)lens[j];
    }
    require(m > 0, ASAT_ERR_INVALID_ARGUMENT, "empty problem");
    require(k == 3 || k == 4, ASAT_ERR_UNSUPPORTED, "the adjoint supports k = 3 or 4 (longest clause)");
    std::vector<int> C((size_t)m * k, 0);
    off = 0;
    int row = 0;
    for (int j = 0; j < nclauses; j++) {
        int q = 0;
        for (int i = 0; i < lens[j]; i++)
            if (lits[off + i]) C[(size_t)row * k + q++] = lits[off + i];
        if (q) row++;
        off += (size_t)lens[j];
    }
    std::vector<int> ptr((size_t)n + 2, 0);
    for (int l : C)
        if (l) ptr[(size_t)std::abs(l) + 1]++;
    for (int i = 0; i <= n; i++) ptr[i + 1] += ptr[i];
    // var-major position of every slot: variables in order, each variable's
    // slots ascending (the reference's accumulation order)
    std::vector<int> spos(std::max<size_t>(C.size(), 1), -1);
    std::vector<int> slot(std::max(ptr[n + 1], 1));
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    for (size_t s = 0; s < C.size(); s++)
        if (C[s]) {
            spos[s] = fill[std::abs(C[s])]++;
            slot[(size_t)spos[s]] = (int)s;
        }
    if (!A.stream) HIP_CHECK(hipStreamCreateWithFlags(&A.stream, hipStreamNonBlocking));
    HIP_CHECK(hipStreamSynchronize(A.stream));
    A.free_batch();
    (void)hipFree(A.d_lit); (void)hipFree(A.d_csr_ptr); (void)hipFree(A.d_csr_slot); (void)hipFree(A.d_slot_pos);
    HIP_CHECK(hipMalloc((void**)&A.d_lit, sizeof(int) * C.size()));
    HIP_CHECK(hipMalloc((void**)&A.d_csr_ptr, sizeof(int) * ptr.size()));
    HIP_CHECK(hipMalloc((void**)&A.d_csr_slot, sizeof(int) * slot.size()));
    HIP_CHECK(hipMalloc((void**)&A.d_slot_pos, sizeof(int) * spos.size()));
    HIP_CHECK(hipMemcpy(A.d_lit, C.data(), sizeof(int) * C.size(), hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(A.d_csr_ptr, ptr.data(), sizeof(int) * ptr.size(), hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(A.d_csr_slot, slot.data(), sizeof(int) * slot.size(), hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(A.d_slot_pos, spos.data(), sizeof(int) * spos.size(), hipMemcpyHostToDevice));
    A.n = n; A.m = m; A.k = k;
    A.D = (long long)((n + m + 1 + 63) / 64) * 64;
    A.L = (long long)m * k;
    A.lits_in.assign(lits, lits + off);
    A.lens_in.assign(lens, lens + nclauses);
    A.have_problem = true;
    A.have_forward = false;
}

static void alloc_batch(Adj& A, int ntraj)
{
    A.free_batch();
    A.ntraj = ntraj;
    A.ncheck = (A.T + A.C - 1) / A.C + 1;
    const size_t V = (size_t)ntraj * A.D;
    HIP_CHECK(hipMalloc((void**)&A.ckpt, sizeof(float) * V * A.ncheck));
    HIP_CHECK(hipMalloc((void**)&A.seg, sizeof(float) * V * (A.C + 1)));
    HIP_CHECK(hipMalloc((void**)&A.Y, sizeof(float) * V * 6));
    HIP_CHECK(hipMalloc((void**)&A.Yseg, sizeof(float) * V * 6 * A.C));
    HIP_CHECK(hipMalloc((void**)&A.kk, sizeof(float) * V * 6));
    HIP_CHECK(hipMalloc((void**)&A.c32, sizeof(float) * (size_t)ntraj * A.L));
    HIP_CHECK(hipMalloc((void**)&A.c64, sizeof(double) * (size_t)ntraj * A.L));
    HIP_CHECK(hipMalloc((void**)&A.lam, sizeof(double) * V));
    HIP_CHECK(hipMalloc((void**)&A.mu, sizeof(double) * V));
    HIP_CHECK(hipMalloc((void**)&A.nu, sizeof(double) * V * 6));
    HIP_CHECK(hipMalloc((void**)&A.hlog, sizeof(double) * (size_t)ntraj * std::max(A.T, 1)));
    HIP_CHECK(hipMalloc((void**)&A.hcol, sizeof(double) * (size_t)ntraj * std::max(A.T, 1)));
    HIP_CHECK(hipMalloc((void**)&A.hcolr, sizeof(double) * (size_t)ntraj * std::max(A.T, 1)));
    const int nblk = ceil_div(A.m, BLK);
    HIP_CHECK(hipMalloc((void**)&A.vpart, sizeof(double) * (size_t)ntraj * nblk));
    HIP_CHECK(hipMalloc((void**)&A.dV, sizeof(double) * ntraj));
    HIP_CHECK(hipMalloc((void**)&A.ctl, sizeof(TCtl) * ntraj));
    HIP_CHECK(hipMalloc((void**)&A.hstep, sizeof(double) * ntraj));
    HIP_CHECK(hipMalloc((void**)&A.eparts, sizeof(double) * (size_t)ntraj * (ceil_div(A.m, BLK) + ceil_div(A.n, BLK))));
    HIP_CHECK(hipMemset(A.seg, 0, sizeof(float) * V * (A.C + 1)));
    HIP_CHECK(hipMemset(A.Y, 0, sizeof(float) * V * 6));
    HIP_CHECK(hipMemset(A.kk, 0, sizeof(float) * V * 6));
    HIP_CHECK(hipMemset(A.ckpt, 0, sizeof(float) * V * A.ncheck));
}

// f = RHS(Y) for all trajectories (FP32, forward arithmetic)
template <typename T>
static void rhs_all(Adj& A, const T* Yv, T* out, T* contrib)
{
    const Dev d = A.dev();
    dim3 gc(ceil_div(A.m, BLK), ntraj8(A.ntraj)), gv(ceil_div(A.n, BLK), ntraj8(A.ntraj));
    if (A.k == 3) hipLaunchKernelGGL((k_rhs_clause<T, 3>), gc, dim3(BLK), 0, A.stream, d, Yv, out, contrib);
    else hipLaunchKernelGGL((k_rhs_clause<T, 4>), gc, dim3(BLK), 0, A.stream, d, Yv, out, contrib);
    hipLaunchKernelGGL(k_rhs_var<T>, gv, dim3(BLK), 0, A.stream, d, (const T*)contrib, out);
}

// one frozen Cash-Karp step for every trajectory with step sizes hcol;
// fills Y[0..5] (stage inputs) and kk[0..5]; writes x_next if non-null
template <typename T>
static void frozen_step(Adj& A, const T* X, T* Xn, T* Y, T* kk, T* contrib)
{
    const Dev d = A.dev();
    const size_t V = (size_t)A.ntraj * A.D;
    T* Yl[6];
    T* kl[6];
    for (int i = 0; i < 6; i++) { Yl[i] = Y + i * V; kl[i] = kk + i * V; }
    dim3 ge(ceil_div((long long)A.n + A.m, BLK), ntraj8(A.ntraj));
    HIP_CHECK(hipMemcpyAsync(Yl[0], X, sizeof(T) * V, hipMemcpyDeviceToDevice, A.stream));
    rhs_all<T>(A, Yl[0], kl[0], contrib);
    hipLaunchKernelGGL((k_stage<T, 2>), ge, dim3(BLK), 0, A.stream, d, X, kl[0], kl[1], kl[2], kl[3], kl[4], A.hcur, Yl[1]);
    rhs_all<T>(A, Yl[1], kl[1], contrib);
    hipLaunchKernelGGL((k_stage<T, 3>), ge, dim3(BLK), 0, A.stream, d, X, kl[0], kl[1], kl[2], kl[3], kl[4], A.hcur, Yl[2]);
    rhs_all<T>(A, Yl[2], kl[2], contrib);
    hipLaunchKernelGGL((k_stage<T, 4>), ge, dim3(BLK), 0, A.stream, d, X, kl[0], kl[1], kl[2], kl[3], kl[4], A.hcur, Yl[3]);
    rhs_all<T>(A, Yl[3], kl[3], contrib);
    hipLaunchKernelGGL((k_stage<T, 5>), ge, dim3(BLK), 0, A.stream, d, X, kl[0], kl[1], kl[2], kl[3], kl[4], A.hcur, Yl[4]);
    rhs_all<T>(A, Yl[4], kl[4], contrib);
    hipLaunchKernelGGL((k_stage<T, 6>), ge, dim3(BLK), 0, A.stream, d, X, kl[0], kl[1], kl[2], kl[3], kl[4], A.hcur, Yl[5]);
    rhs_all<T>(A, Yl[5], kl[5], contrib);
    if (Xn) {
        dim3 gf(ceil_div((long long)A.n + A.m + 1, BLK), ntraj8(A.ntraj));
        hipLaunchKernelGGL(k_final<T>, gf, dim3(BLK), 0, A.stream, d, X, kl[0], kl[2], kl[3], kl[5], A.hcur, Xn);
    }
}


// V (and grad V into lam if want_grad) at X for every trajectory
template <typename TX>
static void energy(Adj& A, const TX* X, bool want_grad)
{
    const Dev d = A.dev();
    const int nblk = ceil_div(A.m, BLK);
    dim3 gc(nblk, ntraj8(A.ntraj)), gv(ceil_div(A.n, BLK), ntraj8(A.ntraj));
    double* g = A.lam;
    if (A.k == 3) hipLaunchKernelGGL((k_gradv_clause<TX, 3>), gc, dim3(BLK), 0, A.stream, d, X, g, A.c64, A.vpart, nblk);
    else hipLaunchKernelGGL((k_gradv_clause<TX, 4>), gc, dim3(BLK), 0, A.stream, d, X, g, A.c64, A.vpart, nblk);
    if (want_grad) {
        hipLaunchKernelGGL(k_rhs_var<double>, gv, dim3(BLK), 0, A.stream, d, (const double*)A.c64, g);
        hipLaunchKernelGGL(k_zero_slot0, dim3(ceil_div(A.ntraj, 64)), dim3(64), 0, A.stream, g, A.D, A.ntraj);
    }
    hipLaunchKernelGGL(k_sum_parts, dim3(A.ntraj), dim3(64), 0, A.stream, (const double*)A.vpart, nblk, A.dV);
}

// one attempt (adaptive) or one frozen step of every trajectory: six fused stages
template <int K, bool FZ>
static void fused_stages(Adj& A, const FArgs& fa)
{
    const dim3 gc(fa.nblk_c, ntraj8(A.ntraj)), gv(fa.nblk_v, ntraj8(A.ntraj));
    hipLaunchKernelGGL((k_fclause<K, 1, FZ>), gc, dim3(BLK), 0, A.stream, fa);
    hipLaunchKernelGGL((k_fvar<1, FZ>), gv, dim3(BLK), 0, A.stream, fa);
    hipLaunchKernelGGL((k_fclause<K, 2, FZ>), gc, dim3(BLK), 0, A.stream, fa);
    hipLaunchKernelGGL((k_fvar<2, FZ>), gv, dim3(BLK), 0, A.stream, fa);
    hipLaunchKernelGGL((k_fclause<K, 3, FZ>), gc, dim3(BLK), 0, A.stream, fa);
    hipLaunchKernelGGL((k_fvar<3, FZ>), gv, dim3(BLK), 0, A.stream, fa);
    hipLaunchKernelGGL((k_fclause<K, 4, FZ>), gc, dim3(BLK), 0, A.stream, fa);
    hipLaunchKernelGGL((k_fvar<4, FZ>), gv, dim3(BLK), 0, A.stream, fa);
    hipLaunchKernelGGL((k_fclause<K, 5, FZ>), gc, dim3(BLK), 0, A.stream, fa);
    hipLaunchKernelGGL((k_fvar<5, FZ>), gv, dim3(BLK), 0, A.stream, fa);
    hipLaunchKernelGGL((k_fclause<K, 6, FZ>), gc, dim3(BLK), 0, A.stream, fa);
    hipLaunchKernelGGL((k_fvar<6, FZ>), gv, dim3(BLK), 0, A.stream, fa);
}

static FArgs fused_args(Adj& A, const float* X, float* Xn)
{
    FArgs fa;
    memset(&fa, 0, sizeof(fa));
    fa.d = A.dev();
    fa.X = X;
    fa.Xn = Xn;
    const size_t V = (size_t)A.ntraj * A.D;
    for (int i = 0; i < 6; i++) fa.k[i] = A.kk + i * V;
    fa.V = (long long)V;
    fa.c = A.c32;
    fa.nblk_c = ceil_div(A.m, BLK);
    fa.nblk_v = ceil_div(A.n, BLK);
    return fa;
}

// the FP32 frozen recompute of one step with the current h column: the six
// stage inputs into Y (the adjoint's linearisation points), x_next into Xn
// unless null
static void frozen_step_fused(Adj& A, const float* X, float* Xn, float* Y)
{
    FArgs fa = fused_args(A, X, Xn);
    fa.Y = Y;
    fa.hcol = A.hcur;
    if (A.k == 3) fused_stages<3, true>(A, fa);
    else fused_stages<4, true>(A, fa);
}

// Adaptive forward of every trajectory in lockstep: T accepted steps each,
// step sizes logged in hlog, checkpoints c = 0..ceil(T/C) (x at step c*C,
// the last one x_T).  X0 holds the start states in the internal layout.
static void forward_lockstep(Adj& A, const std::vector<float>& X0)
{
    const Dev d = A.dev();
    const size_t V = (size_t)A.ntraj * A.D;
    float* X = A.seg;          // seg has C + 1 >= 2 slots
    float* Xn = A.seg + V;
    HIP_CHECK(hipMemcpyAsync(X, X0.data(), sizeof(float) * V, hipMemcpyHostToDevice, A.stream));
    HIP_CHECK(hipMemcpyAsync(A.ckpt, X, sizeof(float) * V, hipMemcpyDeviceToDevice, A.stream));
    hipLaunchKernelGGL(k_ctl_init, dim3(ceil_div(A.ntraj, 64)), dim3(64), 0, A.stream, A.ctl, A.hstep, A.ntraj);
    A.hcur = A.hstep;
    const dim3 ge(ceil_div((long long)A.n + A.m + 1, BLK), ntraj8(A.ntraj));
    FArgs fa = fused_args(A, X, Xn);
    fa.S = A.Y;
    fa.ctl = A.ctl;
    fa.parts = A.eparts;
    const int nblk = fa.nblk_c + fa.nblk_v;
    std::vector<TCtl> hc((size_t)A.ntraj);
    const int per_round = 16;
    // every trajectory needs at least T attempts; the round loop ends when none is active
    for (long long round = 0;; round++) {
        for (int r = 0; r < per_round; r++) {
            if (A.k == 3) fused_stages<3, false>(A, fa);
            else fused_stages<4, false>(A, fa);
            hipLaunchKernelGGL(k_ctl_batch, dim3(A.ntraj), dim3(BLK), 0, A.stream, A.ctl, (const double*)A.eparts,
                               nblk, A.eps, A.T, A.C, A.hlog, A.hstep);
            hipLaunchKernelGGL(k_accept, ge, dim3(BLK), 0, A.stream, d, (const TCtl*)A.ctl, (const float*)Xn, X,
                               A.ckpt, (long long)A.ntraj);
        }
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipMemcpyAsync(hc.data(), A.ctl, sizeof(TCtl) * hc.size(), hipMemcpyDeviceToHost, A.stream));
        HIP_CHECK(hipStreamSynchronize(A.stream));
        int active = 0;
        long long attempts = 0;
        for (const TCtl& c : hc) {
            if (c.underflow) throw asat::Error(ASAT_UNDERFLOW, "step size underflow in the forward run");
            active += c.active;
            attempts += c.attempts;
        }
        A.fwd_attempts = attempts;
        A.fwd_lockstep = (round + 1) * per_round;
        if (!active) break;
        // each round accepts or rejects every active trajectory per attempt;
        // a trajectory that needs more than 64 attempts per step is stuck
        require(round * per_round < 64LL * A.T + 64, ASAT_ERR_HIP, "forward run made no progress");
    }
}

// step sizes of step s for every trajectory: hT[s][traj] (and FP32-rounded hTr)
__global__ void k_transpose_h(const double* hlog, int ntraj, int T, double* hT, double* hTr)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)ntraj * T) return;
    const int t = (int)(i / T), s = (int)(i % T);
    const double h = hlog[i];
    hT[(long long)s * ntraj + t] = h;
    hTr[(long long)s * ntraj + t] = (double)(float)h;
}

static void make_columns(Adj& A)
{
    const long long n = (long long)A.ntraj * A.T;
    hipLaunchKernelGGL(k_transpose_h, dim3(ceil_div(n, 256)), dim3(256), 0, A.stream, (const double*)A.hlog,
                       A.ntraj, A.T, A.hcol, A.hcolr);
}

static void load_h(Adj& A, int s)
{
    A.hcur = A.hcol + (size_t)s * A.ntraj;
    A.hcurr = A.hcolr + (size_t)s * A.ntraj;
}

}  // namespace adj
}  // namespace asat

// ===========================================================================
// C ABI
// ===========================================================================
using namespace asat;
using namespace asat::adj;

struct asat_adj { Adj a; };

#define AAPI_BEGIN try {
#define AAPI_END                                                 \
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

int asat_adj_create(asat_adj_t** out)
{
    AAPI_BEGIN
    require(out != nullptr, ASAT_ERR_INVALID_ARGUMENT, "null argument: out");
    *out = new asat_adj();
    AAPI_END
}

int asat_adj_destroy(asat_adj_t* a)
{
    AAPI_BEGIN
    delete a;
    AAPI_END
}

int asat_adj_set_problem(asat_adj_t* a, const int* literals, const int* lengths, int nclauses)
{
    AAPI_BEGIN
    require(a != nullptr, ASAT_ERR_INVALID_ARGUMENT, "null adjoint");
    set_problem(a->a, literals, lengths, nclauses);
    AAPI_END
}

int asat_adj_set_params(asat_adj_t* a, double eps, int steps, int checkpoint_every)
{
    AAPI_BEGIN
    require(a != nullptr, ASAT_ERR_INVALID_ARGUMENT, "null adjoint");
    require(!(eps < 1e-16 || eps > 1.0), ASAT_ERR_INVALID_ARGUMENT, "epsilon must be between 1e-16 and 1");
    require(steps >= 1 && checkpoint_every >= 1, ASAT_ERR_INVALID_ARGUMENT, "steps and checkpoint interval must be positive");
    a->a.eps = eps; a->a.T = steps; a->a.C = checkpoint_every;
    a->a.have_forward = false;
    AAPI_END
}

int asat_adj_forward(asat_adj_t* a, int ntraj, const float* pub0, double* V_out)
{
    AAPI_BEGIN
    require(a && pub0 && ntraj > 0, ASAT_ERR_INVALID_ARGUMENT, "null argument or no trajectories");
    Adj& A = a->a;
    require(A.have_problem && A.T > 0, ASAT_ERR_NOT_CONFIGURED, "adjoint problem/parameters not set");
    alloc_batch(A, ntraj);
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    HIP_CHECK(hipEventRecord(e0, A.stream));
    const size_t V = (size_t)A.ntraj * A.D;
    std::vector<float> x0(V, 0.f);
    for (int t = 0; t < ntraj; t++) {
        x0[(size_t)t * A.D] = -1.0f;
        std::copy(pub0 + (size_t)t * (A.n + A.m), pub0 + (size_t)(t + 1) * (A.n + A.m), x0.begin() + (size_t)t * A.D + 1);
    }
    forward_lockstep(A, x0);
    make_columns(A);
    // V(T) from the last checkpoint (= x_T)
    const int last = (A.T + A.C - 1) / A.C;
    energy(A, A.ckpt + (size_t)last * V, false);
    HIP_CHECK(hipEventRecord(e1, A.stream));
    HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    A.fwd_ms = ms;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (V_out) HIP_CHECK(hipMemcpy(V_out, A.dV, sizeof(double) * ntraj, hipMemcpyDeviceToHost));
    A.have_forward = true;
    AAPI_END
}

int asat_adj_get_steps(const asat_adj_t* a, double* h_out)
{
    AAPI_BEGIN
    require(a && h_out, ASAT_ERR_INVALID_ARGUMENT, "null argument");
    require(a->a.have_forward, ASAT_ERR_NOT_CONFIGURED, "no forward run");
    HIP_CHECK(hipMemcpy(h_out, a->a.hlog, sizeof(double) * (size_t)a->a.ntraj * a->a.T, hipMemcpyDeviceToHost));
    AAPI_END
}

int asat_adj_replay(asat_adj_t* a, int ntraj, const float* pub0, const double* h, double* V_out, float* xT_out)
{
    AAPI_BEGIN
    require(a && pub0 && h && ntraj > 0, ASAT_ERR_INVALID_ARGUMENT, "null argument or no trajectories");
    Adj& A = a->a;
    require(A.have_problem && A.T > 0, ASAT_ERR_NOT_CONFIGURED, "adjoint problem/parameters not set");
    alloc_batch(A, ntraj);
    const size_t V = (size_t)ntraj * A.D;
    HIP_CHECK(hipMemcpy(A.hlog, h, sizeof(double) * (size_t)ntraj * A.T, hipMemcpyHostToDevice));
    make_columns(A);
    std::vector<float> x0(V, 0.f);
    for (int t = 0; t < ntraj; t++) {
        x0[(size_t)t * A.D] = -1.0f;
        std::copy(pub0 + (size_t)t * (A.n + A.m), pub0 + (size_t)(t + 1) * (A.n + A.m), x0.begin() + (size_t)t * A.D + 1);
    }
    HIP_CHECK(hipMemcpy(A.seg, x0.data(), sizeof(float) * V, hipMemcpyHostToDevice));
    float* cur = A.seg;
    float* nxt = A.seg + V;
    for (int s = 0; s < A.T; s++) {
        load_h(A, s);
        frozen_step_fused(A, cur, nxt, A.Y);
        std::swap(cur, nxt);
    }
    energy(A, cur, false);
    HIP_CHECK(hipStreamSynchronize(A.stream));
    if (V_out) HIP_CHECK(hipMemcpy(V_out, A.dV, sizeof(double) * ntraj, hipMemcpyDeviceToHost));
    if (xT_out) {
        std::vector<float> xt(V);
        HIP_CHECK(hipMemcpy(xt.data(), cur, sizeof(float) * V, hipMemcpyDeviceToHost));
        for (int t = 0; t < ntraj; t++)
            std::copy(xt.begin() + (size_t)t * A.D + 1, xt.begin() + (size_t)t * A.D + 1 + A.n + A.m,
                      xT_out + (size_t)t * (A.n + A.m));
    }
    A.have_forward = false;
    AAPI_END
}

int asat_adj_replay_f64(asat_adj_t* a, int ntraj, const double* pub0, const double* h, double* V_out)
{
    AAPI_BEGIN
    require(a && pub0 && h && V_out && ntraj > 0, ASAT_ERR_INVALID_ARGUMENT, "null argument or no trajectories");
    Adj& A = a->a;
    require(A.have_problem && A.T > 0, ASAT_ERR_NOT_CONFIGURED, "adjoint problem/parameters not set");
    alloc_batch(A, ntraj);
    const size_t V = (size_t)ntraj * A.D;
    HIP_CHECK(hipMemcpy(A.hlog, h, sizeof(double) * (size_t)ntraj * A.T, hipMemcpyHostToDevice));
    make_columns(A);
    std::vector<double> x0(V, 0.0);
    for (int t = 0; t < ntraj; t++) {
        x0[(size_t)t * A.D] = -1.0;
        std::copy(pub0 + (size_t)t * (A.n + A.m), pub0 + (size_t)(t + 1) * (A.n + A.m), x0.begin() + (size_t)t * A.D + 1);
    }
    // FP64 state: the current and next state, six stage inputs, six stage derivatives
    double* buf = nullptr;
    HIP_CHECK(hipMalloc((void**)&buf, sizeof(double) * V * 14));
    double* cur = buf;
    double* nxt = buf + V;
    HIP_CHECK(hipMemcpy(cur, x0.data(), sizeof(double) * V, hipMemcpyHostToDevice));
    for (int s = 0; s < A.T; s++) {
        load_h(A, s);
        frozen_step<double>(A, cur, nxt, buf + 2 * V, buf + 8 * V, A.c64);
        std::swap(cur, nxt);
    }
    energy<double>(A, cur, false);
    HIP_CHECK(hipStreamSynchronize(A.stream));
    HIP_CHECK(hipMemcpy(V_out, A.dV, sizeof(double) * ntraj, hipMemcpyDeviceToHost));
    HIP_CHECK(hipFree(buf));
    A.have_forward = false;
    AAPI_END
}

int asat_adj_backward(asat_adj_t* a, double* grad_out)
{
    AAPI_BEGIN
    require(a && grad_out, ASAT_ERR_INVALID_ARGUMENT, "null argument");
    Adj& A = a->a;
    require(A.have_forward, ASAT_ERR_NOT_CONFIGURED, "asat_adj_forward must run first");
    const Dev d = A.dev();
    const size_t V = (size_t)A.ntraj * A.D;
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    HIP_CHECK(hipEventRecord(e0, A.stream));
    const int nseg = (A.T + A.C - 1) / A.C;
    // lam = grad V(x_T)
    energy(A, A.ckpt + (size_t)nseg * V, true);
    dim3 ge(ceil_div((long long)A.n + A.m + 1, BLK), ntraj8(A.ntraj)), ge1(ceil_div((long long)A.n + A.m, BLK), ntraj8(A.ntraj));
    dim3 gc(ceil_div(A.m, BLK), ntraj8(A.ntraj)), gv(ceil_div(A.n, BLK), ntraj8(A.ntraj));
    for (int sg = nseg - 1; sg >= 0; sg--) {
        const int s0 = sg * A.C, s1 = std::min(A.T, s0 + A.C);
        // recompute the segment's states from its checkpoint
        HIP_CHECK(hipMemcpyAsync(A.seg, A.ckpt + (size_t)sg * V, sizeof(float) * V, hipMemcpyDeviceToDevice, A.stream));
        // ... keeping every step's six stage inputs (6 V floats per step)
        for (int s = s0; s < s1; s++) {
            load_h(A, s);
            frozen_step_fused(A, A.seg + (size_t)(s - s0) * V, s < s1 - 1 ? A.seg + (size_t)(s - s0 + 1) * V : nullptr,
                              A.Yseg + (size_t)(s - s0) * 6 * V);
        }
        for (int s = s1 - 1; s >= s0; s--) {
            load_h(A, s);
            for (int l = 6; l >= 1; l--) {
                hipLaunchKernelGGL(k_mu, ge, dim3(BLK), 0, A.stream, d, l, (const double*)A.lam,
                                   (const double*)A.nu, A.hcurr, A.mu);
                double* nul = A.nu + (size_t)(l - 1) * V;
                const float* Yl = A.Yseg + ((size_t)(s - s0) * 6 + (l - 1)) * V;
                if (A.k == 3) hipLaunchKernelGGL(k_vjp_clause<3>, gc, dim3(BLK), 0, A.stream, d, Yl, (const double*)A.mu, nul, A.c64);
                else hipLaunchKernelGGL(k_vjp_clause<4>, gc, dim3(BLK), 0, A.stream, d, Yl, (const double*)A.mu, nul, A.c64);
                hipLaunchKernelGGL(k_rhs_var<double>, gv, dim3(BLK), 0, A.stream, d, (const double*)A.c64, nul);
            }
            hipLaunchKernelGGL(k_lam, ge1, dim3(BLK), 0, A.stream, d, A.lam, (const double*)A.nu);
        }
    }
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipEventRecord(e1, A.stream));
    HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    A.bwd_ms = ms;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    std::vector<double> lam(V);
    HIP_CHECK(hipMemcpy(lam.data(), A.lam, sizeof(double) * V, hipMemcpyDeviceToHost));
    for (int t = 0; t < A.ntraj; t++)
        std::copy(lam.begin() + (size_t)t * A.D + 1, lam.begin() + (size_t)t * A.D + 1 + A.n + A.m,
                  grad_out + (size_t)t * (A.n + A.m));
    AAPI_END
}

int asat_adj_timing(const asat_adj_t* a, double* forward_ms, double* backward_ms)
{
    AAPI_BEGIN
    require(a != nullptr, ASAT_ERR_INVALID_ARGUMENT, "null adjoint");
    if (forward_ms) *forward_ms = a->a.fwd_ms;
    if (backward_ms) *backward_ms = a->a.bwd_ms;
    AAPI_END
}

int asat_adj_attempts(const asat_adj_t* a, long long* trajectory_attempts, long long* lockstep_attempts)
{
    AAPI_BEGIN
    require(a != nullptr, ASAT_ERR_INVALID_ARGUMENT, "null adjoint");
    if (trajectory_attempts) *trajectory_attempts = a->a.fwd_attempts;
    if (lockstep_attempts) *lockstep_attempts = a->a.fwd_lockstep;
    AAPI_END
}

}  // extern "C"
