// Shared definitions of the AnalogSat MI355X library: error handling,
// the device-side step-control block, the Cash-Karp tableau, wave64 helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "analogsat_hip.h"

#define ASAT_HD __host__ __device__
#include "asat_powctrl.h"

namespace asat {

constexpr int GEN_KMAX = 32;      // longest clause supported by the generic kernels

// ---------------------------------------------------------------------------
// errors: every failing HIP call becomes an exception carrying a status code;
// the extern "C" layer turns it into a return code + thread-local message.
// ---------------------------------------------------------------------------
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void hip_check(hipError_t e, const char* what, const char* file, int line)
{
    if (e != hipSuccess) {
        char buf[512];
        snprintf(buf, sizeof(buf), "HIP error %d (%s) in %s at %s:%d", (int)e, hipGetErrorString(e), what, file,
                 line);
        throw Error(ASAT_ERR_HIP, buf);
    }
}
#define HIP_CHECK(x) ::asat::hip_check((x), #x, __FILE__, __LINE__)

inline void require(bool ok, int code, const char* msg)
{
    if (!ok) throw Error(code, msg);
}

// ---------------------------------------------------------------------------
// device-resident step control for one trajectory (plain solver).  Written by
// the control kernel only; every stage kernel reads it at entry.
// ---------------------------------------------------------------------------
struct alignas(64) Ctrl {
    double h;          // step size of the current attempt
    double hnext;      // step size proposed at the start of the current step (xscal uses it)
    double time;       // analog time
    double hdid;       // last accepted step
    double errmax;     // last error norm / eps
    double eps;
    double floor;      // constant term of xscal (1e-20 = reference TINY)
    double* hlog;      // optional: accepted step sizes, hlog[steps-1] = hdid
    long long hlog_len;
    long long steps;      // accepted steps
    long long attempts;   // attempts incl. rejected
    long long rhscount;   // RHS evaluations
    long long target;     // stop once steps reaches this
    int parity;        // which of the two state buffers holds x
    int step_start;    // 1: this attempt begins a new step (k1 must be computed)
    int active;        // 0 once target reached or underflow
    int underflow;
    int violations;    // scratch for the violated-clause count
    int pad_[3];
};

// ---------------------------------------------------------------------------
// Cash-Karp tableau (Numerical Recipes in C, 2nd ed. rkck; reference values
// AnalogSAT/cpu_sat/CpuRungeKutta.h:83-112).  State-type values are the
// double constants rounded to T, the error coefficients are double.
// ---------------------------------------------------------------------------
template <typename T>
struct Tab {
    static constexpr T b21 = (T)0.2;
    static constexpr T b31 = (T)(3.0 / 40.0), b32 = (T)(9.0 / 40.0);
    static constexpr T b41 = (T)0.3, b42 = (T)-0.9, b43 = (T)1.2;
    static constexpr T b51 = (T)(-11.0 / 54.0), b52 = (T)2.5, b53 = (T)(-70.0 / 27.0), b54 = (T)(35.0 / 27.0);
    static constexpr T b61 = (T)(1631.0 / 55296.0), b62 = (T)(175.0 / 512.0), b63 = (T)(575.0 / 13824.0),
                       b64 = (T)(44275.0 / 110592.0), b65 = (T)(253.0 / 4096.0);
    static constexpr T c1 = (T)(37.0 / 378.0), c3 = (T)(250.0 / 621.0), c4 = (T)(125.0 / 594.0),
                       c6 = (T)(512.0 / 1771.0);
};
struct TabD {
    static constexpr double c1 = 37.0 / 378.0, c3 = 250.0 / 621.0, c4 = 125.0 / 594.0, c6 = 512.0 / 1771.0;
    static constexpr double dc5 = -277.0 / 14336.0;
    static constexpr double dc1 = c1 - 2825.0 / 27648.0, dc3 = c3 - 18575.0 / 48384.0,
                            dc4 = c4 - 13525.0 / 55296.0, dc6 = c6 - 0.25;
    static constexpr double TINY = 1e-20, SAFETY = 0.9, PGROW = -0.2, PSHRNK = -0.25, ERRCON = 1.89e-4,
                            HFACTOR = 0.1;
};

// stage input of stage J (2..6) from x and k1..k(J-1), in the reference's
// operation order (AnalogSAT/cpu_sat/CpuRungeKutta.cpp:190-210)
template <typename T, int J>
__device__ __forceinline__ T stage_input(T x, T k1, T k2, T k3, T k4, T k5, T hs)
{
    using B = Tab<T>;
    if constexpr (J == 2) return x + B::b21 * hs * k1;
    else if constexpr (J == 3) return x + hs * (B::b31 * k1 + B::b32 * k2);
    else if constexpr (J == 4) return x + hs * (B::b41 * k1 + B::b42 * k2 + B::b43 * k3);
    else if constexpr (J == 5) return x + hs * (B::b51 * k1 + B::b52 * k2 + B::b53 * k3 + B::b54 * k4);
    else return x + hs * (B::b61 * k1 + B::b62 * k2 + B::b63 * k3 + B::b64 * k4 + B::b65 * k5);
}

// final combination (xnew) and error term |xerr / xscal| of one element
// floor: constant term of xscal, TINY = 1e-20 in the reference
// (AnalogSAT/cpu_sat/CpuRungeKutta.cpp:84-85)
template <typename T>
__device__ __forceinline__ void final_combination(T x, T k1, T k3, T k4, T k5, T k6, T hs, double h, double hnext,
                                                  double floor, T& xnew, double& err)
{
    using B = Tab<T>;
    xnew = x + hs * (B::c1 * k1 + B::c3 * k3 + B::c4 * k4 + B::c6 * k6);
    double xerr = h * (TabD::dc1 * (double)k1 + TabD::dc3 * (double)k3 + TabD::dc4 * (double)k4 +
                       TabD::dc5 * (double)k5 + TabD::dc6 * (double)k6);
    double xscal = fabs((double)x) + fabs((double)k1 * hnext) + floor;
    err = fabs(xerr / xscal);
}

// ---------------------------------------------------------------------------
// per-clause arithmetic shared by every path (reference order:
// AnalogSAT/cpu_sat/CpuSat.cpp:206-229, AnalogSAT/cpu_sat/cpu_helper.h:23-113)
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T spin_factor(int l, const T* sp)
{
    T s = l != 0 ? sp[l < 0 ? -l : l] : (T)-1.0;
    return l < 0 ? (T)1.0 + s : (T)1.0 - s;
}

// product of f[0..K-1] except position Q, left to right (reference KmiProd order)
template <typename T, int K, int Q>
__device__ __forceinline__ T kmi_prod(const T (&f)[K])
{
    T p = (T)1.0;
    bool first = true;
#pragma unroll
    for (int q = 0; q < K; q++) {
        if (q == Q) continue;
        p = first ? f[q] : p * f[q];
        first = false;
    }
    return p;
}

template <typename T, int K, int Q>
__device__ __forceinline__ void write_contribs(const int (&l)[K], const T (&f)[K], T r, T* out)
{
    if constexpr (Q < K) {
        if (l[Q] != 0) {
            T rr = r * kmi_prod<T, K, Q>(f);
            rr *= l[Q] < 0 ? (T)-2.0 : (T)2.0;
            out[Q] = rr;
            write_contribs<T, K, Q + 1>(l, f, r, out);
        }
    }
}

// tanh-family slot term, knorm a |prod| tanh(q f) c (AnalogSAT/cpu_sat/
// cpu_helper.h:115-125, AnalogSAT/cpu_sat/CpuSatTanh.cpp:262-275), r = knorm a.
// For K = 3..10 the reference's templated helper calls the C library's double
// tanh (it is declared before the file's using-directive), so in FP32 that
// product is formed in double and rounded once; otherwise the float overload.
template <typename T>
__device__ __forceinline__ T tanh_term(T r, T prod, T f, T q, bool neg, bool ex)
{
    const T ap = prod < (T)0 ? -prod : prod;
    T rr;
    if constexpr (sizeof(T) == 4) {
        if (ex) rr = (T)((double)(r * ap) * tanh((double)(q * f)));
        else rr = r * ap * tanhf(q * f);
    } else {
        (void)ex;
        rr = r * ap * tanh(q * f);
    }
    return neg ? -rr : rr;
}

// ---------------------------------------------------------------------------
// wave64 / block reductions (DPP-free, via __shfl_xor on 64 lanes)
// fmax ignores NaN like the CPU reference's std::fmax.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// A lane permutation within rows of 16 lanes (DPP), for a double.  Every
// lane of the wave must be active.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, 0xF, 0xF, false);
    const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

__device__ __forceinline__ double readlane_f64(double v, int lane)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, lane);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), lane);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// wave_max without LDS round trips: within each row of 16 lanes by DPP
// (swap neighbours, swap pairs, half-row mirror, row mirror), then the four
// row maxima read from lanes 0, 16, 32 and 48.  Same value as wave_max (the
// maximum does not depend on the order), wave-uniform.  Every lane must be
// active.
__device__ __forceinline__ double wave_max_dpp(double v)
{
    v = fmax(v, dpp_f64<0xB1>(v));    // quad_perm [1,0,3,2]
    v = fmax(v, dpp_f64<0x4E>(v));    // quad_perm [2,3,0,1]
    v = fmax(v, dpp_f64<0x141>(v));   // row_half_mirror
    v = fmax(v, dpp_f64<0x140>(v));   // row_mirror
    return fmax(fmax(readlane_f64(v, 0), readlane_f64(v, 16)), fmax(readlane_f64(v, 32), readlane_f64(v, 48)));
}

__device__ __forceinline__ int wave_sum(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

template <int BLOCK>
__device__ __forceinline__ double block_max(double v, double* sh)
{
    v = wave_max(v);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) sh[w] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x < 64) {
        r = threadIdx.x < BLOCK / 64 ? sh[threadIdx.x] : 0.0;
        r = wave_max(r);
    }
    __syncthreads();
    return r;   // valid in thread 0
}

template <int BLOCK>
__device__ __forceinline__ int block_sum(int v, int* sh)
{
    v = wave_sum(v);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) sh[w] = v;
    __syncthreads();
    int r = 0;
    if (threadIdx.x < 64) {
        r = threadIdx.x < BLOCK / 64 ? sh[threadIdx.x] : 0;
        r = wave_sum(r);
    }
    __syncthreads();
    return r;
}

// block reductions whose result every thread receives (all threads of the
// block must call them; no thread may have exited)
template <int BLOCK>
__device__ __forceinline__ double block_max_all(double v, double* sh)
{
    v = wave_max(v);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) sh[w] = v;
    __syncthreads();
    double r = sh[0];
#pragma unroll
    for (int i = 1; i < BLOCK / 64; i++) r = fmax(r, sh[i]);
    __syncthreads();
    return r;
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// deterministic block sum (fixed butterfly and wave order), in every thread
template <int BLOCK>
__device__ __forceinline__ double block_sum_all(double v, double* sh)
{
    v = wave_sum(v);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) sh[w] = v;
    __syncthreads();
    double r = sh[0];
#pragma unroll
    for (int i = 1; i < BLOCK / 64; i++) r += sh[i];
    __syncthreads();
    return r;
}

template <int BLOCK>
__device__ __forceinline__ int block_sum_all(int v, int* sh)
{
    v = wave_sum(v);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) sh[w] = v;
    __syncthreads();
    int r = 0;
#pragma unroll
    for (int i = 0; i < BLOCK / 64; i++) r += sh[i];
    __syncthreads();
    return r;
}

// Left-to-right sum s + v[e0] + ... + v[e1 - 1] (the reference's
// accumulation order), with the loads issued four at a time so that one
// memory latency covers four terms.
template <typename T>
__device__ __forceinline__ T sum_run(T s, const T* v, int e0, int e1)
{
    int e = e0;
    for (; e + 4 <= e1; e += 4) {
        const T a = v[e], b = v[e + 1], c = v[e + 2], d = v[e + 3];
        s += a;
        s += b;
        s += c;
        s += d;
    }
    for (; e < e1; e++) s += v[e];
    return s;
}

// the same through an index array: s + v[idx[e0]] + ... + v[idx[e1 - 1]]
template <typename T, typename I>
__device__ __forceinline__ T sum_run_idx(T s, const T* v, const I* idx, int e0, int e1)
{
    int e = e0;
    for (; e + 4 <= e1; e += 4) {
        const I i0 = idx[e], i1 = idx[e + 1], i2 = idx[e + 2], i3 = idx[e + 3];
        const T a = v[i0], b = v[i1], c = v[i2], d = v[i3];
        s += a;
        s += b;
        s += c;
        s += d;
    }
    for (; e < e1; e++) s += v[idx[e]];
    return s;
}

inline int ceil_div(long long a, long long b) { return (int)((a + b - 1) / b); }

}  // namespace asat
