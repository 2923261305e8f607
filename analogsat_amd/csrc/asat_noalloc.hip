// asat_solve_noalloc: asat_rk_solve for one small problem without any device
// allocation.  For callers that cannot call hipMalloc (code under HIP graph
// capture, tight serving loops, harnesses that forbid it).
//
// Every device buffer the one-workgroup kernel (asat_small.h) needs -- the
// control block, the solver block, the state, the literals, each slot's
// var-major position and each variable's range -- lives in `scratch`, one file-scope device array
// of ASAT_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends.  The
// host side never passes the array's name to a kernel or a copy: on each
// device's first call it asks hipGetSymbolAddress for that device's address
// and uses that pointer from then on.  Every call on a GPU shares its array
// (each call writes all of it that it reads), so calls that overlap on one
// GPU race; calls on different GPUs do not.
//
// Host staging uses a static host buffer of the same size under a mutex, so
// a call allocates nothing on the host either.  The clause order is the
// input order; results equal asat_rk_solve's on a system with the same
// problem, bit for bit (each variable's terms are added in the input clause
// order on both paths).
#include "asat_common.h"
#include "asat_small.h"

#include <cstring>
#include <mutex>

namespace asat {
extern thread_local std::string g_last_error;
}

using namespace asat;

static __device__ __align__(256) unsigned char scratch[ASAT_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;                                  // guards the two tables and the staging buffer
unsigned char* g_dev_scratch[MAX_DEVICES] = {};   // `scratch` of each device (hipGetSymbolAddress)
alignas(256) unsigned char g_stage[ASAT_NOALLOC_SCRATCH_BYTES];

constexpr size_t align256(size_t v) { return (v + 255) / 256 * 256; }

// the current device's address of `scratch`, looked up once per device
unsigned char* device_scratch()
{
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    require(dev >= 0 && dev < MAX_DEVICES, ASAT_ERR_UNSUPPORTED, "device index out of range");
    if (!g_dev_scratch[dev]) {
        void* p = nullptr;
        HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        require(p != nullptr, ASAT_ERR_HIP, "no device address for the scratch array");
        g_dev_scratch[dev] = (unsigned char*)p;
    }
    return g_dev_scratch[dev];
}

// sections of the scratch array (byte offsets), for one problem
struct Sections {
    size_t ctrl, solve, x, lit, vptr, spos, end;
};

Sections sections(int n, int m, int k, long long alloc, int esize)
{
    Sections s;
    s.ctrl = 0;
    s.solve = align256(sizeof(Ctrl));
    s.x = s.solve + align256(sizeof(SmallSolve));
    s.lit = s.x + align256((size_t)alloc * esize);
    s.vptr = s.lit + align256((size_t)m * k * sizeof(int));
    s.spos = s.vptr + align256((size_t)(n + 2) * sizeof(int));
    s.end = s.spos + align256((size_t)m * k * sizeof(int));
    return s;
}

template <typename T, int KT, int BS>
void launch(const SmallArgs& a, size_t lds)
{
    static const hipError_t attr = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_small<T, KT, BS, ASAT_FAMILY_CTDS>),
                                                       hipFuncAttributeMaxDynamicSharedMemorySize, SMALL_LDS_MAX);
    HIP_CHECK(attr);
    hipLaunchKernelGGL((k_small<T, KT, BS, ASAT_FAMILY_CTDS>), dim3(1), dim3(BS), lds, 0, a);
    HIP_CHECK(hipGetLastError());
}

template <typename T>
void launch_small(const SmallArgs& a, size_t lds)
{
    const bool wide = std::max(a.m, a.n) > 512;
    if (a.k == 3) wide ? launch<T, 3, 1024>(a, lds) : launch<T, 3, 256>(a, lds);
    else if (a.k == 4) wide ? launch<T, 4, 1024>(a, lds) : launch<T, 4, 256>(a, lds);
    else wide ? launch<T, 0, 1024>(a, lds) : launch<T, 0, 256>(a, lds);
}

int solve_noalloc(int precision, const int* literals, const int* lengths, int nclauses, const void* pub0,
                  double eps, int batch, long long max_steps, double max_time, int time_float, void* pub_out,
                  asat_solve_result_t* out, asat_rk_stats_t* stats)
{
    require(precision == ASAT_F64 || precision == ASAT_F32, ASAT_ERR_INVALID_ARGUMENT, "unknown precision");
    require(pub0 && pub_out && out, ASAT_ERR_INVALID_ARGUMENT, "null argument");
    require(nclauses > 0 && literals && lengths, ASAT_ERR_INVALID_ARGUMENT, "empty SAT problem");
    require(eps >= 1e-16 && eps <= 1.0, ASAT_ERR_INVALID_ARGUMENT, "Epsilon must be in [1e-16, 1]");
    require(batch >= 1, ASAT_ERR_INVALID_ARGUMENT, "Batch size must be positive");
    require(max_steps >= 1, ASAT_ERR_INVALID_ARGUMENT, "Max iterations must be positive");
    require(max_time > 0.0, ASAT_ERR_INVALID_ARGUMENT, "Max time must be positive");
    const int esize = precision == ASAT_F64 ? 8 : 4;
    // SatProblem::AddClauses (AnalogSAT/problem/SatProblem.cpp:64-95): zero
    // literals and empty clauses dropped, N = max |lit|, K = longest clause
    int n = 0, m = 0, k = 0;
    long long off = 0, nl = 0;
    for (int j = 0; j < nclauses; j++) {
        require(lengths[j] >= 0, ASAT_ERR_INVALID_ARGUMENT, "negative clause length");
        int len1 = 0;
        for (int i = 0; i < lengths[j]; i++) {
            const int l = literals[off + i];
            if (l == 0) continue;
            require(l != INT32_MIN, ASAT_ERR_INVALID_ARGUMENT, "literal out of range");
            n = std::max(n, std::abs(l));
            len1++;
        }
        if (len1 > 0) { m++; k = std::max(k, len1); nl += len1; }
        off += lengths[j];
    }
    require(m > 0 && k > 0, ASAT_ERR_INVALID_ARGUMENT, "empty SAT problem");
    require(k <= GEN_KMAX, ASAT_ERR_UNSUPPORTED, "clauses longer than 32 literals are not supported");
    const long long aoff = (n + 1 + 63) / 64 * 64;
    const long long alloc = aoff + (m + 63LL) / 64 * 64;
    const long long lds = small_lds_bytes(n, m, k, alloc, nl, esize);
    const Sections sec = sections(n, m, k, alloc, esize);
    require(lds <= SMALL_LDS_MAX && sec.end <= ASAT_NOALLOC_SCRATCH_BYTES, ASAT_ERR_INVALID_ARGUMENT,
            "problem too large for the allocation-free path (its integrator must fit in one workgroup's LDS "
            "and its buffers in ASAT_NOALLOC_SCRATCH_BYTES)");

    std::lock_guard<std::mutex> lock(g_mu);
    unsigned char* dev = device_scratch();
    unsigned char* st = g_stage;
    memset(st, 0, sec.end);
    // control block as after CudaRungeKutta::Configure / Reset: h = 1e-9
    Ctrl* c = (Ctrl*)(st + sec.ctrl);
    c->h = 1e-9; c->hnext = 1e-9; c->eps = eps; c->floor = TabD::TINY;
    c->step_start = 1;
    SmallSolve* ss = (SmallSolve*)(st + sec.solve);
    ss->result = -1;
    // state [-1, s_1..s_N, pad | a_1..a_M (input clause order), pad]
    if (esize == 8) {
        double* x = (double*)(st + sec.x);
        const double* p = (const double*)pub0;
        x[0] = -1.0;
        for (int i = 0; i < n; i++) x[1 + i] = p[i];
        for (int j = 0; j < m; j++) x[aoff + j] = p[n + j];
    } else {
        float* x = (float*)(st + sec.x);
        const float* p = (const float*)pub0;
        x[0] = -1.0f;
        for (int i = 0; i < n; i++) x[1 + i] = p[i];
        for (int j = 0; j < m; j++) x[aoff + j] = p[n + j];
    }
    // clause matrix (zero padded to K) and each slot's var-major position,
    // a variable's terms in the reference's accumulation order (clauses in
    // order, literals in order)
    int* lit = (int*)(st + sec.lit);
    int* vptr = (int*)(st + sec.vptr);
    int* spos = (int*)(st + sec.spos);
    off = 0;
    int row = 0;
    for (int j = 0; j < nclauses; j++) {
        int q = 0;
        for (int i = 0; i < lengths[j]; i++) {
            const int l = literals[off + i];
            if (l == 0) continue;
            lit[(size_t)row * k + q++] = l;
        }
        if (q > 0) row++;
        off += lengths[j];
    }
    long long nlist = 0;
    const int runlen = small_build_lists(lit, nullptr, n, m, k, alloc, esize, vptr, spos, &nlist);
    // small_build_lists takes uniform runs only where they fit in SMALL_LDS_MAX
    const long long lds_run = small_lds_bytes(n, m, k, alloc, nlist, esize);
    HIP_CHECK(hipMemcpy(dev, st, sec.end, hipMemcpyHostToDevice));

    SmallArgs a;
    memset(&a, 0, sizeof(a));
    a.lit = (const int*)(dev + sec.lit);
    a.vptr = (const int*)(dev + sec.vptr);
    a.spos = (const int*)(dev + sec.spos);
    a.x = dev + sec.x;
    a.ctrl = (Ctrl*)(dev + sec.ctrl);
    a.solve = (SmallSolve*)(dev + sec.solve);
    a.max_steps = max_steps;
    a.max_time = max_time;
    a.time_float = time_float;
    a.batch = batch;
    a.n = n; a.m = m; a.k = k; a.aoff = (int)aoff; a.alloc = (int)alloc; a.nlist = (int)nlist;
    a.runlen = runlen;
    a.family = ASAT_FAMILY_CTDS;
    a.knorm = std::pow(2.0, -k);
    // relaunch until the loop ends (each launch runs a bounded number of attempts)
    long long last = 0;
    for (;;) {
        if (esize == 8) launch_small<double>(a, (size_t)lds_run);
        else launch_small<float>(a, (size_t)lds_run);
        HIP_CHECK(hipMemcpy(st, dev, sec.x, hipMemcpyDeviceToHost));
        if (ss->result >= 0) break;
        require(c->attempts > last, ASAT_ERR_HIP, "integrator made no progress");
        last = c->attempts;
    }
    HIP_CHECK(hipMemcpy(st + sec.x, dev + sec.x, (size_t)alloc * esize, hipMemcpyDeviceToHost));
    if (esize == 8) {
        const double* x = (const double*)(st + sec.x);
        double* p = (double*)pub_out;
        for (int i = 0; i < n; i++) p[i] = x[1 + i];
        for (int j = 0; j < m; j++) p[n + j] = x[aoff + j];
    } else {
        const float* x = (const float*)(st + sec.x);
        float* p = (float*)pub_out;
        for (int i = 0; i < n; i++) p[i] = x[1 + i];
        for (int j = 0; j < m; j++) p[n + j] = x[aoff + j];
    }
    out->result = ss->result;
    out->steps = ss->solver_steps;
    out->last_violations = ss->last_viol;
    out->on_device = 1;
    if (stats) {
        stats->time = c->time; stats->hdid = c->hdid; stats->hnext = c->hnext; stats->errmax = c->errmax;
        stats->steps = c->steps; stats->attempts = c->attempts; stats->rhscount = c->rhscount;
    }
    return ASAT_OK;
}

}  // namespace

extern "C" int asat_solve_noalloc(int precision, const int* literals, const int* lengths, int nclauses,
                                  const void* pub0, double eps, int batch, long long max_steps, double max_time,
                                  int time_float, void* pub_out, asat_solve_result_t* out, asat_rk_stats_t* stats)
{
    try {
        return solve_noalloc(precision, literals, lengths, nclauses, pub0, eps, batch, max_steps, max_time,
                             time_float, pub_out, out, stats);
    } catch (const Error& e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return ASAT_ERR_HIP;
    }
}
