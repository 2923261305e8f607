// Host-side problem generation for the MI355X library.
//
// asat_problem_fastsat reproduces FastSatMaker (AnalogSAT/problem/FastSatMaker.cpp:59-107)
// including its random stream, so a seed names the same instance on both
// sides.  The stream is GCC's std::default_random_engine (minstd_rand0:
// x <- 16807 x mod (2^31 - 1), seed s -> s mod m, 0 -> 1) feeding
// uniform_real_distribution<double>(0,1), i.e. generate_canonical<double,53>,
// which combines two engine outputs as ((u1-1) + (u2-1) r) / r^2, r = 2^31-2.
// The clause rule: k variables floor(u*n)+1, sorted ascending, redrawn while
// any repeats, then each literal kept positive when u < 0.5, else negated.
//
// asat_problem_randomsat reproduces RandomSatMaker
// (AnalogSAT/problem/RandomSatMaker.cpp:59-125): the same draw, and a clause
// equal to one already accepted (same signed literals in sorted order) is
// drawn again.  The reference looks clauses up in a hash set; only equality
// decides, so the set's hash function does not change the instance.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <unordered_set>

#include "analogsat_hip.h"

namespace asat {
extern thread_local std::string g_last_error;

namespace {

class MinStd {
public:
    explicit MinStd(unsigned long seed)
    {
        uint64_t v = (uint64_t)seed % kMod;
        x_ = v == 0 ? 1 : v;
    }
    static MinStd from_state(uint64_t x)
    {
        MinStd g(1);
        g.x_ = x;
        return g;
    }
    uint64_t state() const { return x_; }
    uint64_t next()
    {
        x_ = (x_ * 16807ULL) % kMod;
        return x_;
    }
    // generate_canonical<double, 53>: two 31-bit draws
    double canonical()
    {
        const double r = 2147483646.0;
        double sum = (double)(next() - 1);
        sum += (double)(next() - 1) * r;
        double v = sum / (r * r);
        return v >= 1.0 ? std::nextafter(1.0, 0.0) : v;
    }

private:
    static constexpr uint64_t kMod = 2147483647ULL;
    uint64_t x_;
};

void sort_small(int* v, int k)
{
    for (int i = 1; i < k; i++) {
        int x = v[i], j = i - 1;
        while (j >= 0 && v[j] > x) { v[j + 1] = v[j]; j--; }
        v[j + 1] = x;
    }
}

}  // namespace
}  // namespace asat

static int fastsat_from(asat::MinStd& g, int n, int m, int k, int* literals_out)
{
    if (n <= 0 || m < 0 || k <= 0 || k > 64 || k > n || !literals_out) {
        asat::g_last_error = "fastsat: need n > 0, m >= 0, 0 < k <= min(n, 64) and an output array";
        return ASAT_ERR_INVALID_ARGUMENT;
    }
    int clause[64];
    for (int j = 0; j < m; j++) {
        bool repeated;
        do {
            for (int i = 0; i < k; i++) clause[i] = (int)std::floor(g.canonical() * n) + 1;
            asat::sort_small(clause, k);
            repeated = false;
            for (int i = 0; i + 1 < k; i++) repeated |= clause[i] == clause[i + 1];
        } while (repeated);
        for (int i = 0; i < k; i++) clause[i] *= g.canonical() < 0.5 ? 1 : -1;
        std::memcpy(literals_out + (size_t)j * k, clause, sizeof(int) * (size_t)k);
    }
    return ASAT_OK;
}

// Accepted clauses of one RandomSatMaker call, keyed by clause index into the
// output array (the literals stay where they were written).
struct ClauseKey {
    const int* base;
    int k;
    size_t operator()(int j) const
    {
        uint64_t h = 1469598103934665603ULL;
        for (int i = 0; i < k; i++) { h ^= (uint32_t)base[(size_t)j * k + i]; h *= 1099511628211ULL; }
        return (size_t)h;
    }
    bool operator()(int a, int b) const
    {
        return std::memcmp(base + (size_t)a * k, base + (size_t)b * k, sizeof(int) * (size_t)k) == 0;
    }
};

static int randomsat_from(asat::MinStd& g, int n, int m, int k, int* literals_out)
{
    if (n <= 0 || m < 0 || k <= 0 || k > 64 || k > n || !literals_out) {
        asat::g_last_error = "randomsat: need n > 0, m >= 0, 0 < k <= min(n, 64) and an output array";
        return ASAT_ERR_INVALID_ARGUMENT;
    }
    // C(n,k) 2^k distinct clauses exist; asking for more would draw forever
    double distinct = 1.0;
    for (int i = 0; i < k; i++) distinct = distinct * (double)(n - i) / (double)(i + 1) * 2.0;
    if ((double)m > distinct) {
        asat::g_last_error = "randomsat: more clauses requested than distinct clauses exist";
        return ASAT_ERR_INVALID_ARGUMENT;
    }
    ClauseKey key{literals_out, k};
    std::unordered_set<int, ClauseKey, ClauseKey> seen((size_t)m + 16, key, key);
    for (int j = 0; j < m; j++) {
        int* clause = literals_out + (size_t)j * k;
        for (;;) {
            bool repeated;
            do {
                for (int i = 0; i < k; i++) clause[i] = (int)std::floor(g.canonical() * n) + 1;
                asat::sort_small(clause, k);
                repeated = false;
                for (int i = 0; i + 1 < k; i++) repeated |= clause[i] == clause[i + 1];
            } while (repeated);
            for (int i = 0; i < k; i++) clause[i] *= g.canonical() < 0.5 ? 1 : -1;
            if (seen.insert(j).second) break;
        }
    }
    return ASAT_OK;
}

extern "C" int asat_problem_randomsat(int n, int m, int k, unsigned long seed, int* literals_out)
{
    asat::MinStd g(seed);
    return randomsat_from(g, n, m, k, literals_out);
}

extern "C" int asat_problem_randomsat_stream(int n, int m, int k, unsigned long long* engine, int* literals_out)
{
    if (!engine || *engine == 0 || *engine >= 2147483647ULL) {
        asat::g_last_error = "engine state must be in [1, 2^31 - 2]";
        return ASAT_ERR_INVALID_ARGUMENT;
    }
    asat::MinStd g = asat::MinStd::from_state(*engine);
    int rc = randomsat_from(g, n, m, k, literals_out);
    if (rc == ASAT_OK) *engine = g.state();
    return rc;
}

extern "C" int asat_problem_fastsat(int n, int m, int k, unsigned long seed, int* literals_out)
{
    asat::MinStd g(seed);
    return fastsat_from(g, n, m, k, literals_out);
}

extern "C" int asat_problem_seed_engine(unsigned long seed, unsigned long long* engine)
{
    if (!engine) {
        asat::g_last_error = "null engine";
        return ASAT_ERR_INVALID_ARGUMENT;
    }
    *engine = asat::MinStd(seed).state();
    return ASAT_OK;
}

extern "C" int asat_problem_fastsat_stream(int n, int m, int k, unsigned long long* engine, int* literals_out)
{
    if (!engine || *engine == 0 || *engine >= 2147483647ULL) {
        asat::g_last_error = "engine state must be in [1, 2^31 - 2]";
        return ASAT_ERR_INVALID_ARGUMENT;
    }
    asat::MinStd g = asat::MinStd::from_state(*engine);
    int rc = fastsat_from(g, n, m, k, literals_out);
    if (rc == ASAT_OK) *engine = g.state();
    return rc;
}
