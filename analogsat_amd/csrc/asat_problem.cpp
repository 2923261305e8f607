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
#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>

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
