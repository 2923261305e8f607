/*
 * CPU oracle for the AnalogSat CTDS integrator -- TEST INFRASTRUCTURE ONLY.
 *
 * This library checks the results of the HIP implementation.  Only the test
 * suite, the smoke check and the CPU-baseline leg of the benchmark may load
 * it; the product library never links or calls it.
 *
 * Contents (each restates a piece of the reference's CPU path, see the
 * file:line citations next to each routine):
 *   - asat_oracle_impl.h, instantiated twice:
 *       *_f64  state double, control double  (== CpuSat<double>+CpuRungeKutta<double>)
 *       *_mix  state float,  control double  (definition of the GPU FP32 mode)
 *   - problem container semantics (SatProblem::AddClauses)
 *   - FastSatMaker's clause generator and CpuRandom's uniform stream, both
 *     driven by a restatement of libstdc++'s minstd_rand0 + generate_canonical
 *
 * Pinning (see tests/test_oracle_*.py): FP64 RHS, step sequences, violation
 * counts, generator and uniform streams are compared bit for bit with the
 * reference's own sources compiled into oracle/_ref (golden fixtures under
 * tests/golden/), and the RHS/tableau/controller against closed forms.
 * The *_mix mode has no reference counterpart: it is pinned by closed forms,
 * by agreement with *_f64 to float accuracy on single RHS evaluations, and by
 * the GPU FP32 path agreeing with it.
 *
 * Build: gcc -O2 -ffp-contract=off -fPIC -shared (oracle/Makefile).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* Step-controller pow.  Mode 0 (default): glibc pow, as the reference calls it
   (AnalogSAT/cpu_sat/CpuRungeKutta.cpp:243,254).  Mode 1: correctly rounded
   x^-0.25 / x^-0.2 (one double-double Newton step on x r^q = 1 plus the
   first-order factor for -0.2 != -1/5), the definition the GPU controller
   uses; glibc's pow differs from it in about 0.1% of arguments. */
static int g_pow_mode = 0;
void orc_set_pow_mode(int mode) { g_pow_mode = mode; }
int orc_get_pow_mode(void) { return g_pow_mode; }

/* Accumulation of the bias term's sum of a: 0 = serial sum in the state type
   (the reference, CpuSat.cpp:224 / CpuSatTanh.cpp:215), 1 = sum in double,
   rounded to the state type once (the GPU's definition; its tree order
   changes a double sum of at most 2^31 terms only far below float rounding). */
static int g_mean_mode = 0;
void orc_set_mean_mode(int mode) { g_mean_mode = mode; }
int orc_get_mean_mode(void) { return g_mean_mode; }

static void dd_mul_d(double *hi, double *lo, double b)
{
    double p = *hi * b;
    double e = fma(*hi, b, -p);
    e = fma(*lo, b, e);
    double s = p + e;
    *lo = e - (s - p);
    *hi = s;
}

double orc_pow_cr(double x, int q)
{
    const double y = q == 4 ? -0.25 : -0.2;
    if (!(x > 0.0) || !(x < 1.0e300) || !(x > 1.0e-300)) return pow(x, y);
    double r = pow(x, y);
    double hi = r, lo = 0.0;
    for (int i = 1; i < q; i++) dd_mul_d(&hi, &lo, r);
    dd_mul_d(&hi, &lo, x);
    double e = (1.0 - hi) - lo;
    double corr = r * e / (double)q;
    if (q == 5) corr -= r * (1.1102230246251565404e-17 * log(x));
    return r + corr;
}

static double ctrl_pow(double x, double y)
{
    if (g_pow_mode == 1 && (y == -0.25 || y == -0.2)) return orc_pow_cr(x, y == -0.25 ? 4 : 5);
    return pow(x, y);
}

#define CT_FABS fabs
#define CT_FMAX fmax
#define CT_FMIN fmin
#define CT_POW(x, y) ctrl_pow((double)(x), (double)(y))

#define ST double
#define CT double
#define SFX _f64
#define ST_SIN sin
#define ST_TANH tanh
#include "asat_oracle_impl.h"
#undef ST
#undef CT
#undef SFX
#undef ST_SIN
#undef ST_TANH

#define ST float
#define CT double
#define SFX _mix
#define ST_SIN sinf
#define ST_TANH tanhf
#include "asat_oracle_impl.h"
#undef ST
#undef CT
#undef SFX
#undef ST_SIN
#undef ST_TANH

#undef CT_FABS
#undef CT_FMAX
#undef CT_FMIN
#undef CT_POW

/* extended precision (x87 80-bit long double): CpuSat<long double> and
   CpuRungeKutta<long double> (AnalogSAT/cpu_sat/CpuSat.cpp:393,
   AnalogSAT/cpu_sat/CpuRungeKutta.cpp:273), for FP32/FP64 error studies */
#define CT_FABS fabsl
#define CT_FMAX fmaxl
#define CT_FMIN fminl
#define CT_POW(x, y) powl((x), (y))
#define ST long double
#define CT long double
#define SFX _ld
#define ST_SIN sinl
#define ST_TANH tanhl
#include "asat_oracle_impl.h"
#undef ST
#undef CT
#undef SFX
#undef ST_SIN
#undef ST_TANH
#undef CT_FABS
#undef CT_FMAX
#undef CT_FMIN
#undef CT_POW

/* ---------------------------------------------------------------------------
 * Problem container: SatProblem::AddClauses (AnalogSAT/problem/SatProblem.cpp:64-95)
 * Drops zero literals and empty clauses; N = max |lit|, K = max clause length.
 * Writes the padded clause-major matrix C[j*K + q] (zeros pad short clauses,
 * CpuSat::InitProblem, AnalogSAT/cpu_sat/CpuSat.cpp:352-377).
 * Call with C == NULL to get the sizes only.
 * ------------------------------------------------------------------------- */
int orc_problem_build(const int *literals, const int *lengths, int nclauses,
                      int *out_n, int *out_m, int *out_k, int *C)
{
    int n = 0, m = 0, k = 0;
    size_t off = 0;
    for (int j = 0; j < nclauses; j++) {
        int len1 = 0;
        for (int i = 0; i < lengths[j]; i++) {
            int lit = literals[off + i];
            if (lit == 0) continue;
            int a = lit < 0 ? -lit : lit;
            if (a > n) n = a;
            len1++;
        }
        if (len1 > 0) { m++; if (len1 > k) k = len1; }
        off += (size_t)lengths[j];
    }
    *out_n = n; *out_m = m; *out_k = k;
    if (!C) return 0;
    memset(C, 0, sizeof(int) * (size_t)m * (size_t)k);
    off = 0;
    int row = 0;
    for (int j = 0; j < nclauses; j++) {
        int q = 0;
        for (int i = 0; i < lengths[j]; i++) {
            int lit = literals[off + i];
            if (lit == 0) continue;
            C[(size_t)row * k + q++] = lit;
        }
        if (q > 0) row++;
        off += (size_t)lengths[j];
    }
    return 0;
}

/* ---------------------------------------------------------------------------
 * libstdc++ std::default_random_engine (= minstd_rand0: x <- 16807 x mod 2^31-1,
 * seed s -> s mod m, 0 -> 1) and uniform_real_distribution<double>(0,1), which
 * draws generate_canonical<double,53>: two engine outputs combined as
 * ((u1-1) + (u2-1) r) / r^2 with r = 2^31 - 2, clamped below 1.
 * (Published behaviour of GCC's <random>; the reference uses it in
 * AnalogSAT/problem/FastSatMaker.h:54-55 and AnalogSAT/cpu_sat/CpuRandom.h.)
 * ------------------------------------------------------------------------- */
typedef struct { uint64_t x; } orc_lcg_t;

static void lcg_seed(orc_lcg_t *g, uint64_t s)
{
    uint64_t v = s % 2147483647ULL;
    g->x = v == 0 ? 1 : v;
}

static uint64_t lcg_next(orc_lcg_t *g)
{
    g->x = (g->x * 16807ULL) % 2147483647ULL;
    return g->x;
}

static double canonical_double(orc_lcg_t *g)
{
    const double r = 2147483646.0;
    double sum = 0.0, tmp = 1.0;
    for (int i = 0; i < 2; i++) {
        sum += (double)(lcg_next(g) - 1) * tmp;
        tmp *= r;
    }
    double ret = sum / tmp;
    if (ret >= 1.0) ret = nextafter(1.0, 0.0);
    return ret;
}

/* CpuRandom<double>(seed).GenerateUniform(out, len)  (AnalogSAT/cpu_sat/CpuRandom.cpp:25-52) */
void orc_uniform_f64(unsigned int seed, double *out, int len)
{
    orc_lcg_t g;
    lcg_seed(&g, seed);
    for (int i = 0; i < len; i++) out[i] = canonical_double(&g);
}

/* CpuSat::SetRandomState from a CpuRandom<double>(seed) stream
   (AnalogSAT/cpu_sat/CpuSat.cpp:122-133): public [s = 2u-1, a = u+1] */
void orc_random_state_f64(unsigned int seed, int n, int m, double *pub)
{
    orc_uniform_f64(seed, pub, n + m);
    for (int i = 0; i < n; i++) pub[i] = pub[i] * 2.0 - 1.0;
    for (int i = n; i < n + m; i++) pub[i] = pub[i] + 1.0;
}

/* LocalSort (AnalogSAT/util/utils.cpp:44-56): ascending */
static void local_sort(int *v, int len)
{
    if (len <= 1) return;
    if (len == 2) { if (v[0] > v[1]) { int t = v[0]; v[0] = v[1]; v[1] = t; } return; }
    if (len == 3) {
        int t;
        if (v[0] > v[1]) { t = v[0]; v[0] = v[1]; v[1] = t; }
        if (v[0] > v[2]) { t = v[0]; v[0] = v[2]; v[2] = t; }
        if (v[1] > v[2]) { t = v[1]; v[1] = v[2]; v[2] = t; }
        return;
    }
    for (int i = 1; i < len; i++) {          /* insertion sort: same ascending result */
        int x = v[i], j = i - 1;
        while (j >= 0 && v[j] > x) { v[j + 1] = v[j]; j--; }
        v[j + 1] = x;
    }
}

/* FastSatMaker(seed) with Set_N/Set_M/Set_K, MakeSatProblem
   (AnalogSAT/problem/FastSatMaker.cpp:59-107): m clauses of k distinct
   variables floor(u*n)+1, sorted, redrawn on repeats, then each literal kept
   positive if u < 0.5 else negated.  out: m*k literals, clause-major. */
void orc_fastsat(int n, int m, int k, unsigned long seed, int *out)
{
    orc_lcg_t g;
    lcg_seed(&g, (uint64_t)seed);
    int clause[64];
    for (int j = 0; j < m; j++) {
        int rep;
        do {
            for (int i = 0; i < k; i++) clause[i] = (int)floor(canonical_double(&g) * n) + 1;
            local_sort(clause, k);
            rep = 0;
            for (int i = 0; i < k - 1; i++) if (clause[i] == clause[i + 1]) rep = 1;
        } while (rep);
        for (int i = 0; i < k; i++) clause[i] *= canonical_double(&g) < 0.5 ? 1 : -1;
        memcpy(out + (size_t)j * k, clause, sizeof(int) * (size_t)k);
    }
}

/* RandomSatMaker(seed) with Set_N/Set_M/Set_K, MakeSatProblem
   (AnalogSAT/problem/RandomSatMaker.cpp:59-125): each clause is drawn as in
   FastSatMaker (same stream, same rule), and a clause equal to one already
   accepted (same signed literals in the same sorted order) is drawn again.
   The reference keeps the accepted clauses in an unordered_set keyed by a
   hash of the literals; only equality decides, so any hash gives the same
   instance.  Returns 0, or -1 if m exceeds the number of distinct clauses
   C(n,k) 2^k (where the reference would draw forever). */
static uint64_t clause_hash(const int *c, int k)
{
    uint64_t h = 1469598103934665603ULL;
    for (int i = 0; i < k; i++) { h ^= (uint32_t)c[i]; h *= 1099511628211ULL; }
    return h;
}

int orc_randomsat(int n, int m, int k, unsigned long seed, int *out)
{
    double distinct = 1.0;                   /* C(n,k) 2^k, as a double */
    for (int i = 0; i < k; i++) distinct = distinct * (double)(n - i) / (double)(i + 1) * 2.0;
    if ((double)m > distinct) return -1;
    size_t cap = 16;
    while (cap < 2 * (size_t)m + 16) cap <<= 1;
    int *slot = (int *)malloc(sizeof(int) * cap);   /* index of an accepted clause, -1 empty */
    for (size_t i = 0; i < cap; i++) slot[i] = -1;
    orc_lcg_t g;
    lcg_seed(&g, (uint64_t)seed);
    for (int j = 0; j < m; j++) {
        int *clause = out + (size_t)j * k;
        for (;;) {
            int rep;
            do {
                for (int i = 0; i < k; i++) clause[i] = (int)floor(canonical_double(&g) * n) + 1;
                local_sort(clause, k);
                rep = 0;
                for (int i = 0; i < k - 1; i++) if (clause[i] == clause[i + 1]) rep = 1;
            } while (rep);
            for (int i = 0; i < k; i++) clause[i] *= canonical_double(&g) < 0.5 ? 1 : -1;
            size_t at = (size_t)clause_hash(clause, k) & (cap - 1);
            int seen = 0;
            while (slot[at] >= 0) {
                if (memcmp(out + (size_t)slot[at] * k, clause, sizeof(int) * (size_t)k) == 0) { seen = 1; break; }
                at = (at + 1) & (cap - 1);
            }
            if (!seen) { slot[at] = j; break; }
        }
    }
    free(slot);
    return 0;
}

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
   easy as 1, 2, 3", SC'11): ten rounds of two 32x32->64 multiplies with the
   published multipliers and Weyl key increments.  Not a reference function:
   the reference draws its device initial conditions from cuRAND MTGP32
   (AnalogSAT/cuda_base/CudaRandom.cu:24-60), a closed library replaced by
   this counter-based generator. */
void orc_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4])
{
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; r++) {
        const uint64_t p0 = (uint64_t)M0 * c0, p1 = (uint64_t)M1 * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += W0;
        k1 += W1;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* The uniform stream of the library's ISatRandom: element i of a draw at
   stream offset `offset` is Philox of counter {idx lo, idx hi, 0x5EED, 0}
   (idx = offset + i) under key {seed lo, seed hi}; a double takes the top 53
   bits of (word0 << 32 | word1), a float the top 24 bits of word0. */
void orc_philox_uniform(unsigned long long seed, unsigned long long offset, int len, int f32, void *out)
{
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (int i = 0; i < len; i++) {
        const unsigned long long idx = offset + (unsigned long long)i;
        const uint32_t ctr[4] = {(uint32_t)idx, (uint32_t)(idx >> 32), 0x5EEDu, 0u};
        uint32_t w[4];
        orc_philox4x32_10(ctr, key, w);
        if (f32) ((float *)out)[i] = (float)(w[0] >> 8) * 0x1.0p-24f;
        else ((double *)out)[i] = (double)((((uint64_t)w[0] << 32) | w[1]) >> 11) * 0x1.0p-53;
    }
}
