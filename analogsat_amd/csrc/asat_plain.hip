// Plain single-trajectory CTDS solver for MI355X (gfx950): one problem, one
// trajectory, the whole chip.  Replaces CudaSat1/2/3 + CudaRungeKutta
// (AnalogSAT/cuda_sat/CudaSat3.cu:33-430, AnalogSAT/cuda_sat/CudaRungeKutta.cu:56-484).
//
// Data layout in HBM (T = double for ASAT_F64, float for ASAT_F32):
//   lit[m*K + q]      int32 signed literal (+v / -v, 0 = padding), clause-major,
//                     padded to whole tiles of TC clauses
//   contrib[p]        per-slot contribution 2 c a K K_mi of the current stage.
//                     Variables are split into groups (at most about 2048 float
//                     / 1024 double variables, halved until the largest region
//                     fits the LDS); a group's contributions form one
//                     region (64-element aligned, at most 144 KB), in slot
//                     (clause) order.  Between a var pass and the next clause
//                     pass the same array holds the slot spins: each slot's
//                     stage spin at the slot's position
//   spos, srank       per tile: the positions of the tile's slots in ascending
//                     order (int32), and each slot's rank in that order (uint16,
//                     the literal's sign in bit 15; literal q of the tile's
//                     clauses at q*TC + clause, so each load is one run)
//   runs              big FP32 tiles (up to 4608 clauses): a word per variable
//                     group in place of spos (the tile's positions in one
//                     region are consecutive: first rank, first offset)
//   pw16, pv0         big FP32 tiles: per clause its primary variable relative
//                     to the tile's first (pv0), the primary literal's position
//                     and sign, in place of pvar and the primary slots' ranks
//   csr_ptr[0..N+1]   var-major CSR row pointers
//   csr_loc[e]        uint16 offset of each of a variable's contributions in
//                     its group's region, in slot order
//   state buffer      [ -1, s_1..s_N, pad | a_1..a_M, pad ], aux at aoff
//   k1..k5            same layout as the state (stage derivatives)
//   tmp               spin part of the next stage's input (aux stage inputs are
//                     recomputed inside the clause pass and never stored)
// Problems outside the fast paths use a generic clause pass (K other than 3
// and 4: pos[m*K + q] per slot) and a gather var pass (a group whose region
// exceeds the LDS: csr_pos global positions, ~400 KB groups, or one
// workgroup per variable when rows are long); both gather the stage spins by
// variable from x / tmp.
//
// One RK45 attempt = 6 x (clause pass + var pass) + 1 control kernel:
//   clause pass J: per tile, read the slot spins through the tile's sorted
//                  positions (one run per region), K_m, aux stage input and
//                  k_J(a_m) = a K^2, then store the tile's contributions at
//                  the same positions
//   var pass J:    one workgroup per group reads the group's region into LDS
//                  with coalesced loads, adds each variable's contributions
//                  in slot order (== the reference's accumulation order),
//                  then k_J(s_i), the spin part of stage J+1's input, and
//                  writes the region back as stage J+1's slot spins
//   J = 6 passes also form x_new and |xerr/xscal| and a per-block FP64 max;
//   the control kernel finishes the max-norm and runs the Cash-Karp
//   accept/reject/step-size logic in FP64 on the device.
// Step(count) fills the slot spins of x, enqueues attempts without host
// round trips and reads the control block back once per round of attempts.
#include "asat_common.h"
#include "asat_small.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <vector>

namespace asat {

thread_local std::string g_last_error;

constexpr int BLK = 256;          // stage-kernel block size (4 waves)
constexpr int CTRL_BLK = 1024;    // control kernel block size

static inline long long round_up(long long a, long long b) { return (a + b - 1) / b * b; }

// ===========================================================================
// problem (host)
// ===========================================================================
constexpr int TC = 2048;                  // clauses per tile of the tiled clause pass (one workgroup)
constexpr int TC_BIG = 4608;              // FP32 3-literal systems with direct primary slots: the most
                                          // clauses a tile may have (its size is chosen per problem)
constexpr int TC_MAX = TC_BIG;
constexpr unsigned PW_VMASK = 0x1FFFu;    // big tiles' primary words: relative variable bits,
constexpr int PW_QSHIFT = 13;             // then the primary literal's position, then its sign
constexpr int TILE_BT = 1024;             // threads per tile workgroup (a clause each)
constexpr int VBLK = 1024;                // var-pass workgroup (one variable group)
constexpr int VAR_LDS_BYTES = 147456;     // LDS holding one group's contribution region
constexpr int VAR_LDS_CU = 163840 - 1024; // LDS two var workgroups may share on one CU
constexpr int REGION_ALIGN = 64;          // region starts, in elements
constexpr bool REORDER_CLAUSES = true;    // tiled pass: process clauses sorted by a primary variable

struct HostProblem {
    int n = 0, m = 0, k = 0;
    std::vector<int> lit;          // m*k, zero padded, clause-major, in processing order
    std::vector<int> perm;         // processing position -> clause index (empty: identity)
    // contribution layout: variables split into groups of vg; the
    // contributions of a group's variables form one region, in slot (clause)
    // order; region starts are REGION_ALIGN-element aligned
    int vg = 0, ngroups = 0;
    long long max_region = 0;      // largest region incl. alignment padding
    std::vector<long long> gstart; // ngroups+1
    std::vector<int> pos;          // m*k: the slot's contribution position, -1 for padding
    std::vector<int> csr_ptr;      // n+2
    std::vector<int> csr_pos;      // contribution positions of each variable, in slot order
    std::vector<uint16_t> csr_loc; // the same relative to the variable's region start
    // tiled clause pass: per tile of TC clauses, its slots sorted by position
    std::vector<int> spos;         // ntiles*TC*k positions (-1 past the tile's last slot)
    // big tiles: per tile, a run table in place of spos.
    // Word g of tile t: bits 0-15 the first rank of group g's run in the
    // tile's sorted list, bits 16-31 the run's first position relative to
    // the region start gstart[g]; word ngroups holds the tile's slot count
    // (run_stride words per tile)
    std::vector<uint32_t> runs;
    int run_stride = 0;
    // big tiles: primary words in place of pvar and of the primary slots'
    // ranks.  pw16[j] holds the clause's primary variable relative to its
    // tile's first (pv0[tile]) in bits 0-12, the primary literal's position
    // in bits 13-14 (3: padding clause) and its sign in bit 15; srank then
    // holds K - 1 ranks per clause, the other literals in clause order
    std::vector<uint16_t> pw16;
    std::vector<int> pv0;
    std::vector<uint16_t> srank;   // per slot: its rank in the tile's sorted list, the literal's sign
                                   // in bit 15, 0xFFFF for padding; per tile, literal q of every
                                   // clause first (index tile*TC*K + q*TC + clause in the tile)
    // primary variables (tiled pass): pvar[j] is the primary variable of
    // processing-order clause j and pq[j] the first of its literals on it
    std::vector<int> pvar;
    std::vector<unsigned char> pq;
    // direct primary slots (tiled pass with LDS regions): a clause's primary
    // slot is not in any region; its spin is read from the stage vector by
    // variable and its term stored at the clause's processing position of
    // pcontrib.  Regions hold the other slots.  pc0[g] is the first clause
    // whose primary variable is in group g (ngroups+1 entries).
    bool split = false;
    std::vector<int> pc0;
    int tc = TC;                   // tiled pass: clauses per tile
};

// SatProblem::AddClauses semantics (AnalogSAT/problem/SatProblem.cpp:64-95) +
// CpuSat::InitProblem clause matrix (AnalogSAT/cpu_sat/CpuSat.cpp:352-377).
static void parse_clauses(const int* literals, const int* lengths, int nclauses, HostProblem& P)
{
    require(nclauses >= 0, ASAT_ERR_INVALID_ARGUMENT, "negative clause count");
    require(nclauses == 0 || (literals && lengths), ASAT_ERR_INVALID_ARGUMENT, "null clause arrays");
    int n = 0, m = 0, k = 0;
    size_t off = 0;
    for (int j = 0; j < nclauses; j++) {
        require(lengths[j] >= 0, ASAT_ERR_INVALID_ARGUMENT, "negative clause length");
        int len1 = 0;
        for (int i = 0; i < lengths[j]; i++) {
            int l = literals[off + i];
            if (l == 0) continue;
            require(l != INT32_MIN, ASAT_ERR_INVALID_ARGUMENT, "literal out of range");
            n = std::max(n, std::abs(l));
            len1++;
        }
        if (len1 > 0) { m++; k = std::max(k, len1); }
        off += (size_t)lengths[j];
    }
    require(m > 0 && k > 0, ASAT_ERR_INVALID_ARGUMENT, "empty SAT problem");
    require(k <= GEN_KMAX, ASAT_ERR_UNSUPPORTED, "clauses longer than 32 literals are not supported");
    require((long long)m * k < (1LL << 31) - (long long)TC_MAX * k, ASAT_ERR_UNSUPPORTED, "more than 2^31 literal slots");
    P.n = n; P.m = m; P.k = k;
    P.lit.assign((size_t)m * k, 0);
    off = 0;
    int row = 0;
    for (int j = 0; j < nclauses; j++) {
        int q = 0;
        for (int i = 0; i < lengths[j]; i++) {
            int l = literals[off + i];
            if (l == 0) continue;
            P.lit[(size_t)row * k + q++] = l;
        }
        if (q > 0) row++;
        off += (size_t)lengths[j];
    }
}

// Processing order of the tiled clause pass.  Every clause gets one primary
// variable (greedily its least-loaded variable, so that each variable is
// primary for about M/N clauses) and clauses are processed sorted by it
// (stable).  A wave then reads its clauses' primary spins from a few adjacent
// lines, and a tile stores its primary contributions as one run.  Literal
// order inside a clause is kept, and so is every per-variable sum order.
static void reorder_clauses(HostProblem& P)
{
    const int K = P.k, m = P.m, n = P.n;
    std::vector<int> load((size_t)n + 1, 0), prim((size_t)m, 0);
    for (int j = 0; j < m; j++) {
        int best = 0;
        for (int q = 0; q < K; q++) {
            const int l = P.lit[(size_t)j * K + q];
            if (!l) continue;
            const int v = std::abs(l);
            if (best == 0 || load[v] < load[best]) best = v;
        }
        prim[j] = best;
        load[best]++;
    }
    P.pvar.assign((size_t)m, 0);
    P.pq.assign((size_t)m, 0);
    std::vector<int> start((size_t)n + 2, 0);
    for (int j = 0; j < m; j++) start[(size_t)prim[j] + 1]++;
    for (int v = 0; v <= n; v++) start[v + 1] += start[v];
    P.perm.assign((size_t)m, 0);
    for (int j = 0; j < m; j++) P.perm[(size_t)start[prim[j]]++] = j;
    std::vector<int> lit2(P.lit.size());
    for (int j = 0; j < m; j++)
        std::copy_n(P.lit.begin() + (size_t)P.perm[j] * K, K, lit2.begin() + (size_t)j * K);
    P.lit.swap(lit2);
    for (int j = 0; j < m; j++) {
        const int v = prim[(size_t)P.perm[j]];
        P.pvar[(size_t)j] = v;
        int q = 0;
        while (std::abs(P.lit[(size_t)j * K + q]) != v) q++;
        P.pq[(size_t)j] = (unsigned char)q;
    }
}

// Contribution layout for groups of vg variables, and the transposed
// var-major slot list (role of CudaSat3::InitProblem,
// AnalogSAT/cuda_sat/CudaSat3.cu:266-307, without the pad-to-32).
// pv > 0 (tiled pass with primary variables): direct primary slots.  The
// regions then hold only the other slots, and a variable's list entry for a
// primary slot points past its group's region: the LDS var pass loads the
// group's primary terms (pcontrib from pc0[g] rounded down to a multiple of pv
// elements, its 16-byte load unit) right behind the region.
static void build_layout(HostProblem& P, int vg, int pv)
{
    const int n = P.n, K = P.k;
    const bool split = pv > 0 && !P.pvar.empty();
    P.split = split;
    P.vg = vg;
    P.ngroups = (n + vg - 1) / vg;
    auto is_prim = [&](size_t s) { return split && P.pq[s / K] == (unsigned char)(s % K); };
    std::vector<long long> gcount((size_t)P.ngroups, 0);
    for (size_t s = 0; s < P.lit.size(); s++) {
        int l = P.lit[s];
        if (l && !is_prim(s)) gcount[(size_t)(std::abs(l) - 1) / vg]++;
    }
    P.pc0.clear();
    if (split) {
        // clauses are sorted by primary variable: group g's primaries are the
        // processing range [pc0[g], pc0[g+1])
        P.pc0.assign((size_t)P.ngroups + 1, P.m);
        for (int j = P.m - 1; j >= 0; j--) P.pc0[(size_t)(P.pvar[(size_t)j] - 1) / vg] = j;
        for (int g = P.ngroups - 1; g >= 0; g--) P.pc0[g] = std::min(P.pc0[g], P.pc0[(size_t)g + 1]);
    }
    P.gstart.assign((size_t)P.ngroups + 1, 0);
    P.max_region = 0;
    for (int g = 0; g < P.ngroups; g++) {
        const long long sz = round_up(std::max<long long>(gcount[g], 1), REGION_ALIGN);
        P.gstart[g + 1] = P.gstart[g] + sz;
        long long lds = sz;
        if (split) lds += round_up(P.pc0[(size_t)g + 1] - (P.pc0[g] / pv) * pv, pv);
        P.max_region = std::max(P.max_region, lds);
    }
    require(P.gstart[P.ngroups] < (1LL << 31), ASAT_ERR_UNSUPPORTED, "contribution array exceeds 2^31 entries");
    P.pos.assign(P.lit.size(), -1);
    std::vector<long long> gfill(P.gstart.begin(), P.gstart.end() - 1);
    for (size_t s = 0; s < P.lit.size(); s++) {
        int l = P.lit[s];
        if (l && !is_prim(s)) P.pos[s] = (int)gfill[(size_t)(std::abs(l) - 1) / vg]++;
    }
    // per-variable lists of contribution positions; slots visited in the
    // original clause order, which is the reference's accumulation order
    P.csr_ptr.assign((size_t)n + 2, 0);
    for (size_t s = 0; s < P.lit.size(); s++) {
        int l = P.lit[s];
        if (l) P.csr_ptr[(size_t)std::abs(l) + 1]++;
    }
    for (int i = 0; i <= n; i++) P.csr_ptr[i + 1] += P.csr_ptr[i];
    P.csr_pos.assign((size_t)P.csr_ptr[n + 1], 0);
    std::vector<long long> cloc(split ? P.csr_pos.size() : 0);
    std::vector<int> fill(P.csr_ptr.begin(), P.csr_ptr.end() - 1);
    std::vector<int> row((size_t)P.m);
    for (int j = 0; j < P.m; j++) row[P.perm.empty() ? j : P.perm[j]] = j;
    for (int mo = 0; mo < P.m; mo++)
        for (int q = 0; q < K; q++) {
            const size_t s = (size_t)row[mo] * K + q;
            int l = P.lit[s];
            if (!l) continue;
            const int v = std::abs(l);
            const int e = fill[v]++;
            P.csr_pos[(size_t)e] = P.pos[s];
            if (split) {
                const size_t g = (size_t)(v - 1) / vg;
                cloc[(size_t)e] = is_prim(s) ? (P.gstart[g + 1] - P.gstart[g]) + row[mo] - (P.pc0[g] / pv) * pv
                                             : P.pos[s] - P.gstart[g];
            }
        }
    P.csr_loc.clear();
    if (P.max_region <= 65536) {
        P.csr_loc.resize(P.csr_pos.size());
        for (int i = 1; i <= n; i++) {
            const long long base = P.gstart[(size_t)(i - 1) / vg];
            for (int e = P.csr_ptr[i]; e < P.csr_ptr[i + 1]; e++)
                P.csr_loc[e] = (uint16_t)(split ? cloc[(size_t)e] : P.csr_pos[e] - base);
        }
    }
}

// Big tiles: a tile's slots of one group may hold the run's positions in any
// order (ranks follow positions, and the var pass adds each variable's terms
// in the order of its offset list, whatever the positions).  The var pass
// reads the k-th list entry of each of a wave's variables with one LDS
// instruction, and writes the slot spins back the same way; entries of such
// a read that share an LDS bank serialise.  Greedy, run by run: each slot
// takes the free position whose bank its (wave, k) bucket has used least,
// counting the group's primary entries (fixed, behind the region) first.
// Banks are counted modulo 32 (a wave's access served as two 32-lane
// halves); modulo 64 measured the same within noise.  N = 1e6: the six var
// passes 172.6 -> 169.8 us per attempt, 2331 -> 2342 attempts/s.
constexpr int LDS_BANKS = 32;
static void spread_run_banks(HostProblem& P, int pv)
{
    const int K = P.k, vg = P.vg, n = P.n, m = P.m;
    const size_t per = (size_t)P.tc * K;
    const size_t L = P.lit.size();
    std::vector<int> row((size_t)m);
    for (int j = 0; j < m; j++) row[P.perm.empty() ? j : P.perm[j]] = j;
    // each slot's index in its variable's list (the reference's order)
    std::vector<int> kidx(L, 0);
    {
        std::vector<int> fill((size_t)n + 1, 0);
        for (int mo = 0; mo < m; mo++)
            for (int q = 0; q < K; q++) {
                const size_t sl = (size_t)row[mo] * K + q;
                const int l = P.lit[sl];
                if (l) kidx[sl] = fill[(size_t)std::abs(l)]++;
            }
    }
    int kmax = 1;
    for (int v = 1; v <= n; v++) kmax = std::max(kmax, P.csr_ptr[(size_t)v + 1] - P.csr_ptr[v]);
    // a variable's wave in its group's var workgroup (VBLK threads, a
    // variable each, the group's later variables in further passes)
    const int waves_pass = VBLK / 64;
    const int nw = ceil_div(vg, VBLK) * waves_pass;
    auto wave_of = [&](int v) {
        const int lv = (v - 1) % vg;
        return (lv / VBLK) * waves_pass + (lv % VBLK) / 64;
    };
    // the groups' non-primary slots, in slot order
    std::vector<long long> gofs((size_t)P.ngroups + 1, 0);
    for (size_t sl = 0; sl < L; sl++)
        if (P.lit[sl] && P.pos[sl] >= 0) gofs[(size_t)(std::abs(P.lit[sl]) - 1) / vg + 1]++;
    for (int g = 0; g < P.ngroups; g++) gofs[(size_t)g + 1] += gofs[g];
    std::vector<int> gsl((size_t)gofs[P.ngroups]);
    {
        std::vector<long long> f(gofs.begin(), gofs.end() - 1);
        for (size_t sl = 0; sl < L; sl++)
            if (P.lit[sl] && P.pos[sl] >= 0) gsl[(size_t)f[(size_t)(std::abs(P.lit[sl]) - 1) / vg]++] = (int)sl;
    }
    std::vector<uint16_t> used((size_t)nw * kmax * LDS_BANKS);
    std::vector<int> cand;
    for (int g = 0; g < P.ngroups; g++) {
        std::fill(used.begin(), used.end(), (uint16_t)0);
        const long long rlen = P.gstart[(size_t)g + 1] - P.gstart[g];
        const int pb = P.pc0[g] / pv * pv;
        for (int j = P.pc0[g]; j < P.pc0[(size_t)g + 1]; j++) {
            const size_t sl = (size_t)j * K + P.pq[(size_t)j];
            const long long off = rlen + j - pb;
            used[((size_t)wave_of(P.pvar[(size_t)j]) * kmax + kidx[sl]) * LDS_BANKS + (size_t)(off % LDS_BANKS)]++;
        }
        for (long long a = gofs[g]; a < gofs[(size_t)g + 1];) {
            // one run: the group's slots of one tile, positions p0 .. p0+c-1
            const size_t t = (size_t)gsl[(size_t)a] / per;
            long long b = a;
            int p0 = INT_MAX;
            while (b < gofs[(size_t)g + 1] && (size_t)gsl[(size_t)b] / per == t) p0 = std::min(p0, P.pos[(size_t)gsl[(size_t)b++]]);
            const int c = (int)(b - a);
            cand.assign((size_t)c, 1);
            for (long long e = a; e < b; e++) {
                const size_t sl = (size_t)gsl[(size_t)e];
                const int v = std::abs(P.lit[sl]);
                uint16_t* u = &used[((size_t)wave_of(v) * kmax + kidx[sl]) * LDS_BANKS];
                int best = -1;
                for (int i = 0; i < c; i++) {
                    if (!cand[(size_t)i]) continue;
                    const long long off = p0 + i - P.gstart[g];
                    if (best < 0 || u[off % LDS_BANKS] < u[(p0 + best - P.gstart[g]) % LDS_BANKS]) best = i;
                }
                cand[(size_t)best] = 0;
                P.pos[sl] = p0 + best;
                u[(p0 + best - P.gstart[g]) % LDS_BANKS]++;
            }
            a = b;
        }
    }
    // the offset lists follow the new positions
    for (size_t sl = 0; sl < L; sl++) {
        const int l = P.lit[sl];
        if (!l || P.pos[sl] < 0) continue;
        const int v = std::abs(l);
        const size_t e = (size_t)P.csr_ptr[v] + kidx[sl];
        P.csr_pos[e] = P.pos[sl];
        if (!P.csr_loc.empty()) P.csr_loc[e] = (uint16_t)(P.pos[sl] - P.gstart[(size_t)(v - 1) / vg]);
    }
}

// Per tile of TC clauses: the tile's slots in ascending position order.
// Positions inside a region follow slot order, so this is a stable counting
// sort of the tile's slots by variable group; the clause pass then writes
// each group's run of contributions with neighbouring lanes.  With direct
// primary slots only the other slots are sorted (K - 1 per clause), and a
// primary slot's rank entry is 0x4000 with the sign bit.
static void build_tile_order(HostProblem& P)
{
    const int K = P.k, TCt = P.tc;
    const int NR = P.split ? K - 1 : K;
    const long long ntiles = ceil_div((long long)P.m, TCt);
    const size_t per = (size_t)TCt * K, nper = (size_t)TCt * NR;
    // run tables (big tiles): a tile's positions in one region are
    // consecutive (tiles are ranges of slots, region positions follow slot
    // order), so the sorted list is one run per group
    const bool rt = P.split && TCt != TC;
    require(!rt || (long long)(P.ngroups + 1) * 4 <= (long long)nper, ASAT_ERR_UNSUPPORTED,
            "run tables larger than the positions they replace");
    P.runs.clear();
    P.run_stride = 0;
    if (rt) {
        P.run_stride = (int)round_up((long long)P.ngroups + 1, 16);
        P.runs.assign((size_t)ntiles * P.run_stride, 0);
    } else {
        P.spos.assign((size_t)ntiles * nper, -1);
    }
    // big tiles store K - 1 ranks per clause (the primary one is in pw16)
    const bool cp = P.split && TCt != TC;
    const int KR = cp ? NR : K;
    const size_t rper = (size_t)TCt * KR;
    P.srank.assign((size_t)ntiles * rper, 0xFFFF);
    P.pw16.clear();
    P.pv0.clear();
    if (cp) {
        P.pw16.assign((size_t)ntiles * TCt, (uint16_t)(3u << PW_QSHIFT));
        P.pv0.assign((size_t)ntiles, 0);
        for (long long t = 0; t < ntiles; t++) {
            const long long j0 = t * TCt, j1 = std::min<long long>(P.m, j0 + TCt);
            const int v0 = P.pvar[(size_t)j0];
            P.pv0[(size_t)t] = v0;
            for (long long j = j0; j < j1; j++) {
                const int rel = P.pvar[(size_t)j] - v0;
                require(rel >= 0 && rel <= (int)PW_VMASK, ASAT_ERR_UNSUPPORTED, "tile's primary variables too far apart");
                const int q = P.pq[(size_t)j];
                const unsigned sg = P.lit[(size_t)j * K + q] < 0 ? 0x8000u : 0u;
                P.pw16[(size_t)j] = (uint16_t)(rel | (q << PW_QSHIFT) | sg);
            }
        }
    }
    // a tile's slots of one group hold consecutive positions (in any order):
    // a slot's rank is its run's first rank plus its offset from the run's
    // lowest position
    std::vector<int> cnt((size_t)P.ngroups + 1, 0);
    std::vector<int> rmin((size_t)P.ngroups + 1, INT_MAX);
    std::vector<int> touched;
    touched.reserve(per);
    for (long long t = 0; t < ntiles; t++) {
        const size_t s0 = (size_t)t * per, s1 = std::min(P.lit.size(), s0 + per);
        const size_t r0 = (size_t)t * nper;
        touched.clear();
        for (size_t s = s0; s < s1; s++) {
            int l = P.lit[s];
            if (!l || P.pos[s] < 0) continue;
            int g = (std::abs(l) - 1) / P.vg;
            if (cnt[g]++ == 0) touched.push_back(g);
            rmin[g] = std::min(rmin[g], P.pos[s]);
        }
        std::sort(touched.begin(), touched.end());
        int acc = 0;
        if (rt) {
            // every group's word: its first rank (untouched groups: that of
            // the next run); region offsets are filled from the slots below
            uint32_t* w = P.runs.data() + (size_t)t * P.run_stride;
            size_t i = 0;
            for (int g = 0; g <= P.ngroups; g++) {
                while (i < touched.size() && touched[i] < g) acc += cnt[touched[i++]];
                w[g] = (uint32_t)acc;
            }
            require(acc <= (int)nper, ASAT_ERR_UNSUPPORTED, "run table exceeds the tile");
            acc = 0;
        }
        for (int g : touched) { int c = cnt[g]; cnt[g] = acc; acc += c; }
        for (size_t s = s0; s < s1; s++) {
            int l = P.lit[s];
            if (!l) continue;
            const size_t ls = s - s0;   // tile-local slot: clause ls / K, literal ls % K
            const unsigned sign = l < 0 ? 0x8000u : 0u;
            if (P.pos[s] < 0) {   // a direct primary slot
                if (!cp) P.srank[(size_t)t * rper + (ls % K) * TCt + ls / K] = (uint16_t)(0x4000u | sign);
                continue;
            }
            size_t qs = ls % K;   // the slot's rank row: with primary words, among the other literals
            if (cp && qs > P.pq[(size_t)t * TCt + ls / K]) qs--;
            uint16_t& rk = P.srank[(size_t)t * rper + qs * TCt + ls / K];
            int g = (std::abs(l) - 1) / P.vg;
            const size_t j = (size_t)cnt[g] + (size_t)(P.pos[s] - rmin[g]);
            if (rt) {
                uint32_t& w = P.runs[(size_t)t * P.run_stride + g];
                if (j == (w & 0xFFFFu)) {   // the run's lowest position
                    const long long off = P.pos[s] - P.gstart[g];
                    require(off >= 0 && off < 65536, ASAT_ERR_UNSUPPORTED, "run offset exceeds 16 bits");
                    w |= (uint32_t)off << 16;
                }
            } else {
                P.spos[r0 + j] = P.pos[s];
            }
            rk = (uint16_t)(j | sign);
        }
        for (int g : touched) { cnt[g] = 0; rmin[g] = INT_MAX; }
    }
}

// ===========================================================================
// device kernels
// ===========================================================================
template <typename T>
struct StageArgs {
    const int* lit;
    const int* pos;        // contribution position of each slot (-1: padding); generic-K pass
    const int* spos;       // tiled pass: per tile, positions in ascending order (-1: none)
    const uint16_t* srank; // tiled pass: rank of each slot in its tile's sorted list
    const unsigned* runs;  // big tiles: per tile, run table in place of spos
    int run_stride;        // words per tile of the run tables
    const int* csr_ptr;
    const int* csr_pos;    // per-variable contribution positions (gather var pass)
    const uint16_t* csr_loc;   // the same relative to the region start (LDS var pass)
    const long long* gstart;   // region starts, ngroups+1
    const int* pvar;       // direct primary slots: primary variable of each clause (processing order)
    const uint16_t* pw16;  // big tiles: primary word of each clause (in place of pvar)
    const int* pv0;        // big tiles: each tile's first primary variable
    const int* pc0;        // direct primary slots: first clause of each group's primaries, ngroups+1
    T* pcontrib;           // direct primary slots: each clause's primary term (processing order)
    T* contrib;
    T* xa;       // state buffer 0
    T* xb;       // state buffer 1 (x = parity ? xb : xa; x_new is the other)
    T* k1; T* k2; T* k3; T* k4; T* k5;
    T* tmp;      // spin stage input (n+1)
    // slot spins: between a var pass and the next clause pass, contrib holds
    // the next stage's spins in contribution layout (each slot's spin at the
    // slot's contribution position); 0: the clause pass gathers from x / tmp
    int slot_spins;
    double* partial;   // per-block error maxima: clause blocks first, then var blocks
    double* apart;     // bias term: per clause-block sums of the stage aux values
    const Ctrl* ctrl;  // null: standalone RHS (x = xa, output to k1, no tmp)
    int n, m, k, aoff;
    int nblk_clause;   // clause-pass grid (its J = 6 error partials come first)
    int vg;            // variables per group
    int var_region_cap;   // LDS var pass: elements reserved for a region
    int var_stagger;      // LDS var pass: CUs whose second workgroup starts late (0: none)
    int vg_blocks;     // gather var pass: blocks per variable group
    int ngroups;
    int clause_per_xcd;   // clause-pass chunks (BLK clauses, or TC-clause tiles) per XCD
    int ntiles;           // tiled pass: tiles
    int tile_clauses;     // tiled pass: clauses per tile
    int split;            // direct primary slots (tiled pass + LDS var pass)
    int family;           // ASAT_FAMILY_CTDS or ASAT_FAMILY_TANH
    int bias;             // sine bias term on (b > 0)
    T knorm;
    T q;                  // tanh family parameter
    T bb, alpha;          // b knorm^2 and M / N (bias term)
};

// XCD-contiguous clause-pass mapping: XCD x (blocks b with b % 8 == x) walks
// the clause chunks [x*per, (x+1)*per) in order.  Contribution positions are
// assigned in clause order inside each region, so every XCD fills its own
// contiguous stretch of each region and its L2 merges the partial lines
// before they are written back.
__device__ __forceinline__ int clause_of_thread(int per)
{
    const int b = blockIdx.x;
    return ((b & 7) * per + (b >> 3)) * BLK + threadIdx.x;
}

// XCD-affine var-pass mapping for the gather var pass: workgroups are
// dispatched to the 8 XCDs round-robin, so block b runs on XCD b % 8.  All
// vg_blocks blocks of one variable group are given to the same XCD, and each
// XCD walks its groups in order, so a group's contribution region stays in
// that XCD's L2 while its variables gather from it.  Returns the block's
// first variable (1-based) and last variable, or 0 for an idle block.
__device__ __forceinline__ int var_block_range(int vg, int vg_blocks, int ngroups, int n, int& last)
{
    const int b = blockIdx.x;
    const int xcd = b & 7, t = b >> 3;
    const int grp = (t / vg_blocks) * 8 + xcd;
    last = 0;
    if (grp >= ngroups) return 0;
    const long long g0 = (long long)grp * vg + 1;
    const long long i0 = g0 + (long long)(t % vg_blocks) * BLK;
    const long long gl = std::min<long long>(g0 + vg - 1, n);
    if (i0 > gl) return 0;
    last = (int)std::min<long long>(i0 + BLK - 1, gl);
    return (int)i0;
}

// Global loads and stores of the solver kernels: ordinary cached accesses.
// Non-temporal hints on the once-per-pass streams were measured on the
// N = 1e6 instance and cost 4% of the attempt rate: the index arrays and the
// state vectors are read again by the next passes, and the memory-side cache
// keeps part of them between passes.
template <typename V>
__device__ __forceinline__ V ld_g(const V* p)
{
    return *p;
}
template <typename V>
__device__ __forceinline__ void st_g(V* p, V v)
{
    *p = v;
}
// non-temporal accesses of the practical-rate probes (plain streams)
template <typename V>
__device__ __forceinline__ V ld_stream(const V* p)
{
    return __builtin_nontemporal_load(p);
}
template <typename V>
__device__ __forceinline__ void st_stream(V* p, V v)
{
    __builtin_nontemporal_store(v, p);
}

// ---- clause pass, compile-time K (Ex order), tiled -----------------------
// Slot ranks: bits 0-13 the rank in the tile's sorted list, bit 14 a direct
// primary slot (its rank bits 0), bit 15 the literal's sign, 0xFFFF padding.
constexpr unsigned RANK_MASK = 0x3FFFu;
constexpr unsigned RANK_PRIMARY = 0x4000u;

// Slot terms of one clause, written to LDS at the slots' ranks in the tile's
// position order (a direct primary slot's term goes to pt instead).  CTDS:
// r = K_m 2^-k a_m times the product of the other factors in KmiProd order,
// times +-2; tanh: tanh_term with r = 2^-k a_m.  Stops at the first padding
// literal (KmiFunc / KmiFuncTanh).
template <typename T, int K, int Q, bool PS>
__device__ __forceinline__ void slot_terms(const int (&l)[K], const T (&f)[K], T r, bool tanh_family, T q,
                                           const uint16_t* rank, T* s_out, T& pt)
{
    if constexpr (Q < K) {
        if (l[Q] != 0) {
            const T p = kmi_prod<T, K, Q>(f);
            T rr;
            if (tanh_family) {
                rr = tanh_term<T>(r, p, f[Q], q, l[Q] < 0, true);
            } else {
                rr = r * p;
                rr *= l[Q] < 0 ? (T)-2.0 : (T)2.0;
            }
            if (PS && (rank[Q] & RANK_PRIMARY)) pt = rr;
            else s_out[rank[Q] & RANK_MASK] = rr;
            slot_terms<T, K, Q + 1, PS>(l, f, r, tanh_family, q, rank, s_out, pt);
        }
    }
}

// Phase 1 of a tile for G of the thread's clauses (its clauses u0 .. u0+G-1,
// see k_clause_tile): K_m, the aux stage input, k_J(a_m) and the slot terms,
// which go to s_out at their ranks (direct primary terms into pt).  Every
// load that does not depend on another is issued first: the slot ranks
// (padded to whole tiles, read unguarded), the literal rows when the spins
// are gathered by variable, and the aux values.
// SS (slot spins): s_out already holds each rank's stage spin (phase 0), and
// psv the clauses' direct primary spins; the literal signs come from the
// ranks' bit 15.  Otherwise sp is the N-element stage spin vector, gathered
// by variable.
// CP (primary words, big tiles, K = 3): a clause's two stored ranks are its
// other literals' in clause order; its primary word names the primary
// literal's position and sign, from which the three slot ranks are formed.
template <typename T, int K, int J, int BT, bool SS, bool PS, int G, int CPT, bool CP>
__device__ __forceinline__ void tile_group(const StageArgs<T>& a, long long c0, int TS, int u0, const T* x, T* xnew,
                                           const T* sp, const T (&psv)[CPT], T hs, double h, double hnext,
                                           double floor, bool tanh_family, T* s_out, double& err, double& asum)
{
    int l[G][K];
    T sv[G][K];
    T ax[G], a1[G], a2[G], a3[G], a4[G], a5[G];
    uint16_t rk[G][K];
    static_assert(!CP || K == 3, "primary words: 3-literal clauses");
    unsigned pw[G], r2[G][2];
#pragma unroll
    for (int g = 0; g < G; g++) {
        const int lc = (u0 + g) * BT + (int)threadIdx.x;   // the clause in the tile (< TS: a clause of it)
        const long long m = c0 + lc;
        if constexpr (!SS) {
#pragma unroll
            for (int q = 0; q < K; q++) l[g][q] = lc < TS ? ld_g(a.lit + m * K + q) : 0;
        }
        if constexpr (CP) {
            pw[g] = lc < TS ? (unsigned)ld_g(a.pw16 + m) : (3u << PW_QSHIFT);
#pragma unroll
            for (int q = 0; q < 2; q++) r2[g][q] = lc < TS ? (unsigned)ld_g(a.srank + c0 * 2 + (long long)q * TS + lc) : 0xFFFFu;
        } else {
#pragma unroll
            for (int q = 0; q < K; q++) rk[g][q] = lc < TS ? ld_g(a.srank + c0 * K + (long long)q * TS + lc) : 0xFFFF;
        }
    }
#pragma unroll
    for (int g = 0; g < G; g++) {
        const int lc = (u0 + g) * BT + (int)threadIdx.x;
        const long long m = c0 + lc;
        const size_t ai = (size_t)a.aoff + m;
        ax[g] = a1[g] = a2[g] = a3[g] = a4[g] = a5[g] = (T)0;
        if (lc < TS && m < a.m) {
            ax[g] = ld_g(x + ai);
            if constexpr (J > 1) a1[g] = ld_g(a.k1 + ai);
            if constexpr (J > 2) a2[g] = ld_g(a.k2 + ai);
            if constexpr (J > 3) a3[g] = ld_g(a.k3 + ai);
            if constexpr (J > 4) a4[g] = ld_g(a.k4 + ai);
            if constexpr (J > 5) a5[g] = ld_g(a.k5 + ai);
        }
    }
    if constexpr (CP) {
        // primary position 3: a padding clause (its stored ranks are 0xFFFF)
#pragma unroll
        for (int g = 0; g < G; g++) {
            const unsigned pq = (pw[g] >> PW_QSHIFT) & 3u;
            const unsigned pr = RANK_PRIMARY | (pw[g] & 0x8000u);
            rk[g][0] = (uint16_t)(pq == 0 ? pr : r2[g][0]);
            rk[g][1] = (uint16_t)(pq == 1 ? pr : (pq == 0 ? r2[g][0] : r2[g][1]));
            rk[g][2] = (uint16_t)(pq == 2 ? pr : (pq == 3 ? 0xFFFFu : r2[g][1]));
        }
    }
    if constexpr (SS) {
        // each rank is read here and overwritten with its slot term below by
        // the same thread (the owner of the slot), so one buffer serves both
#pragma unroll
        for (int g = 0; g < G; g++) {
#pragma unroll
            for (int q = 0; q < K; q++) {
                const unsigned r = rk[g][q];
                l[g][q] = r == 0xFFFFu ? 0 : ((r & 0x8000u) ? -1 : 1);
                if (r == 0xFFFFu) sv[g][q] = (T)-1.0;
                else if (PS && (r & RANK_PRIMARY)) sv[g][q] = psv[u0 + g];
                else sv[g][q] = s_out[r & RANK_MASK];
            }
        }
    } else {
#pragma unroll
        for (int g = 0; g < G; g++) {
#pragma unroll
            for (int q = 0; q < K; q++) {
                const int v = l[g][q] < 0 ? -l[g][q] : l[g][q];
                sv[g][q] = v != 0 ? sp[v] : (T)-1.0;
            }
        }
    }
#pragma unroll
    for (int g = 0; g < G; g++) {
        const int lc = (u0 + g) * BT + (int)threadIdx.x;
        const long long m = c0 + lc;
        T pt = (T)0;
        if (lc < TS && m < a.m) {
            T f[K];
            T km = a.knorm;
#pragma unroll
            for (int q = 0; q < K; q++) {
                f[q] = l[g][q] < 0 ? (T)1.0 + sv[g][q] : (T)1.0 - sv[g][q];
                km *= f[q];
            }
            const size_t ai = (size_t)a.aoff + m;
            T am;
            if constexpr (J == 1) am = ax[g];
            else am = stage_input<T, J>(ax[g], a1[g], a2[g], a3[g], a4[g], a5[g], hs);
            T kj = am * km * km;
            asum += (double)am;
            if constexpr (J == 1) st_g(a.k1 + ai, kj);
            else if constexpr (J == 2) st_g(a.k2 + ai, kj);
            else if constexpr (J == 3) st_g(a.k3 + ai, kj);
            else if constexpr (J == 4) st_g(a.k4 + ai, kj);
            else if constexpr (J == 5) st_g(a.k5 + ai, kj);
            else {
                T xn;
                double e;
                final_combination<T>(ax[g], a1[g], a3[g], a4[g], a5[g], kj, hs, h, hnext, floor, xn, e);
                err = fmax(err, e);
                st_g(xnew + ai, xn);
            }
            const T r = tanh_family ? a.knorm * am : km * a.knorm * am;
            slot_terms<T, K, 0, PS>(l[g], f, r, tanh_family, a.q, rk[g], s_out, pt);
        }
        // pcontrib is padded to whole tiles: padding clauses store 0
        if constexpr (PS)
            if (lc < TS) st_g(a.pcontrib + m, pt);
    }
}

// One workgroup of BT threads per tile of TCt clauses (XCD-contiguous tile
// ranges), TCt / BT clauses per thread.
// Phase 0, a thread per rank (slot spins only): the tile's sorted positions,
// and the stage spins read through them into s_out (neighbouring lanes,
// neighbouring positions of one region); with direct primary slots also the
// thread's clauses' primary spins, from the stage spin vector.
// Phase 1, a thread per clause, in groups of G clauses (tile_group): the
// clause arithmetic, slot terms into s_out at their ranks; direct primary
// terms (PS) are stored at the clauses' own positions of pcontrib.
// Phase 2, a thread per rank: store the terms through the sorted positions.
// Big tiles (TCt = TC_BIG) hold up to 4608 clauses, a size chosen per
// problem (tile_clauses) so that the tiles fill whole rounds of resident
// workgroups: about 4100-clause tiles give each region runs of about 8
// entries per tile in place of 4, and no nearly empty last round (one
// extra tile running alone cost about 5 us per pass).  Phase 1 takes their
// clauses one at a time, which keeps them at 8 waves per SIMD without
// spills; two at a time spilled at J = 6 and ran 1-4 us slower per pass.
template <typename T, int K, int J, int BT, int TCt, bool SS, bool PS>
__global__ __launch_bounds__(BT) __attribute__((amdgpu_waves_per_eu(TCt > TC ? 8 : 1)))
void k_clause_tile(StageArgs<T> a)
{
    constexpr int CPT = (TCt + BT - 1) / BT;
    constexpr int G = TCt > TC ? 1 : (CPT < 2 ? CPT : 2);
    constexpr int CPTG = (CPT + G - 1) / G * G;   // clause slots of the groups (past CPT: none of the tile)
    constexpr int NR = PS ? K - 1 : K;
    constexpr int NP = CPT * NR;
    static_assert(TCt * NR <= (int)RANK_MASK + 1, "a tile's ranks must fit the rank field");
    // the sorted positions of big tiles wait in LDS for phase 2
    constexpr bool PL = TCt > TC;
    __shared__ double sh[BT / 64];
    __shared__ T s_out[TCt * NR];
    __shared__ int s_pos[PL ? TCt * NR : 1];
    const Ctrl* c = a.ctrl;
    int parity = 0;
    double h = 0.0, hnext = 0.0, floor = TabD::TINY;
    if (c) {
        if (!c->active) return;
        if (J == 1 && !c->step_start) return;
        parity = c->parity; h = c->h; hnext = c->hnext; floor = c->floor;
    }
    const T hs = (T)h;
    const T* x = parity ? a.xb : a.xa;
    T* xnew = parity ? a.xa : a.xb;
    const T* sp = SS ? a.contrib : ((J == 1) ? x : a.tmp);
    const bool tanh_family = a.family == ASAT_FAMILY_TANH;
    const int b = blockIdx.x;
    double err = 0.0, asum = 0.0;
    const int tile = (b & 7) * a.clause_per_xcd + (b >> 3);
    if (tile < a.ntiles) {
        const int TS = PL ? a.tile_clauses : TCt;
        const long long c0 = (long long)tile * TS;
        int ps[NP];
        T psv[CPTG];
        int pv[CPTG];
        if constexpr (SS && PS) {
            // padding clauses have primary variable 0 (or the tile's first
            // primary variable): x[0] / tmp[0], unused
            if constexpr (PL) {
                const int v0 = a.pv0[tile];
#pragma unroll
                for (int u = 0; u < CPTG; u++) {
                    const int lc = u * BT + (int)threadIdx.x;
                    pv[u] = lc < TS ? v0 + (int)(ld_g(a.pw16 + c0 + lc) & PW_VMASK) : 0;
                }
            } else {
#pragma unroll
                for (int u = 0; u < CPTG; u++) {
                    const int lc = u * BT + (int)threadIdx.x;
                    pv[u] = lc < TS ? ld_g(a.pvar + c0 + lc) : 0;
                }
            }
        }
        if constexpr (PL) {
            // the sorted positions from the tile's run table: group g's
            // ranks [j0(g), j0(g+1)) sit at consecutive positions from
            // gstart[g] + off(g); ranks past the tile's slots get -1
            const unsigned* rt = a.runs + (long long)tile * a.run_stride;
            for (int g = (int)threadIdx.x; g < a.ngroups; g += BT) {
                const unsigned w0 = ld_g(rt + g), w1 = ld_g(rt + g + 1);
                const int j0 = (int)(w0 & 0xFFFFu), j1 = (int)(w1 & 0xFFFFu);
                const int base = (int)ld_g(a.gstart + g) + (int)(w0 >> 16) - j0;
                for (int j = j0; j < j1; j++) s_pos[j] = base + j;
            }
            const int nsl = (int)(ld_g(rt + a.ngroups) & 0xFFFFu);
#pragma unroll
            for (int i = 0; i < NP; i++) {
                const int lr = i * BT + (int)threadIdx.x;
                if (lr >= nsl && lr < TS * NR) s_pos[lr] = -1;
            }
            if constexpr (SS) __syncthreads();
#pragma unroll
            for (int i = 0; i < NP; i++) {
                const int lr = i * BT + (int)threadIdx.x;
                ps[i] = (SS && lr < TS * NR) ? s_pos[lr] : -1;
            }
        } else {
            const int* sp_t = a.spos + c0 * NR;
#pragma unroll
            for (int i = 0; i < NP; i++) {
                const int lr = i * BT + (int)threadIdx.x;
                ps[i] = lr < TS * NR ? ld_g(sp_t + lr) : -1;
            }
        }
        if constexpr (SS) {
            if constexpr (PS) {
                const T* svec = (J == 1) ? x : a.tmp;
#pragma unroll
                for (int u = 0; u < CPTG; u++) psv[u] = ld_g(svec + pv[u]);
            }
            // LDS holds TCt * NR ranks; NP * BT may be more (TCt not a
            // multiple of BT), and a tile has TS * NR <= TCt * NR of them
            T so[NP];
#pragma unroll
            for (int i = 0; i < NP; i++) so[i] = ps[i] >= 0 ? sp[ps[i]] : (T)0;
#pragma unroll
            for (int i = 0; i < NP; i++) {
                const int lr = i * BT + (int)threadIdx.x;
                if (lr < TS * NR) s_out[lr] = so[i];
            }
        }
        if constexpr (SS) __syncthreads();
#pragma unroll
        for (int u0 = 0; u0 < CPT; u0 += G)
            tile_group<T, K, J, BT, SS, PS, G, CPTG, PL>(a, c0, TS, u0, x, xnew, sp, psv, hs, h, hnext, floor,
                                                          tanh_family, s_out, err, asum);
        __syncthreads();
#pragma unroll
        for (int i = 0; i < NP; i++) {
            const int lr = i * BT + (int)threadIdx.x;
            if (lr < TS * NR) {
                const int p = PL ? s_pos[lr] : ps[i];
                if (p >= 0) a.contrib[p] = s_out[lr];
            }
        }
    }
    if (a.bias) {
        const double bs = block_sum_all<BT>(asum, sh);
        if (threadIdx.x == 0) a.apart[blockIdx.x] = bs;
    }
    if constexpr (J == 6) {
        double bm = block_max<BT>(err, sh);
        if (threadIdx.x == 0) a.partial[blockIdx.x] = bm;
    }
}

// ---- clause pass, runtime K (any K <= 32; K > 10 uses CalculateAll order) --
template <typename T, int J>
__global__ __launch_bounds__(BLK) void k_clause_gen(StageArgs<T> a)
{
    __shared__ double sh[BLK / 64];
    __shared__ T fsh[GEN_KMAX * BLK];
    const Ctrl* c = a.ctrl;
    int parity = 0;
    double h = 0.0, hnext = 0.0, floor = TabD::TINY;
    if (c) {
        if (!c->active) return;
        if (J == 1 && !c->step_start) return;
        parity = c->parity; h = c->h; hnext = c->hnext; floor = c->floor;
    }
    const T hs = (T)h;
    const T* x = parity ? a.xb : a.xa;
    T* xnew = parity ? a.xa : a.xb;
    const T* sp = (J == 1) ? x : a.tmp;
    const int K = a.k;
    const int m = clause_of_thread(a.clause_per_xcd);
    T* f = fsh + threadIdx.x;   // f[q*BLK]
    double err = 0.0, asum = 0.0;
    if (m < a.m) {
        const int* lm = a.lit + (size_t)m * K;
        T km = a.knorm;
        for (int q = 0; q < K; q++) {
            T fq = spin_factor<T>(lm[q], sp);
            km *= fq;
            f[q * BLK] = fq;
        }
        const size_t ai = (size_t)a.aoff + m;
        T am, k1 = 0, k3 = 0, k4 = 0, k5 = 0;
        if constexpr (J == 1) {
            am = x[ai];
        } else {
            k1 = a.k1[ai];
            T k2 = J > 2 ? a.k2[ai] : (T)0;
            k3 = J > 3 ? a.k3[ai] : (T)0;
            k4 = J > 4 ? a.k4[ai] : (T)0;
            k5 = J > 5 ? a.k5[ai] : (T)0;
            am = stage_input<T, J>(x[ai], k1, k2, k3, k4, k5, hs);
        }
        T kj = am * km * km;
        asum = (double)am;
        if constexpr (J == 1) a.k1[ai] = kj;
        else if constexpr (J == 2) a.k2[ai] = kj;
        else if constexpr (J == 3) a.k3[ai] = kj;
        else if constexpr (J == 4) a.k4[ai] = kj;
        else if constexpr (J == 5) a.k5[ai] = kj;
        else {
            T xn;
            final_combination<T>(x[ai], k1, k3, k4, k5, kj, hs, h, hnext, floor, xn, err);
            xnew[ai] = xn;
        }
        T* out = a.contrib;
        const int* pm = a.pos + (size_t)m * K;
        if (a.family == ASAT_FAMILY_TANH) {
            // CpuSatTanh: templated helper for K = 3..10, generic loop otherwise
            const T r = a.knorm * am;
            const bool ex = K >= 3 && K <= 10;
            for (int q = 0; q < K; q++) {
                int l = lm[q];
                if (l == 0) break;
                T p = (T)1.0;
                bool first = true;
                for (int qq = 0; qq < K; qq++) {
                    if (qq == q) continue;
                    p = first ? f[qq * BLK] : p * f[qq * BLK];
                    first = false;
                }
                out[pm[q]] = tanh_term<T>(r, p, f[q * BLK], a.q, l < 0, ex);
            }
        } else if (K <= 10) {
            const T r = km * a.knorm * am;
            for (int q = 0; q < K; q++) {
                int l = lm[q];
                if (l == 0) break;
                T p = (T)1.0;
                bool first = true;
                for (int qq = 0; qq < K; qq++) {
                    if (qq == q) continue;
                    p = first ? f[qq * BLK] : p * f[qq * BLK];
                    first = false;
                }
                T rr = r * p;
                rr *= l < 0 ? (T)-2.0 : (T)2.0;
                out[pm[q]] = rr;
            }
        } else {
            for (int q = 0; q < K; q++) {
                int l = lm[q];
                if (l == 0) continue;
                T temp = km;
                for (int qq = 0; qq < K; qq++) temp *= (qq == q) ? (T)1.0 : f[qq * BLK];
                out[pm[q]] = (T)2.0 * temp * a.knorm * am * (l < 0 ? (T)-1.0 : (T)1.0);
            }
        }
    }
    if (a.bias) {
        const double bs = block_sum_all<BLK>(asum, sh);
        if (threadIdx.x == 0) a.apart[blockIdx.x] = bs;
    }
    if constexpr (J == 6) {
        double bm = block_max<BLK>(err, sh);
        if (threadIdx.x == 0) a.partial[blockIdx.x] = bm;
    }
}

// ---- var pass: deterministic per-variable sums + spin stage combination ---
// The state values the stage combination of variable i reads (x and the
// earlier stage derivatives it needs), loaded ahead of the sums.
template <typename T>
struct VarIn {
    T x, k1, k2, k3, k4, k5;
};
template <typename T, int J>
__device__ __forceinline__ VarIn<T> var_load(const StageArgs<T>& a, int i, const T* x, bool fresh)
{
    VarIn<T> v;
    v.x = ld_g(x + i);
    v.k1 = v.k2 = v.k3 = v.k4 = v.k5 = (T)0;
    if (J > 1 || !fresh) v.k1 = ld_g(a.k1 + i);
    if (J > 2 && J < 6) v.k2 = ld_g(a.k2 + i);
    if (J > 3) v.k3 = ld_g(a.k3 + i);
    if (J > 4) v.k4 = ld_g(a.k4 + i);
    if (J > 5) v.k5 = ld_g(a.k5 + i);
    return v;
}

// Stage combination of one variable, shared by both var passes.  kj is the
// variable's summed derivative (or, on a retry's J = 1, read back from k1).
// Returns the spin the next clause pass reads: stage J+1's input, or x_new.
template <typename T, int J>
__device__ __forceinline__ T var_stage(const StageArgs<T>& a, int i, bool fresh, T kj, const VarIn<T>& v, T* xnew,
                                       T hs, double h, double hnext, double floor, double& err)
{
    const T xi = v.x;
    T out = (T)0;
    if constexpr (J == 1) {
        if (fresh) st_g(a.k1 + i, kj);
        else kj = v.k1;
        out = stage_input<T, 2>(xi, kj, 0, 0, 0, 0, hs);
        if (a.tmp) a.tmp[i] = out;
    } else if constexpr (J == 2) {
        st_g(a.k2 + i, kj);
        out = stage_input<T, 3>(xi, v.k1, kj, 0, 0, 0, hs);
        a.tmp[i] = out;
    } else if constexpr (J == 3) {
        st_g(a.k3 + i, kj);
        out = stage_input<T, 4>(xi, v.k1, v.k2, kj, 0, 0, hs);
        a.tmp[i] = out;
    } else if constexpr (J == 4) {
        st_g(a.k4 + i, kj);
        out = stage_input<T, 5>(xi, v.k1, v.k2, v.k3, kj, 0, hs);
        a.tmp[i] = out;
    } else if constexpr (J == 5) {
        st_g(a.k5 + i, kj);
        out = stage_input<T, 6>(xi, v.k1, v.k2, v.k3, v.k4, kj, hs);
        a.tmp[i] = out;
    } else {
        double e;
        final_combination<T>(xi, v.k1, v.k3, v.k4, v.k5, kj, hs, h, hnext, floor, out, e);
        err = fmax(err, e);
        st_g(xnew + i, out);
    }
    return out;
}

// Bias term (AnalogSAT/cpu_sat/CpuSat.cpp:232-239, CudaSatBase.cu:75-121):
// mean(a) alpha b knorm^2 pi/2, with mean(a) the stage aux mean.  Every var
// workgroup adds the clause blocks' partial sums in the same order, so all of
// them use the same value.
template <typename T, int BLOCK>
__device__ __forceinline__ T bias_prefactor(const StageArgs<T>& a, double* sh)
{
    double v = 0.0;
    for (int i = threadIdx.x; i < a.nblk_clause; i += BLOCK) v += a.apart[i];
    const double tot = block_sum_all<BLOCK>(v, sh);
    const T mean = (T)tot / (T)a.m;
    return mean * a.alpha * a.bb * (T)1.570796326794897;
}

template <typename T>
__device__ __forceinline__ T bias_sine(T s)
{
    if constexpr (sizeof(T) == 4) return sinf(s * (T)3.141592653589793);
    else return sin(s * (T)3.141592653589793);
}

// LDS var pass: one workgroup per variable group.  The group's whole
// contribution region and its variables' 16-bit region offsets (one
// contiguous CSR range) are read into LDS with coalesced 16-byte loads, all
// issued before the first LDS write; then each thread adds its variables'
// runs from LDS in slot order (the reference's accumulation order).  With
// slot spins on, each thread then writes its variables' output spins (stage
// J+1's input, or x_new) over their own region entries, and the workgroup
// streams the region back to contrib as the next clause pass's slot spins.
// Each region entry belongs to one slot of one variable, so the entries a
// thread reads and overwrites are its own.
// FP32: two 1024-thread workgroups per CU (LDS holds two regions) need 8
// waves per SIMD, which caps the scalar registers: at J = 6 the compiler
// otherwise takes 90 of them, only one workgroup fits, and the pass took 44
// instead of 36 us.  FP64 systems fit one var workgroup per CU anyway.
template <typename T, int J>
__global__ __launch_bounds__(VBLK) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 8 : 1)))
void k_var_region(StageArgs<T> a)
{
    constexpr int V = 16 / (int)sizeof(T);
    constexpr int R = 4;              // 16-byte loads in flight per thread and round
    typedef T vT __attribute__((ext_vector_type(V)));
    typedef unsigned int v4u __attribute__((ext_vector_type(4)));
    __shared__ double sh[VBLK / 64];
    // dynamic LDS: the region (largest region x sizeof(T)), then the offsets
    extern __shared__ __attribute__((aligned(16))) unsigned char var_lds[];
    const Ctrl* c = a.ctrl;
    int parity = 0;
    double h = 0.0, hnext = 0.0, floor = TabD::TINY;
    bool fresh = true;
    if (c) {
        if (!c->active) return;
        parity = c->parity; h = c->h; hnext = c->hnext; floor = c->floor;
        fresh = (J != 1) || c->step_start;
    }
    const T hs = (T)h;
    const T* x = parity ? a.xb : a.xa;
    T* xnew = parity ? a.xa : a.xb;
    const int g = blockIdx.x;
    // The dispatcher gives each CU one workgroup before any gets a second, so
    // workgroups var_stagger .. 2 var_stagger - 1 are the second ones of the
    // first round.  Started together, a CU's two workgroups would load, add
    // and write back in lockstep, and HBM would idle while both add from LDS;
    // the second starts about 8000 cycles later, so their phases alternate
    // (N = 1e6: 31-34 -> 28-29 us per pass at J = 1..5, 36 -> 32 us at J = 6).
    if (g >= a.var_stagger && g < 2 * a.var_stagger) __builtin_amdgcn_s_sleep(127);
    const int i0 = g * a.vg + 1, i1 = (int)std::min<long long>((long long)(g + 1) * a.vg, a.n);
    T* reg = (T*)var_lds;
    const uint16_t* loc = (const uint16_t*)(var_lds + (size_t)a.var_region_cap * sizeof(T));
    T* s_next = a.slot_spins ? a.contrib : nullptr;
    const long long r0 = a.gstart[g];
    const int nreg = (int)(a.gstart[g + 1] - r0) / V;
    const int rlen = nreg * V;
    // each thread's state values, loaded before the region (groups of at most
    // 2 * VBLK variables; larger groups load them in the loop)
    constexpr int VP = 2;
    const bool pre = i1 - i0 + 1 <= VP * VBLK;
    VarIn<T> vin[VP];
    if (pre) {
#pragma unroll
        for (int u = 0; u < VP; u++) {
            const int i = i0 + (int)threadIdx.x + u * VBLK;
            if (i <= i1) vin[u] = var_load<T, J>(a, i, x, fresh);
        }
    }
    int ebase = 0;
    if (fresh || s_next) {
        // region sizes are multiples of REGION_ALIGN and the offsets of a group
        // number at most its region size; the host sized the LDS for both and
        // padded csr_loc by 8 entries for the rounded-up 16-byte loads.  A
        // retry's J = 1 pass needs only the offsets (for the slot spins).
        const int nv = fresh ? nreg : 0;
        const int e_first = a.csr_ptr[i0], e_end = a.csr_ptr[i1 + 1];
        ebase = e_first & ~7;
        const int nc = (e_end - ebase + 7) >> 3;
        const vT* src = (const vT*)(a.contrib + r0);
        // direct primary slots: the group's primary terms follow the region
        // in LDS, read from the 16-byte unit holding its first clause
        const vT* psrc = src;
        int npv = 0;
        if (a.split && fresh) {
            const int pb = a.pc0[g] / V * V;
            psrc = (const vT*)(a.pcontrib + pb);
            npv = (a.pc0[g + 1] - pb + V - 1) / V;
        }
        const int nt = nv + npv;
        const v4u* csrc = (const v4u*)(a.csr_loc + ebase);
        vT* dst = (vT*)reg;
        v4u* cdst = (v4u*)loc;
        const int nmax = nt > nc ? nt : nc;
        for (int base = 0; base < nmax; base += R * VBLK) {
            vT rb[R];
            v4u cb[R];
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int v = base + r * VBLK + (int)threadIdx.x;
                if (v < nt) rb[r] = ld_g(v < nv ? src + v : psrc + (v - nv));
                if (v < nc) cb[r] = ld_g(csrc + v);
            }
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int v = base + r * VBLK + (int)threadIdx.x;
                if (v < nt) dst[v] = rb[r];
                if (v < nc) cdst[v] = cb[r];
            }
        }
        __syncthreads();
    }
    double err = 0.0;
    const T pref = (a.bias && fresh) ? bias_prefactor<T, VBLK>(a, sh) : (T)0;
    int u = 0;
    for (int i = i0 + (int)threadIdx.x; i <= i1; i += VBLK, u++) {
        T kj = (T)0;
        const int e0 = a.csr_ptr[i] - ebase, e1 = a.csr_ptr[i + 1] - ebase;
        if (fresh) {
            // four entries in flight per round (offsets, then values), added
            // left to right; lanes past their row read element 0 and skip it
            T sum = (T)0;
            for (int e = e0; e < e1; e += 4) {
                int o[4];
                T v[4];
#pragma unroll
                for (int q = 0; q < 4; q++) o[q] = e + q < e1 ? loc[e + q] : 0;
#pragma unroll
                for (int q = 0; q < 4; q++) v[q] = reg[o[q]];
#pragma unroll
                for (int q = 0; q < 4; q++)
                    if (e + q < e1) sum += v[q];
            }
            kj = sum;
            if (a.bias) kj += pref * bias_sine<T>(J == 1 ? x[i] : a.tmp[i]);
        }
        VarIn<T> in;
        if (pre) in = u == 0 ? vin[0] : vin[1];
        else in = var_load<T, J>(a, i, x, fresh);
        const T out = var_stage<T, J>(a, i, fresh, kj, in, xnew, hs, h, hnext, floor, err);
        if (s_next)   // region entries only (direct primary entries lie past the region)
            for (int e = e0; e < e1; e += 4) {
                int o[4];
#pragma unroll
                for (int q = 0; q < 4; q++) o[q] = e + q < e1 ? loc[e + q] : rlen;
#pragma unroll
                for (int q = 0; q < 4; q++)
                    if (o[q] < rlen) reg[o[q]] = out;
            }
    }
    if (s_next) {
        // the region's entries not owned by a variable (alignment padding)
        // are written out too, with whatever the LDS holds; no slot reads them
        __syncthreads();
        const vT* src = (const vT*)reg;
        vT* dst = (vT*)(s_next + r0);
        // ordinary stores: the next clause pass reads these lines right away
        for (int v = (int)threadIdx.x; v < nreg; v += VBLK) dst[v] = src[v];
    }
    if constexpr (J == 6) {
        double bm = block_max<VBLK>(err, sh);
        if (threadIdx.x == 0) a.partial[a.nblk_clause + blockIdx.x] = bm;
    }
}

// Gather var pass, for groups whose region does not fit in LDS.  A block
// owns up to BLK consecutive variables of one group.  Their CSR entries form
// one contiguous range; the block first gathers that whole range into LDS
// with coalesced csr_pos reads and many independent contribution loads in
// flight, then every thread adds its own variable's run from LDS left to
// right (slot order, as the reference accumulates).
constexpr int VAR_STAGE = 6144;   // staged contributions per round (LDS)

template <typename T, int J>
__global__ __launch_bounds__(BLK) void k_var(StageArgs<T> a)
{
    __shared__ double sh[BLK / 64];
    __shared__ T vals[VAR_STAGE];
    const Ctrl* c = a.ctrl;
    int parity = 0;
    double h = 0.0, hnext = 0.0, floor = TabD::TINY;
    bool fresh = true;
    if (c) {
        if (!c->active) return;
        parity = c->parity; h = c->h; hnext = c->hnext; floor = c->floor;
        fresh = (J != 1) || c->step_start;
    }
    const T hs = (T)h;
    const T* x = parity ? a.xb : a.xa;
    T* xnew = parity ? a.xa : a.xb;
    int i1 = 0;
    const int i0 = var_block_range(a.vg, a.vg_blocks, a.ngroups, a.n, i1);
    const int i = (i0 != 0 && i0 + (int)threadIdx.x <= i1) ? i0 + (int)threadIdx.x : 0;
    double err = 0.0;
    const T pref = (a.bias && fresh) ? bias_prefactor<T, BLK>(a, sh) : (T)0;
    T kj = (T)0;
    if (fresh && i0 != 0) {
        const int base = a.csr_ptr[i0], end = a.csr_ptr[i1 + 1];
        const int e0 = i ? a.csr_ptr[i] : 0, e1 = i ? a.csr_ptr[i + 1] : 0;
        T sum = (T)0;
        for (int r0 = base; r0 < end; r0 += VAR_STAGE) {
            const int r1 = std::min(end, r0 + VAR_STAGE);
            for (int e = r0 + threadIdx.x; e < r1; e += BLK) vals[e - r0] = a.contrib[a.csr_pos[e]];
            __syncthreads();
            const int lo = std::max(e0, r0), hi = std::min(e1, r1);
            sum = sum_run(sum, vals, lo - r0, hi - r0);
            __syncthreads();
        }
        kj = sum;
    }
    if (a.bias && fresh && i != 0) kj += pref * bias_sine<T>(J == 1 ? x[i] : a.tmp[i]);
    if (i != 0) var_stage<T, J>(a, i, fresh, kj, var_load<T, J>(a, i, x, fresh), xnew, hs, h, hnext, floor, err);
    if constexpr (J == 6) {
        double bm = block_max<BLK>(err, sh);
        if (threadIdx.x == 0) a.partial[a.nblk_clause + blockIdx.x] = bm;
    }
}

// Wide var pass, for the gather layout when variables have long rows (at
// least WIDE_MIN_ROW contributions on average): one workgroup per variable.
// The variable's contributions pass through two LDS buffers of WIDE_STAGE
// values: while thread 0 adds one buffer in slot order (so the sum is the
// gather kernel's, and the reference's, exactly), waves 1..3 stage the next
// chunk into the other.
constexpr int WIDE_BLK = 256;
constexpr int WIDE_STAGE = 2048;
constexpr long long WIDE_MIN_ROW = 64;

template <typename T, int J>
__global__ __launch_bounds__(WIDE_BLK) void k_var_wide(StageArgs<T> a)
{
    __shared__ double sh[WIDE_BLK / 64];
    __shared__ __attribute__((aligned(16))) T vals[2][WIDE_STAGE];
    const Ctrl* c = a.ctrl;
    int parity = 0;
    double h = 0.0, hnext = 0.0, floor = TabD::TINY;
    bool fresh = true;
    if (c) {
        if (!c->active) return;
        parity = c->parity; h = c->h; hnext = c->hnext; floor = c->floor;
        fresh = (J != 1) || c->step_start;
    }
    const T hs = (T)h;
    const T* x = parity ? a.xb : a.xa;
    T* xnew = parity ? a.xa : a.xb;
    const int i = (int)blockIdx.x + 1;          // grid: one workgroup per variable 1..n
    double err = 0.0;
    const T pref = (a.bias && fresh) ? bias_prefactor<T, WIDE_BLK>(a, sh) : (T)0;
    T sum = (T)0;
    if (fresh) {
        const int e0 = a.csr_ptr[i], e1 = a.csr_ptr[i + 1];
        const int t = (int)threadIdx.x;
        {   // chunk 0, every thread
            const int r1 = std::min(e1, e0 + WIDE_STAGE);
            for (int e = e0 + t; e < r1; e += WIDE_BLK) vals[0][e - e0] = a.contrib[a.csr_pos[e]];
        }
        __syncthreads();
        for (int r0 = e0, ch = 0; r0 < e1; r0 += WIDE_STAGE, ch++) {
            const int r1 = std::min(e1, r0 + WIDE_STAGE);
            if (t >= 64) {
                // stage chunk ch + 1 into the buffer added in round ch - 1
                const int n1 = std::min(e1, r1 + WIDE_STAGE);
                T* dst = vals[(ch + 1) & 1];
                for (int e = r1 + (t - 64); e < n1; e += WIDE_BLK - 64) dst[e - r1] = a.contrib[a.csr_pos[e]];
            } else if (t == 0) {
                // 16-byte LDS reads, eight ahead; the additions stay in slot order
                constexpr int V = 16 / (int)sizeof(T);
                typedef T vT __attribute__((ext_vector_type(V)));
                const T* v = vals[ch & 1];
                const int len = r1 - r0;
                int e = 0;
                for (; e + 8 * V <= len; e += 8 * V) {
                    vT q[8];
#pragma unroll
                    for (int u = 0; u < 8; u++) q[u] = *(const vT*)(v + e + u * V);
#pragma unroll
                    for (int u = 0; u < 8; u++)
#pragma unroll
                        for (int w = 0; w < V; w++) sum += q[u][w];
                }
                for (; e < len; e++) sum += v[e];
            }
            __syncthreads();
        }
    }
    if (threadIdx.x == 0) {
        T kj = sum;
        if (a.bias && fresh) kj += pref * bias_sine<T>(J == 1 ? x[i] : a.tmp[i]);
        var_stage<T, J>(a, i, fresh, kj, var_load<T, J>(a, i, x, fresh), xnew, hs, h, hnext, floor, err);
    }
    if constexpr (J == 6) {
        double bm = block_max<WIDE_BLK>(err, sh);
        if (threadIdx.x == 0) a.partial[a.nblk_clause + blockIdx.x] = bm;
    }
}

// ---- control: max-norm + Cash-Karp step control (CpuRungeKutta.cpp:217-267)
__global__ __launch_bounds__(CTRL_BLK) void k_control(Ctrl* c, const double* partial, int np)
{
    __shared__ double sh[CTRL_BLK / 64];
    if (!c->active) return;
    double v = 0.0;
    for (int i = threadIdx.x; i < np; i += CTRL_BLK) v = fmax(v, partial[i]);
    v = block_max<CTRL_BLK>(v, sh);
    if (threadIdx.x != 0) return;
    const double errmax = v / c->eps;
    const double h = c->h;
    c->errmax = errmax;
    c->attempts += 1;
    c->rhscount += c->step_start ? 6 : 5;
    if (errmax <= 1.0) {
        double hn = errmax > TabD::ERRCON ? TabD::SAFETY * h * pow_ctrl(errmax, 5) : 5.0 * h;
        c->hdid = h;
        c->time += h;
        if (c->hlog && c->steps < c->hlog_len) c->hlog[c->steps] = h;
        c->steps += 1;
        c->parity ^= 1;
        c->step_start = 1;
        c->hnext = hn;
        c->h = hn;
        if (c->steps >= c->target) c->active = 0;
    } else {
        double htemp = TabD::SAFETY * h * pow_ctrl(errmax, 4);
        double hn = h >= 0.0 ? fmax(htemp, TabD::HFACTOR * h) : fmin(htemp, TabD::HFACTOR * h);
        c->h = hn;
        c->step_start = 0;
        if (c->time + hn == c->time) {
            c->underflow = 1;
            c->active = 0;
        }
    }
}

__global__ void k_ctrl_begin(Ctrl* c, long long count, int parity)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        c->target = c->steps + count;
        c->parity = parity;
        c->active = (c->underflow || count <= 0) ? 0 : 1;
    }
}

__global__ void k_ctrl_set_eps(Ctrl* c, double eps, double floor)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        c->eps = eps;
        c->floor = floor;
    }
}

// ---- violated-clause count (CpuSat::CalculateClauses semantics) ----------
template <typename T>
__global__ __launch_bounds__(BLK) void k_violations(const int* lit, int m, int k, const T* x, int* out)
{
    __shared__ int sh[BLK / 64];
    const int j = blockIdx.x * BLK + threadIdx.x;
    int viol = 0;
    if (j < m) {
        int sat = 0;
        for (int q = 0; q < k && !sat; q++) {
            int l = lit[(size_t)j * k + q];
            if (l == 0) continue;            // padding literal is constant FALSE
            int t = x[l < 0 ? -l : l] > (T)0;
            sat = l < 0 ? !t : t;
        }
        viol = !sat;
    }
    int s = block_sum<BLK>(viol, sh);
    if (threadIdx.x == 0 && s) atomicAdd(out, s);
}

// ---- random numbers: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) ----
__device__ __forceinline__ void philox4x32_10(uint32_t ctr[4], uint32_t key0, uint32_t key1)
{
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)M0 * ctr[0];
        uint64_t p1 = (uint64_t)M1 * ctr[2];
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ ctr[1] ^ key0;
        uint32_t n2 = hi0 ^ ctr[3] ^ key1;
        ctr[0] = n0; ctr[1] = lo1; ctr[2] = n2; ctr[3] = lo0;
        key0 += W0; key1 += W1;
    }
}

template <typename T>
__global__ void k_uniform(T* out, int len, unsigned long long seed, unsigned long long offset)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= len) return;
    unsigned long long idx = offset + (unsigned long long)i;
    uint32_t ctr[4] = {(uint32_t)idx, (uint32_t)(idx >> 32), 0x5EEDu, 0u};
    philox4x32_10(ctr, (uint32_t)seed, (uint32_t)(seed >> 32));
    if constexpr (sizeof(T) == 8) {
        uint64_t v = ((uint64_t)ctr[0] << 32) | ctr[1];
        out[i] = (double)(v >> 11) * 0x1.0p-53;
    } else {
        out[i] = (float)(ctr[0] >> 8) * 0x1.0p-24f;
    }
}

// s = 2u - 1 for spins, a = u + 1 for aux (CpuSat.cpp:127-132); input uniforms
// are laid out in the public order [s, a] inside the state buffer's spin area
// Slot spins of a state: each variable's spin at all of its contribution
// positions (the layout the LDS var pass writes after each stage).
template <typename T>
__global__ __launch_bounds__(BLK) void k_slot_fill(const T* x, const int* csr_ptr, const uint16_t* csr_loc,
                                                   const long long* gstart, int vg, int n, T* sl)
{
    const int i = blockIdx.x * BLK + (int)threadIdx.x + 1;
    if (i > n) return;
    const T v = x[i];
    const int g = (i - 1) / vg;
    T* base = sl + gstart[g];
    const long long rlen = gstart[g + 1] - gstart[g];   // direct primary entries lie past the region
    for (int e = csr_ptr[i]; e < csr_ptr[i + 1]; e++) {
        const int o = csr_loc[e];
        if (o < rlen) base[o] = v;
    }
}

template <typename T>
__global__ void k_set_initial(T* x, int n, int m, int aoff, const T* u, const int* perm)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        x[1 + i] = u[i] * (T)2.0 - (T)1.0;
    } else if (i < (long long)n + m) {
        const int j = (int)(i - n);
        x[aoff + j] = u[n + (perm ? perm[j] : j)] + (T)1.0;
    }
}

// plain streaming copy (16 B per lane): the reference point for HBM byte counters
// Practical HBM rate probes: 16 bytes per lane, STREAM_U loads in flight per
// thread before the first use, non-temporal.
constexpr int STREAM_U = 4;
typedef int stream_v4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void k_stream_copy(const stream_v4* __restrict__ src, stream_v4* __restrict__ dst,
                                                     long long n16)
{
    const long long stride = (long long)gridDim.x * 256 * STREAM_U;
    for (long long base = (long long)blockIdx.x * 256 * STREAM_U + threadIdx.x; base < n16; base += stride) {
        stream_v4 r[STREAM_U];
#pragma unroll
        for (int u = 0; u < STREAM_U; u++) {
            const long long i = base + u * 256;
            if (i < n16) r[u] = ld_stream(src + i);
        }
#pragma unroll
        for (int u = 0; u < STREAM_U; u++) {
            const long long i = base + u * 256;
            if (i < n16) st_stream(dst + i, r[u]);
        }
    }
}
// read-only pass; `out` is written only if the data hash to a value that a
// zero-filled source never gives, which keeps the loads
__global__ __launch_bounds__(256) void k_stream_read(const stream_v4* __restrict__ src, int* out, long long n16)
{
    const long long stride = (long long)gridDim.x * 256 * STREAM_U;
    int acc = 0;
    for (long long base = (long long)blockIdx.x * 256 * STREAM_U + threadIdx.x; base < n16; base += stride) {
        stream_v4 r[STREAM_U];
#pragma unroll
        for (int u = 0; u < STREAM_U; u++) {
            const long long i = base + u * 256;
            r[u] = i < n16 ? ld_stream(src + i) : stream_v4{0, 0, 0, 0};
        }
#pragma unroll
        for (int u = 0; u < STREAM_U; u++) acc ^= r[u].x ^ r[u].y ^ r[u].z ^ r[u].w;
    }
    if (acc == 0x7fffffff) out[0] = acc;
}

__global__ void k_copy16(const int4* __restrict__ src, int4* __restrict__ dst, long long n16)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (long long)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

template <typename T>
__global__ void k_fill_const(T* x)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) x[0] = (T)-1.0;
}

// ===========================================================================
// host objects
// ===========================================================================
template <typename T>
static void dfree(T*& p)
{
    if (p) { (void)hipFree((void*)p); p = nullptr; }
}

struct Sys;

struct State {
    Sys* sys = nullptr;
    void* d = nullptr;     // device buffer of the system's precision
    int alloc = 0;
    unsigned long long gen = 0;   // the system's problem generation when the state was made
};

struct Sys {
    int prec = ASAT_F64;
    int esize = 8;
    double b = 0.0;
    int n = 0, m = 0, k = 0, aoff = 0, alloc = 0;
    long long nslots = 0;
    hipStream_t stream = nullptr;
    int* d_lit = nullptr;
    int* d_pos = nullptr;        // generic-K clause pass only
    int* d_spos = nullptr;       // tiled clause pass (K = 3, 4)
    uint16_t* d_srank = nullptr;
    unsigned* d_runs = nullptr;  // big tiles: run tables in place of d_spos
    int run_stride = 0;
    int* d_csr_ptr = nullptr;
    int* d_csr_pos = nullptr;    // gather var pass only
    uint16_t* d_csr_loc = nullptr;   // LDS var pass only
    long long* d_gstart = nullptr;
    int* d_pvar = nullptr;       // direct primary slots: primary variable per clause
    uint16_t* d_pw16 = nullptr;  // big tiles: primary word per clause (in place of d_pvar)
    int* d_pv0 = nullptr;        // big tiles: first primary variable of each tile
    int* d_pc0 = nullptr;        // direct primary slots: first clause of each group's primaries
    void* d_pcontrib = nullptr;  // direct primary slots: primary term per clause
    bool split = false;          // direct primary slots (tiled pass + LDS var pass)
    int* d_perm = nullptr;       // processing position -> clause index (null: identity)
    std::vector<int> perm;       // host copy, for the public <-> internal a-block order
    // one-workgroup path (asat_small.h) for problems whose integrator fits in
    // LDS: each slot's position in var-major order (a variable's terms in the
    // reference's accumulation order) and each variable's range of positions;
    // small_lds = 0 when the problem is too big
    int* d_vptr = nullptr;
    int* d_spos_small = nullptr;
    int nlist = 0;
    int runlen = 0;                // uniform run length of the lists (0: ragged)
    long long small_lds = 0;
    int vg = 0, ngroups = 0, vg_blocks = 0;
    int family = ASAT_FAMILY_CTDS;
    double q = 1.1;              // tanh family parameter (CpuSatTanh.cpp:39)
    bool tiled = false;          // clause pass: tiled kernel (K = 3, 4) or generic
    int tc = TC;                 // tiled pass: clauses per tile (TC or TC_BIG)
    bool lds_var = false;        // var pass: LDS region kernel or gather kernel
    bool wide_var = false;       // gather path with long rows: one workgroup per variable
    int var_lds = 0;             // LDS region kernel: dynamic LDS bytes (region + offsets)
    int var_stagger = 0;         // LDS region kernel: CUs whose second workgroup starts late
    int var_region_cap = 0;      // LDS region kernel: elements reserved for a region
    void* d_contrib = nullptr;
    double* d_apart = nullptr;   // bias term: per clause-block aux sums
    int* d_count = nullptr;
    bool ready = false;
    // bumped by every SetProblem: states and integrators record the value
    // they were made for, and a later problem (even one of the same size)
    // invalidates them
    unsigned long long gen = 0;

    void free_problem()
    {
        dfree(d_lit); dfree(d_pos); dfree(d_spos); dfree(d_srank); dfree(d_csr_ptr); dfree(d_csr_pos);
        dfree(d_runs);
        run_stride = 0;
        dfree(d_csr_loc); dfree(d_gstart); dfree(d_perm); dfree(d_contrib); dfree(d_apart);
        dfree(d_pvar); dfree(d_pc0); dfree(d_pcontrib); dfree(d_pw16); dfree(d_pv0);
        split = false;
        dfree(d_vptr); dfree(d_spos_small);
        perm.clear();
        nlist = 0;
        runlen = 0;
        small_lds = 0;
    }
    ~Sys()
    {
        free_problem();
        dfree(d_count);
        if (stream) (void)hipStreamDestroy(stream);
    }
    void require_ready() const { require(ready, ASAT_ERR_NOT_CONFIGURED, "SAT problem has not been set"); }
};

// clause-pass and var-pass grid sizes of a configured system
static int clause_grid(const Sys* s) { return 8 * ceil_div(s->tiled ? ceil_div(s->m, s->tc) : ceil_div(s->m, BLK), 8); }
static int var_grid(const Sys* s)
{
    if (s->lds_var) return s->ngroups;
    if (s->wide_var) return s->n;
    return 8 * s->vg_blocks * ceil_div(s->ngroups, 8);
}

struct Rng {
    unsigned long long seed = 0, offset = 0;
};

template <typename V>
static void upload(V*& d, const std::vector<V>& h)
{
    HIP_CHECK(hipMalloc((void**)&d, sizeof(V) * std::max<size_t>(h.size(), 1)));
    if (!h.empty()) HIP_CHECK(hipMemcpy(d, h.data(), sizeof(V) * h.size(), hipMemcpyHostToDevice));
}

// Variable groups for the LDS var pass: about 2048 (float) / 1024 (double)
// variables per group, a multiple of 256 groups once there are more than
// 256 (one workgroup per CU per round), at least 256 groups while groups
// keep 256+ variables; halved until the largest region fits in LDS.  When no
// size fits (a variable with more occurrences than the LDS holds), the gather
// var pass runs with ~400 KB regions.
// LDS of the region var pass: the region and its variables' uint16 offsets
// (at most one per region element), plus room for the rounded 16-byte loads
static int var_lds_bytes(long long max_region, int esize)
{
    return (int)(max_region * esize + round_up(max_region * 2 + 32, 16));
}

static bool choose_layout(HostProblem& P, int esize)
{
    const long long cap = (VAR_LDS_BYTES - 48) / (esize + 2);
    const long long target = esize == 4 ? 2048 : 1024;
    long long G = std::max(ceil_div((long long)P.n, target), std::min(256, ceil_div((long long)P.n, 256)));
    if (G > 256) G = round_up(G, 256);
    int vg = ceil_div((long long)P.n, G);
    // direct primary slots wherever the tiled pass meets the LDS var pass
    // (primary variables exist only for the tiled pass's clause order)
    const int pv = 16 / esize;
    for (;;) {
        build_layout(P, vg, pv);
        if (P.max_region <= cap && !P.csr_loc.empty()) return true;
        if (vg <= 64) break;
        vg = (vg + 1) / 2;
    }
    build_layout(P, esize == 4 ? 8192 : 4096, 0);
    return false;
}

// Var-major term positions of the one-workgroup path, when its LDS
// footprint fits: variable v's terms occupy positions vptr[v] .. vptr[v+1]-1,
// its slots taken with clauses in input order and literals in clause order
// (the order CpuSat::CalculateAllEx adds them, AnalogSAT/cpu_sat/CpuSat.cpp:
// 206-229); spos[j*K + q] is the position of processing-order slot (j, q).
static void build_small_lists(Sys* s, const HostProblem& P)
{
    const int n = P.n, m = P.m, K = P.k;
    std::vector<int> vptr((size_t)n + 2), spos((size_t)m * K), visit;
    if (!P.perm.empty()) {
        visit.resize((size_t)m);
        for (int p = 0; p < m; p++) visit[(size_t)P.perm[p]] = p;
    }
    long long nl = 0;
    const int runlen = small_build_lists(P.lit.data(), visit.empty() ? nullptr : visit.data(), n, m, K, s->alloc,
                                         s->esize, vptr.data(), spos.data(), &nl);
    const long long bytes = small_lds_bytes(n, m, K, s->alloc, nl, s->esize);
    if (bytes > SMALL_LDS_MAX) return;
    upload(s->d_vptr, vptr);
    upload(s->d_spos_small, spos);
    s->nlist = (int)nl;
    s->runlen = runlen;
    s->small_lds = bytes;
}

static void sys_set_problem(Sys* s, const int* lits, const int* lens, int nclauses)
{
    HostProblem P;
    parse_clauses(lits, lens, nclauses, P);
    const bool tiled = P.k == 3 || P.k == 4;
    if (tiled && REORDER_CLAUSES) reorder_clauses(P);
    const bool lds_var = choose_layout(P, s->esize);
    int ncu = 0;
    {
        int dev = 0;
        HIP_CHECK(hipGetDevice(&dev));
        HIP_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    }
    // FP32 3-literal clauses with direct primary slots: big tiles (the
    // kernel takes their clauses one at a time, and two workgroups still fit
    // on a CU), sized so that they fill whole rounds of the 2 x CUs resident
    // workgroups: N = 1e6, M = 4.2e6 gets 1010 tiles of 4160 clauses, two rounds.
    // Problems too small to give every resident workgroup 2048 clauses keep
    // 2048-clause tiles: with few waves per CU a thread's clauses taken one
    // after another expose their load latency (N = 2000: 16-18 us per pass
    // against 5-7 us; M = 6e5 in 1216-clause tiles: 11 us against 8 us)
    P.tc = TC;
    if (tiled && P.split && s->esize == 4 && P.k == 3) {
        const long long slots = 2LL * (ncu > 0 ? ncu : 1);
        const long long rounds = std::max<long long>(1, ceil_div((long long)P.m, slots * TC_BIG));
        const long long ts = round_up(ceil_div((long long)P.m, rounds * slots), 64);
        P.tc = ts < TC ? TC : (int)std::min<long long>(ts, TC_BIG);
        // big tiles rebuild their sorted positions from run tables, a word
        // per variable group: worth it while that is at most a quarter of
        // the positions (ngroups + 1 <= tc (K - 1) / 4)
        if (P.tc != TC && (long long)(P.ngroups + 1) * 4 > (long long)P.tc * (P.k - 1)) P.tc = TC;
        // primary words hold a clause's primary variable relative to its
        // tile's first: tiles spanning more variables keep 2048 clauses
        for (long long j0 = 0; P.tc != TC && j0 < P.m; j0 += P.tc) {
            const long long j1 = std::min<long long>(P.m, j0 + P.tc) - 1;
            if (P.pvar[(size_t)j1] - P.pvar[(size_t)j0] > (int)PW_VMASK) P.tc = TC;
        }
    }
    if (tiled && P.tc != TC) spread_run_banks(P, 16 / s->esize);
    if (tiled) build_tile_order(P);
    HIP_CHECK(hipStreamSynchronize(s->stream));
    s->free_problem();
    s->ready = false;
    s->gen++;
    s->n = P.n; s->m = P.m; s->k = P.k;
    s->aoff = (int)round_up((long long)P.n + 1, 64);
    s->alloc = (int)(s->aoff + round_up(P.m, 64));
    s->nslots = (long long)P.m * P.k;
    s->vg = P.vg;
    s->ngroups = P.ngroups;
    s->vg_blocks = ceil_div(P.vg, BLK);
    s->tiled = tiled;
    s->tc = P.tc;
    s->lds_var = lds_var;
    s->split = P.split;
    require(!P.split || (tiled && lds_var), ASAT_ERR_INVALID_ARGUMENT, "direct primary slots without their passes");
    // few variables with long rows (dense problems such as Ramsey encodings):
    // the gather kernel would add a block's rows one after another, so each
    // variable gets a workgroup of its own
    s->wide_var = !lds_var && s->nslots >= WIDE_MIN_ROW * (long long)P.n;
    s->var_region_cap = lds_var ? (int)P.max_region : 0;
    s->var_lds = lds_var ? var_lds_bytes(P.max_region, s->esize) : 0;
    // two FP32 var workgroups per CU (LDS for both; the FP64 kernel's
    // registers allow one) and more groups than CUs: stagger the pairs
    s->var_stagger = 0;
    if (lds_var && s->esize == 4 && 2LL * s->var_lds <= VAR_LDS_CU && ncu > 0 && P.ngroups > ncu)
        s->var_stagger = ncu;
    // clause arrays padded to whole tiles (padding literal 0, position -1)
    const size_t mpad = (size_t)round_up(P.m, P.tc);
    P.lit.resize(mpad * P.k, 0);
    upload(s->d_lit, P.lit);
    if (tiled) {
        if (P.runs.empty()) upload(s->d_spos, P.spos);
        else upload(s->d_runs, P.runs);
        s->run_stride = P.run_stride;
        upload(s->d_srank, P.srank);
        if (P.split) {
            if (P.pw16.empty()) {
                P.pvar.resize(mpad, 0);   // padding clauses: variable 0
                upload(s->d_pvar, P.pvar);
            } else {
                upload(s->d_pw16, P.pw16);
                upload(s->d_pv0, P.pv0);
            }
            upload(s->d_pc0, P.pc0);
            // one term per clause padded to whole tiles, plus a 16-byte unit
            const size_t np = mpad + 16;
            HIP_CHECK(hipMalloc(&s->d_pcontrib, (size_t)s->esize * np));
            HIP_CHECK(hipMemset(s->d_pcontrib, 0, (size_t)s->esize * np));
        }
    } else {
        P.pos.resize(mpad * P.k, -1);
        upload(s->d_pos, P.pos);
    }
    upload(s->d_csr_ptr, P.csr_ptr);
    if (lds_var) {
        P.csr_loc.resize(P.csr_loc.size() + 8, 0);   // rounded-up 16-byte loads of the last range
        upload(s->d_csr_loc, P.csr_loc);
    }
    else upload(s->d_csr_pos, P.csr_pos);
    upload(s->d_gstart, P.gstart);
    build_small_lists(s, P);
    if (!P.perm.empty()) upload(s->d_perm, P.perm);
    s->perm = std::move(P.perm);
    const size_t ncontrib = (size_t)P.gstart[P.ngroups] + REGION_ALIGN;
    HIP_CHECK(hipMalloc(&s->d_contrib, (size_t)s->esize * ncontrib));
    HIP_CHECK(hipMemset(s->d_contrib, 0, (size_t)s->esize * ncontrib));
    HIP_CHECK(hipMalloc((void**)&s->d_apart, sizeof(double) * clause_grid(s)));
    HIP_CHECK(hipMemset(s->d_apart, 0, sizeof(double) * clause_grid(s)));
    s->ready = true;
}

static State* state_new(Sys* s)
{
    s->require_ready();
    std::unique_ptr<State> st(new State());
    st->sys = s;
    st->alloc = s->alloc;
    st->gen = s->gen;
    HIP_CHECK(hipMalloc(&st->d, (size_t)s->esize * s->alloc));
    HIP_CHECK(hipMemsetAsync(st->d, 0, (size_t)s->esize * s->alloc, s->stream));
    if (s->prec == ASAT_F64) hipLaunchKernelGGL(k_fill_const<double>, dim3(1), dim3(64), 0, s->stream, (double*)st->d);
    else hipLaunchKernelGGL(k_fill_const<float>, dim3(1), dim3(64), 0, s->stream, (float*)st->d);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(s->stream));
    return st.release();
}

static void state_check(const State* st)
{
    require(st && st->sys, ASAT_ERR_INVALID_ARGUMENT, "null state");
    require(st->alloc == st->sys->alloc && st->gen == st->sys->gen && st->sys->ready, ASAT_ERR_INVALID_ARGUMENT,
            "state does not belong to the current problem");
}

static void state_set_public(State* st, const void* pub)
{
    state_check(st);
    Sys* s = st->sys;
    const char* p = (const char*)pub;
    char* d = (char*)st->d;
    const size_t e = (size_t)s->esize;
    std::vector<char> ab;
    const char* pa = p + e * s->n;
    if (!s->perm.empty()) {   // public clause order -> processing order
        ab.resize(e * s->m);
        for (int j = 0; j < s->m; j++) memcpy(ab.data() + e * j, pa + e * s->perm[j], e);
        pa = ab.data();
    }
    HIP_CHECK(hipMemcpyAsync(d + e, p, e * s->n, hipMemcpyHostToDevice, s->stream));
    HIP_CHECK(hipMemcpyAsync(d + e * s->aoff, pa, e * s->m, hipMemcpyHostToDevice, s->stream));
    if (s->prec == ASAT_F64) hipLaunchKernelGGL(k_fill_const<double>, dim3(1), dim3(64), 0, s->stream, (double*)st->d);
    else hipLaunchKernelGGL(k_fill_const<float>, dim3(1), dim3(64), 0, s->stream, (float*)st->d);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(s->stream));
}

static void state_get_public(const State* st, void* pub)
{
    state_check(st);
    Sys* s = st->sys;
    char* p = (char*)pub;
    const char* d = (const char*)st->d;
    const size_t e = (size_t)s->esize;
    HIP_CHECK(hipStreamSynchronize(s->stream));
    HIP_CHECK(hipMemcpy(p, d + e, e * s->n, hipMemcpyDeviceToHost));
    if (s->perm.empty()) {
        HIP_CHECK(hipMemcpy(p + e * s->n, d + e * s->aoff, e * s->m, hipMemcpyDeviceToHost));
    } else {   // processing order -> public clause order
        std::vector<char> ab(e * s->m);
        HIP_CHECK(hipMemcpy(ab.data(), d + e * s->aoff, e * s->m, hipMemcpyDeviceToHost));
        for (int j = 0; j < s->m; j++) memcpy(p + e * (s->n + s->perm[j]), ab.data() + e * j, e);
    }
}

template <typename T>
static StageArgs<T> stage_args(Sys* s, T* xa, T* xb, T* const* kk, T* tmp, double* partial, const Ctrl* c)
{
    StageArgs<T> a;
    memset(&a, 0, sizeof(a));   // padding too: the arguments are compared bytewise (captured attempts)
    a.lit = s->d_lit; a.pos = s->d_pos; a.spos = s->d_spos; a.srank = s->d_srank;
    a.runs = s->d_runs; a.run_stride = s->run_stride;
    a.csr_ptr = s->d_csr_ptr; a.csr_pos = s->d_csr_pos; a.csr_loc = s->d_csr_loc; a.gstart = s->d_gstart;
    a.pvar = s->d_pvar; a.pw16 = s->d_pw16; a.pv0 = s->d_pv0; a.pc0 = s->d_pc0; a.pcontrib = (T*)s->d_pcontrib; a.split = s->split ? 1 : 0;
    a.contrib = (T*)s->d_contrib;
    a.xa = xa; a.xb = xb;
    a.k1 = kk[0]; a.k2 = kk[1]; a.k3 = kk[2]; a.k4 = kk[3]; a.k5 = kk[4];
    a.tmp = tmp; a.partial = partial; a.ctrl = c;
    a.slot_spins = 0;
    a.n = s->n; a.m = s->m; a.k = s->k; a.aoff = s->aoff;
    a.ntiles = ceil_div(s->m, s->tc);
    a.tile_clauses = s->tc;
    a.clause_per_xcd = ceil_div(s->tiled ? a.ntiles : ceil_div(s->m, BLK), 8);
    a.nblk_clause = clause_grid(s);
    a.vg = s->vg;
    a.var_region_cap = s->var_region_cap;
    a.var_stagger = s->var_stagger;
    a.vg_blocks = s->vg_blocks;
    a.ngroups = s->ngroups;
    a.knorm = (T)std::pow(2.0, -s->k);
    a.family = s->family;
    a.q = (T)s->q;
    a.apart = s->d_apart;
    a.bb = (T)s->b * a.knorm * a.knorm;
    a.bias = a.bb > (T)0;
    a.alpha = (T)s->m / (T)s->n;
    return a;
}


template <typename T, int J>
static void launch_clause(const StageArgs<T>& a, const Sys* s, hipStream_t st)
{
    dim3 g(a.nblk_clause), b(BLK);
    const bool tiled = s->tiled;
    const bool ss = a.slot_spins != 0;
    const bool ps = a.split != 0;   // slot spins come with direct primary slots (both need the LDS var pass)
    require(!ss || ps, ASAT_ERR_INVALID_ARGUMENT, "slot spins without direct primary slots");
    const dim3 tb(TILE_BT);
    constexpr int TB = TC_BIG, TS = TC;
    const bool big = s->tc != TC;
    require(!big || (sizeof(T) == 4 && ps && a.k == 3 && a.pw16 && a.pv0 && a.runs), ASAT_ERR_INVALID_ARGUMENT,
            "big clause tiles without their layout");
    if constexpr (sizeof(T) == 4) {
        if (tiled && big) {
            if (ss) hipLaunchKernelGGL((k_clause_tile<T, 3, J, TILE_BT, TB, true, true>), g, tb, 0, st, a);
            else hipLaunchKernelGGL((k_clause_tile<T, 3, J, TILE_BT, TB, false, true>), g, tb, 0, st, a);
            return;
        }
    }
    if (tiled && a.k == 3 && ss) hipLaunchKernelGGL((k_clause_tile<T, 3, J, TILE_BT, TS, true, true>), g, tb, 0, st, a);
    else if (tiled && a.k == 4 && ss) hipLaunchKernelGGL((k_clause_tile<T, 4, J, TILE_BT, TS, true, true>), g, tb, 0, st, a);
    else if (tiled && a.k == 3 && ps) hipLaunchKernelGGL((k_clause_tile<T, 3, J, TILE_BT, TS, false, true>), g, tb, 0, st, a);
    else if (tiled && a.k == 4 && ps) hipLaunchKernelGGL((k_clause_tile<T, 4, J, TILE_BT, TS, false, true>), g, tb, 0, st, a);
    else if (tiled && a.k == 3) hipLaunchKernelGGL((k_clause_tile<T, 3, J, TILE_BT, TS, false, false>), g, tb, 0, st, a);
    else if (tiled && a.k == 4) hipLaunchKernelGGL((k_clause_tile<T, 4, J, TILE_BT, TS, false, false>), g, tb, 0, st, a);
    else hipLaunchKernelGGL((k_clause_gen<T, J>), g, b, 0, st, a);
}

// allow the LDS var pass dynamic LDS beyond the default 64 KB, once per
// instantiation (before the first launch, and outside any stream capture)
template <typename T, int J>
static void var_lds_attr()
{
    static const hipError_t attr = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_var_region<T, J>),
                                                       hipFuncAttributeMaxDynamicSharedMemorySize, VAR_LDS_BYTES);
    HIP_CHECK(attr);
}

template <typename T>
static void var_lds_attrs()
{
    var_lds_attr<T, 1>(); var_lds_attr<T, 2>(); var_lds_attr<T, 3>();
    var_lds_attr<T, 4>(); var_lds_attr<T, 5>(); var_lds_attr<T, 6>();
}

template <typename T, int J>
static void launch_var(const StageArgs<T>& a, const Sys* s, hipStream_t st)
{
    if (s->lds_var) {
        var_lds_attr<T, J>();
        hipLaunchKernelGGL((k_var_region<T, J>), dim3(var_grid(s)), dim3(VBLK), (size_t)s->var_lds, st, a);
    } else if (s->wide_var) {
        hipLaunchKernelGGL((k_var_wide<T, J>), dim3(var_grid(s)), dim3(WIDE_BLK), 0, st, a);
    } else {
        hipLaunchKernelGGL((k_var<T, J>), dim3(var_grid(s)), dim3(BLK), 0, st, a);
    }
}

// One attempt.  tkind >= 0: events e0 / e1 are recorded just before and
// after the kernel of that kind (clause J = kind + 1 for 0..5, var J =
// kind - 5 for 6..11, control for 12).
template <typename T>
static void launch_attempt(const StageArgs<T>& a, const Sys* s, Ctrl* ctrl, const double* partial, int np,
                           hipStream_t st, int tkind = -1, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr)
{
    auto mark = [&](int kind, hipEvent_t e) {
        if (kind == tkind) HIP_CHECK(hipEventRecord(e, st));
    };
    mark(0, e0); launch_clause<T, 1>(a, s, st); mark(0, e1);
    mark(6, e0); launch_var<T, 1>(a, s, st); mark(6, e1);
    mark(1, e0); launch_clause<T, 2>(a, s, st); mark(1, e1);
    mark(7, e0); launch_var<T, 2>(a, s, st); mark(7, e1);
    mark(2, e0); launch_clause<T, 3>(a, s, st); mark(2, e1);
    mark(8, e0); launch_var<T, 3>(a, s, st); mark(8, e1);
    mark(3, e0); launch_clause<T, 4>(a, s, st); mark(3, e1);
    mark(9, e0); launch_var<T, 4>(a, s, st); mark(9, e1);
    mark(4, e0); launch_clause<T, 5>(a, s, st); mark(4, e1);
    mark(10, e0); launch_var<T, 5>(a, s, st); mark(10, e1);
    mark(5, e0); launch_clause<T, 6>(a, s, st); mark(5, e1);
    mark(11, e0); launch_var<T, 6>(a, s, st); mark(11, e1);
    mark(12, e0);
    hipLaunchKernelGGL(k_control, dim3(1), dim3(CTRL_BLK), 0, st, ctrl, partial, np);
    mark(12, e1);
}

template <typename T>
static void sys_rhs(Sys* s, State* dxdt, const State* x)
{
    s->require_ready();
    state_check(dxdt); state_check(x);
    T* kk[5] = {(T*)dxdt->d, nullptr, nullptr, nullptr, nullptr};
    StageArgs<T> a = stage_args<T>(s, (T*)x->d, nullptr, kk, nullptr, nullptr, nullptr);
    // dxdt.SetZero(): padding and the constant slot read as 0
    HIP_CHECK(hipMemsetAsync(dxdt->d, 0, (size_t)s->esize * s->alloc, s->stream));
    launch_clause<T, 1>(a, s, s->stream);
    launch_var<T, 1>(a, s, s->stream);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(s->stream));
}

template <typename T>
static int sys_violations(Sys* s, const State* x)
{
    s->require_ready();
    state_check(x);
    HIP_CHECK(hipMemsetAsync(s->d_count, 0, sizeof(int), s->stream));
    hipLaunchKernelGGL(k_violations<T>, dim3(ceil_div(s->m, BLK)), dim3(BLK), 0, s->stream, s->d_lit, s->m, s->k,
                       (const T*)x->d, s->d_count);
    HIP_CHECK(hipGetLastError());
    int v = 0;
    HIP_CHECK(hipMemcpyAsync(&v, s->d_count, sizeof(int), hipMemcpyDeviceToHost, s->stream));
    HIP_CHECK(hipStreamSynchronize(s->stream));
    return v;
}

// ---------------------------------------------------------------------------
// integrator
// ---------------------------------------------------------------------------
struct RK {
    Sys* sys = nullptr;
    double eps = 1e-4;      // CudaRungeKutta() default (CudaRungeKutta.cu:175)
    double floor = 1e-20;   // xscal constant term (reference TINY)
    double* hlog = nullptr; // optional device log of accepted step sizes (set before Reset)
    long long hlog_len = 0;
    int alloc = 0, prec = -1;
    unsigned long long gen = 0;   // the system's problem generation at Configure
    void* k[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    void* tmp = nullptr;
    int tmp_n = 0;            // tmp holds n+1 values of the problem at Configure
    double* partial = nullptr;
    int npartial = 0;
    Ctrl* d_ctrl = nullptr;
    Ctrl* h_ctrl = nullptr;   // pinned mirror
    int small_mode = 1;              // 1: one-workgroup path where the problem fits, 0: never
    SmallSolve* d_solve = nullptr;   // device SatSolver loop state (one-workgroup path)
    SmallSolve* h_solve = nullptr;   // pinned mirror
    void* pair[2] = {nullptr, nullptr};
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    float last_ms = -1.f;
    bool configured = false;
    // kernel timer (measurement hook): one event pair per attempt
    int timer_kind = -1;
    std::vector<hipEvent_t> tev;
    double timer_ms = 0.0;
    long long timer_launches = 0;
    // small problems: GRAPH_ATTEMPTS attempts captured once and replayed; the
    // key is the exact launch arguments the graph was captured with
    hipGraphExec_t gexec = nullptr;
    std::vector<unsigned char> gkey;

    void drop_graph()
    {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        gexec = nullptr;
        gkey.clear();
    }
    void free_all()
    {
        drop_graph();
        for (auto& p : k) dfree(p);
        dfree(tmp); dfree(partial); dfree(d_ctrl); dfree(d_solve);
        if (h_ctrl) { (void)hipHostFree(h_ctrl); h_ctrl = nullptr; }
        if (h_solve) { (void)hipHostFree(h_solve); h_solve = nullptr; }
    }
    ~RK()
    {
        free_all();
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        for (hipEvent_t e : tev) (void)hipEventDestroy(e);
    }
    void require_configured() const
    {
        require(configured && sys && sys->ready, ASAT_ERR_NOT_CONFIGURED,
                "the RK4 solver has not been configured with an ODE yet");
        // scratch (k1..k5, tmp, partials) is sized for the problem present at
        // Configure; a new problem needs a new Configure
        require(gen == sys->gen, ASAT_ERR_NOT_CONFIGURED,
                "the SAT problem changed since the RK4 solver was configured");
    }
};

// CudaRungeKutta::Reset (h = 1e-9, clock and counters cleared): the control
// block is set on the host and copied in stream order (pinned memory), so a
// solve that follows waits for no round trip.  Every later host write of
// h_ctrl is a device-to-host copy behind this one on the same stream.
static void rk_reset(RK* r)
{
    r->require_configured();
    Ctrl* c = r->h_ctrl;
    memset(c, 0, sizeof(Ctrl));
    c->h = 1e-9; c->hnext = 1e-9; c->eps = r->eps;
    c->floor = r->floor; c->hlog = r->hlog; c->hlog_len = r->hlog_len;
    c->step_start = 1;
    HIP_CHECK(hipMemcpyAsync(r->d_ctrl, c, sizeof(Ctrl), hipMemcpyHostToDevice, r->sys->stream));
    r->pair[0] = r->pair[1] = nullptr;
}

static bool small_path(const RK* r);

static void rk_configure(RK* r, Sys* s)
{
    s->require_ready();
    HIP_CHECK(hipStreamSynchronize(s->stream));
    const int npartial = clause_grid(s) + var_grid(s);
    if (r->configured && r->sys == s && r->alloc == s->alloc && r->prec == s->prec && r->npartial == npartial &&
        r->tmp_n == s->n) {
        // same sizes (SatSolver configures before every Solve): keep the
        // buffers, clear them and reset the control block
        r->drop_graph();
        r->sys = s;
        r->gen = s->gen;
        // the one-workgroup path keeps everything in LDS and reads none of
        // these buffers (small instances: a caller solving start after start)
        if (!small_path(r)) {
            const size_t bytes = (size_t)s->esize * s->alloc;
            for (auto& p : r->k) HIP_CHECK(hipMemsetAsync(p, 0, bytes, s->stream));
            HIP_CHECK(hipMemsetAsync(r->tmp, 0, (size_t)s->esize * (s->n + 1), s->stream));
            HIP_CHECK(hipMemsetAsync(r->partial, 0, sizeof(double) * r->npartial, s->stream));
        }
        rk_reset(r);
        return;
    }
    r->free_all();
    r->configured = false;
    r->sys = s;
    r->alloc = s->alloc;
    r->gen = s->gen;
    r->prec = s->prec;
    const size_t bytes = (size_t)s->esize * s->alloc;
    for (auto& p : r->k) {
        HIP_CHECK(hipMalloc(&p, bytes));
        HIP_CHECK(hipMemset(p, 0, bytes));
    }
    HIP_CHECK(hipMalloc(&r->tmp, (size_t)s->esize * (s->n + 1)));
    r->tmp_n = s->n;
    HIP_CHECK(hipMemset(r->tmp, 0, (size_t)s->esize * (s->n + 1)));
    r->npartial = npartial;
    HIP_CHECK(hipMalloc((void**)&r->partial, sizeof(double) * r->npartial));
    HIP_CHECK(hipMemset(r->partial, 0, sizeof(double) * r->npartial));
    HIP_CHECK(hipMalloc((void**)&r->d_ctrl, sizeof(Ctrl)));
    HIP_CHECK(hipHostMalloc((void**)&r->h_ctrl, sizeof(Ctrl), hipHostMallocDefault));
    memset(r->h_ctrl, 0, sizeof(Ctrl));
    HIP_CHECK(hipMalloc((void**)&r->d_solve, sizeof(SmallSolve)));
    HIP_CHECK(hipHostMalloc((void**)&r->h_solve, sizeof(SmallSolve), hipHostMallocDefault));
    memset(r->h_solve, 0, sizeof(SmallSolve));
    if (!r->ev0) HIP_CHECK(hipEventCreate(&r->ev0));
    if (!r->ev1) HIP_CHECK(hipEventCreate(&r->ev1));
    r->configured = true;
    rk_reset(r);
}

// Stage arguments of the integrator.  With slot spins (tiled clause pass and
// LDS var pass), contrib is first filled with the slot spins of the current
// state x, which the caller may have changed since the last Step.  Inside
// Step the var passes keep contrib holding the next clause pass's spins: a
// rejected attempt's J = 1 clause pass is skipped, and its J = 1 var pass
// (which reuses k1 and reads no contributions) writes stage 2's spins.
template <typename T>
static StageArgs<T> rk_stage_args(RK* r, int parity)
{
    Sys* s = r->sys;
    T* kk[5] = {(T*)r->k[0], (T*)r->k[1], (T*)r->k[2], (T*)r->k[3], (T*)r->k[4]};
    StageArgs<T> a = stage_args<T>(s, (T*)r->pair[0], (T*)r->pair[1], kk, (T*)r->tmp, r->partial, r->d_ctrl);
    if (s->tiled && s->lds_var) {
        a.slot_spins = 1;
        hipLaunchKernelGGL(k_slot_fill<T>, dim3(ceil_div(s->n, BLK)), dim3(BLK), 0, s->stream,
                           (const T*)r->pair[parity], s->d_csr_ptr, s->d_csr_loc, s->d_gstart, s->vg, s->n,
                           (T*)s->d_contrib);
        HIP_CHECK(hipGetLastError());
    }
    return a;
}

// Launch-bound problems (few literal slots) replay a captured graph of
// GRAPH_ATTEMPTS attempts instead of launching 13 kernels per attempt, as
// long as a round still needs that many; the rest are launched one by one.
constexpr long long GRAPH_MAX_SLOTS = 1LL << 20;
constexpr int GRAPH_ATTEMPTS = 8;

template <typename T>
static void ensure_graph(RK* r, const StageArgs<T>& a)
{
    Sys* s = r->sys;
    hipStream_t st = s->stream;
    // the key: the arguments and every launch size the attempts depend on
    std::vector<unsigned char> key(sizeof(a) + 4 * sizeof(int));
    memcpy(key.data(), &a, sizeof(a));
    const int dims[4] = {clause_grid(s), var_grid(s), s->var_lds, s->tiled ? 1 : 0};
    memcpy(key.data() + sizeof(a), dims, sizeof(dims));
    if (r->gexec && r->gkey == key) return;
    r->drop_graph();
    var_lds_attrs<T>();
    hipGraph_t g = nullptr;
    HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    try {
        for (int i = 0; i < GRAPH_ATTEMPTS; i++) launch_attempt<T>(a, s, r->d_ctrl, r->partial, r->npartial, st);
    } catch (...) {
        (void)hipStreamEndCapture(st, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    HIP_CHECK(hipStreamEndCapture(st, &g));
    const hipError_t e = hipGraphInstantiate(&r->gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIP_CHECK(e);
    r->gkey = std::move(key);
}

template <typename T>
static int rk_step_impl(RK* r, State* olds, State* news, int count)
{
    Sys* s = r->sys;
    hipStream_t st = s->stream;
    // canonical ping-pong pair: parity 0 means x lives in pair[0]
    if (!((r->pair[0] == olds->d && r->pair[1] == news->d) || (r->pair[0] == news->d && r->pair[1] == olds->d))) {
        r->pair[0] = olds->d;
        r->pair[1] = news->d;
    }
    const int parity = (r->pair[0] == olds->d) ? 0 : 1;
    HIP_CHECK(hipEventRecord(r->ev0, st));
    StageArgs<T> a = rk_stage_args<T>(r, parity);
    hipLaunchKernelGGL(k_ctrl_begin, dim3(1), dim3(64), 0, st, r->d_ctrl, (long long)count, parity);
    long long remaining = count;
    const long long target = r->h_ctrl->steps + count;
    long long last_attempts = r->h_ctrl->attempts;
    const bool graphed = r->timer_kind < 0 && s->nslots <= GRAPH_MAX_SLOTS && count >= GRAPH_ATTEMPTS;
    if (graphed) ensure_graph<T>(r, a);
    while (remaining > 0) {
        const int tk = r->timer_kind;
        if (graphed) {
            // whole graphs while they fit, the rest one attempt at a time
            const long long full = remaining / GRAPH_ATTEMPTS;
            for (long long i = 0; i < full; i++) HIP_CHECK(hipGraphLaunch(r->gexec, st));
            for (long long i = full * GRAPH_ATTEMPTS; i < remaining; i++)
                launch_attempt<T>(a, s, r->d_ctrl, r->partial, r->npartial, st);
        } else {
            if (tk >= 0)
                while ((long long)r->tev.size() < 2 * remaining) {
                    // timing-only events, read after the stream synchronises:
                    // no system-scope fence, so recording them does not write
                    // back and invalidate the caches between the kernels
                    hipEvent_t e;
                    HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
                    r->tev.push_back(e);
                }
            for (long long i = 0; i < remaining; i++) {
                if (tk >= 0)
                    launch_attempt<T>(a, s, r->d_ctrl, r->partial, r->npartial, st, tk, r->tev[2 * i],
                                      r->tev[2 * i + 1]);
                else
                    launch_attempt<T>(a, s, r->d_ctrl, r->partial, r->npartial, st);
            }
        }
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipMemcpyAsync(r->h_ctrl, r->d_ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
        if (tk >= 0)
            for (long long i = 0; i < remaining; i++) {
                float ms = 0.f;
                HIP_CHECK(hipEventElapsedTime(&ms, r->tev[2 * i], r->tev[2 * i + 1]));
                r->timer_ms += ms;
                r->timer_launches++;
            }
        if (r->h_ctrl->underflow) break;
        remaining = target - r->h_ctrl->steps;
        // every enqueued attempt either accepts, rejects or underflows; no
        // progress here means the control block was left inactive
        require(r->h_ctrl->attempts > last_attempts || remaining <= 0, ASAT_ERR_HIP,
                "integrator made no progress (control block inactive)");
        last_attempts = r->h_ctrl->attempts;
    }
    HIP_CHECK(hipEventRecord(r->ev1, st));
    HIP_CHECK(hipEventSynchronize(r->ev1));
    HIP_CHECK(hipEventElapsedTime(&r->last_ms, r->ev0, r->ev1));
    // hand the newest state to `olds` (the reference swaps the device pointers)
    const int par = r->h_ctrl->parity;
    void* xcur = r->pair[par];
    void* xother = r->pair[par ^ 1];
    olds->d = xcur;
    news->d = xother;
    return r->h_ctrl->underflow ? ASAT_UNDERFLOW : ASAT_OK;
}

// ---------------------------------------------------------------------------
// one-workgroup path (asat_small.h): problems whose whole integrator fits in
// one workgroup's LDS and that use no bias term.  The kernel timer (a
// per-kernel measurement hook) needs the multi-kernel path.
// ---------------------------------------------------------------------------
static bool small_path(const RK* r)
{
    const Sys* s = r->sys;
    return r->small_mode && s->small_lds > 0 && !(s->b > 0.0) && r->timer_kind < 0;
}

template <typename T, int KT, int BS, int FAM = ASAT_FAMILY_CTDS>
static void small_kernel_launch(const SmallArgs& a, size_t lds, hipStream_t st)
{
    static const hipError_t attr = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_small<T, KT, BS, FAM>),
                                                       hipFuncAttributeMaxDynamicSharedMemorySize, SMALL_LDS_MAX);
    HIP_CHECK(attr);
    hipLaunchKernelGGL((k_small<T, KT, BS, FAM>), dim3(1), dim3(BS), lds, st, a);
    HIP_CHECK(hipGetLastError());
}

// 256 threads (a clause and a variable each) up to 512 clauses and
// variables, 1024 threads above; the tanh family runs the generic kernel
template <typename T>
static void small_launch(const Sys* s, const SmallArgs& a, hipStream_t st)
{
    const size_t lds = (size_t)s->small_lds;
    const bool wide = std::max(s->m, s->n) > 512;
    if (s->family == ASAT_FAMILY_TANH) {
        if (wide) small_kernel_launch<T, 0, 1024, ASAT_FAMILY_TANH>(a, lds, st);
        else small_kernel_launch<T, 0, 256, ASAT_FAMILY_TANH>(a, lds, st);
    } else if (s->k == 3) {
        if (wide) small_kernel_launch<T, 3, 1024>(a, lds, st);
        else small_kernel_launch<T, 3, 256>(a, lds, st);
    } else if (s->k == 4) {
        if (wide) small_kernel_launch<T, 4, 1024>(a, lds, st);
        else small_kernel_launch<T, 4, 256>(a, lds, st);
    } else {
        if (wide) small_kernel_launch<T, 0, 1024>(a, lds, st);
        else small_kernel_launch<T, 0, 256>(a, lds, st);
    }
}

static SmallArgs small_args(const RK* r, void* x)
{
    const Sys* s = r->sys;
    SmallArgs a;
    memset(&a, 0, sizeof(a));
    a.lit = s->d_lit; a.vptr = s->d_vptr; a.spos = s->d_spos_small;
    a.x = x;
    a.ctrl = r->d_ctrl;
    a.n = s->n; a.m = s->m; a.k = s->k; a.aoff = s->aoff; a.alloc = s->alloc; a.nlist = s->nlist;
    a.runlen = s->runlen;
    a.family = s->family;
    a.q = s->q;
    a.knorm = std::pow(2.0, -s->k);
    a.batch = 1;
    return a;
}

// Launches the persistent kernel until it reports done: each launch runs at
// most SMALL_ATTEMPT_CAP attempts and leaves its state in the control block
// (and the solver block), so the next launch resumes where it stopped.
template <typename T>
static void small_run(RK* r, const SmallArgs& a)
{
    Sys* s = r->sys;
    hipStream_t st = s->stream;
    HIP_CHECK(hipEventRecord(r->ev0, st));
    long long last = r->h_ctrl->attempts;
    for (;;) {
        small_launch<T>(s, a, st);
        HIP_CHECK(hipMemcpyAsync(r->h_ctrl, r->d_ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, st));
        if (a.solve) HIP_CHECK(hipMemcpyAsync(r->h_solve, r->d_solve, sizeof(SmallSolve), hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
        if (r->h_ctrl->underflow) break;
        if (a.solve ? r->h_solve->result >= 0 : r->h_ctrl->steps >= a.target) break;
        require(r->h_ctrl->attempts > last, ASAT_ERR_HIP, "integrator made no progress");
        last = r->h_ctrl->attempts;
    }
    HIP_CHECK(hipEventRecord(r->ev1, st));
    HIP_CHECK(hipEventSynchronize(r->ev1));
    HIP_CHECK(hipEventElapsedTime(&r->last_ms, r->ev0, r->ev1));
}

static void rk_check_states(const RK* r, const State* olds, const State* news)
{
    state_check(olds); state_check(news);
    require(olds != news && olds->d != news->d, ASAT_ERR_INVALID_ARGUMENT, "old and new state must differ");
    require(olds->sys == r->sys && news->sys == r->sys, ASAT_ERR_INVALID_ARGUMENT,
            "states belong to another system");
}

static int rk_step(RK* r, State* olds, State* news, int count)
{
    r->require_configured();
    rk_check_states(r, olds, news);
    require(count >= 0, ASAT_ERR_INVALID_ARGUMENT, "negative step count");
    if (r->h_ctrl->underflow) return ASAT_UNDERFLOW;
    if (count == 0) return ASAT_OK;
    if (small_path(r)) {
        // the state stays in olds; news is not touched
        SmallArgs a = small_args(r, olds->d);
        a.target = r->h_ctrl->steps + count;
        if (r->prec == ASAT_F64) small_run<double>(r, a);
        else small_run<float>(r, a);
        return r->h_ctrl->underflow ? ASAT_UNDERFLOW : ASAT_OK;
    }
    return r->prec == ASAT_F64 ? rk_step_impl<double>(r, olds, news, count)
                               : rk_step_impl<float>(r, olds, news, count);
}

// SatSolver::DoSolve without a callback (AnalogSAT/solver/SatSolver.h:313-331):
// batches of `batch` accepted steps, the violated-clause count of `olds`
// after each, ending on SAT, underflow, or when the time (compared as a
// float if time_float) or the solver's step count reaches its limit before a
// batch.  The solver's counters start at zero; the integrator's continue
// from where they are (SatSolver::Solve configures it first).  On the
// one-workgroup path the whole loop runs in the persistent kernel; otherwise
// it runs here on the host over rk_step and the violation kernel.
static void rk_solve(RK* r, State* olds, State* news, int batch, long long max_steps, double max_time,
                     int time_float, asat_solve_result_t* out)
{
    r->require_configured();
    rk_check_states(r, olds, news);
    require(batch >= 1, ASAT_ERR_INVALID_ARGUMENT, "Batch size must be positive");
    require(max_steps >= 1, ASAT_ERR_INVALID_ARGUMENT, "Max iterations must be positive");
    require(max_time > 0.0, ASAT_ERR_INVALID_ARGUMENT, "Max time must be positive");
    memset(out, 0, sizeof(*out));
    if (small_path(r) && !r->h_ctrl->underflow) {
        SmallSolve init;
        memset(&init, 0, sizeof(init));
        init.result = -1;
        *r->h_solve = init;
        HIP_CHECK(hipMemcpyAsync(r->d_solve, r->h_solve, sizeof(SmallSolve), hipMemcpyHostToDevice, r->sys->stream));
        SmallArgs a = small_args(r, olds->d);
        a.solve = r->d_solve;
        a.batch = batch;
        a.max_steps = max_steps;
        a.max_time = max_time;
        a.time_float = time_float;
        if (r->prec == ASAT_F64) small_run<double>(r, a);
        else small_run<float>(r, a);
        out->result = r->h_solve->result;
        out->steps = r->h_solve->solver_steps;
        out->last_violations = r->h_solve->last_viol;
        out->on_device = 1;
        return;
    }
    auto tcmp = [&](double t) { return time_float ? (double)(float)t : t; };
    long long steps = 0;
    double time = 0.0;
    int lastv = 0, result = -1;
    while (tcmp(time) < max_time && steps < max_steps) {
        const long long before = r->h_ctrl->steps;
        if (rk_step(r, olds, news, batch) == ASAT_UNDERFLOW) {
            result = SMALL_UNDERFLOW;
            break;
        }
        steps += r->h_ctrl->steps - before;
        time = r->h_ctrl->time;
        lastv = r->prec == ASAT_F64 ? sys_violations<double>(r->sys, olds) : sys_violations<float>(r->sys, olds);
        if (lastv == 0) {
            result = SMALL_SAT;
            break;
        }
    }
    if (result < 0) {
        if (tcmp(time) >= max_time) result = SMALL_MAXTIME;
        else if (steps >= max_steps) result = SMALL_MAXITER;
        else result = 5;   // SAT_UNKNOWN
    }
    out->result = result;
    out->steps = steps;
    out->last_violations = lastv;
    out->on_device = 0;
}

// Times each kernel kind of one attempt over `reps` launches with HIP events
// on the integrator's stream (order: clause J=1..6, var J=1..6, control).
// The control block is saved and restored, so the integrator state is unchanged;
// the scratch buffers (k1..k5, tmp, x_new) are overwritten, which is harmless
// because the next attempt starts a fresh step (step_start == 1 after Step).
template <typename T>
static void rk_profile_impl(RK* r, State* olds, State* news, int reps, float* us)
{
    Sys* s = r->sys;
    hipStream_t st = s->stream;
    if (!((r->pair[0] == olds->d && r->pair[1] == news->d) || (r->pair[0] == news->d && r->pair[1] == olds->d))) {
        r->pair[0] = olds->d;
        r->pair[1] = news->d;
    }
    const int parity = (r->pair[0] == olds->d) ? 0 : 1;
    StageArgs<T> a = rk_stage_args<T>(r, parity);
    Ctrl* backup = nullptr;
    HIP_CHECK(hipMalloc((void**)&backup, sizeof(Ctrl)));
    HIP_CHECK(hipMemcpyAsync(backup, r->d_ctrl, sizeof(Ctrl), hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(k_ctrl_begin, dim3(1), dim3(64), 0, st, r->d_ctrl, (long long)1 << 40, parity);
    auto timeit = [&](int slot, auto&& launch) {
        launch();   // warm
        HIP_CHECK(hipEventRecord(r->ev0, st));
        for (int i = 0; i < reps; i++) launch();
        HIP_CHECK(hipEventRecord(r->ev1, st));
        HIP_CHECK(hipEventSynchronize(r->ev1));
        float ms = 0.f;
        HIP_CHECK(hipEventElapsedTime(&ms, r->ev0, r->ev1));
        us[slot] = ms * 1000.f / (float)reps;
    };
    timeit(0, [&] { launch_clause<T, 1>(a, s, st); });
    timeit(1, [&] { launch_clause<T, 2>(a, s, st); });
    timeit(2, [&] { launch_clause<T, 3>(a, s, st); });
    timeit(3, [&] { launch_clause<T, 4>(a, s, st); });
    timeit(4, [&] { launch_clause<T, 5>(a, s, st); });
    timeit(5, [&] { launch_clause<T, 6>(a, s, st); });
    timeit(6, [&] { launch_var<T, 1>(a, s, st); });
    timeit(7, [&] { launch_var<T, 2>(a, s, st); });
    timeit(8, [&] { launch_var<T, 3>(a, s, st); });
    timeit(9, [&] { launch_var<T, 4>(a, s, st); });
    timeit(10, [&] { launch_var<T, 5>(a, s, st); });
    timeit(11, [&] { launch_var<T, 6>(a, s, st); });
    timeit(12, [&] { hipLaunchKernelGGL(k_control, dim3(1), dim3(CTRL_BLK), 0, st, r->d_ctrl, r->partial, r->npartial); });
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(r->d_ctrl, backup, sizeof(Ctrl), hipMemcpyDeviceToDevice, st));
    HIP_CHECK(hipStreamSynchronize(st));
    (void)hipFree(backup);
}

}  // namespace asat

// ===========================================================================
// C ABI
// ===========================================================================
using namespace asat;

struct asat_sys { Sys s; };
struct asat_state { State st; };
struct asat_rk { RK r; };
struct asat_rng { Rng g; };

#define API_BEGIN try {
#define API_END                                                  \
    }                                                            \
    catch (const asat::Error& e) {                               \
        asat::g_last_error = e.what();                           \
        return e.code;                                           \
    }                                                            \
    catch (const std::bad_alloc&) {                              \
        asat::g_last_error = "host allocation failed";           \
        return ASAT_ERR_HIP;                                     \
    }                                                            \
    catch (const std::exception& e) {                            \
        asat::g_last_error = e.what();                           \
        return ASAT_ERR_HIP;                                     \
    }                                                            \
    return ASAT_OK;

#define CHECK_PTR(p) require((p) != nullptr, ASAT_ERR_INVALID_ARGUMENT, "null argument: " #p)

extern "C" {

const char* asat_last_error(void) { return asat::g_last_error.c_str(); }

const char* asat_version(void) { return "analogsat-mi355x 0.1 (gfx950)"; }

int asat_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int asat_sys_create(int precision, asat_sys_t** out)
{
    API_BEGIN
    CHECK_PTR(out);
    require(precision == ASAT_F64 || precision == ASAT_F32, ASAT_ERR_INVALID_ARGUMENT, "unknown precision");
    std::unique_ptr<asat_sys> p(new asat_sys());
    p->s.prec = precision;
    p->s.esize = precision == ASAT_F64 ? 8 : 4;
    HIP_CHECK(hipStreamCreateWithFlags(&p->s.stream, hipStreamNonBlocking));
    HIP_CHECK(hipMalloc((void**)&p->s.d_count, sizeof(int)));
    *out = p.release();
    API_END
}

int asat_sys_destroy(asat_sys_t* sys)
{
    API_BEGIN
    delete sys;
    API_END
}

int asat_sys_set_problem(asat_sys_t* sys, const int* literals, const int* lengths, int nclauses)
{
    API_BEGIN
    CHECK_PTR(sys);
    sys_set_problem(&sys->s, literals, lengths, nclauses);
    API_END
}

int asat_sys_get_nmk(const asat_sys_t* sys, int* n, int* m, int* k)
{
    API_BEGIN
    CHECK_PTR(sys);
    if (n) *n = sys->s.n;
    if (m) *m = sys->s.m;
    if (k) *k = sys->s.k;
    API_END
}

int asat_sys_set_b(asat_sys_t* sys, double b)
{
    API_BEGIN
    CHECK_PTR(sys);
    sys->s.b = b;
    API_END
}

int asat_sys_get_b(const asat_sys_t* sys, double* b)
{
    API_BEGIN
    CHECK_PTR(sys); CHECK_PTR(b);
    *b = sys->s.b;
    API_END
}

int asat_sys_rhs(asat_sys_t* sys, asat_state_t* dxdt, const asat_state_t* x, double t)
{
    (void)t;   // autonomous system (the reference ignores time too)
    API_BEGIN
    CHECK_PTR(sys); CHECK_PTR(dxdt); CHECK_PTR(x);
    if (sys->s.prec == ASAT_F64) sys_rhs<double>(&sys->s, &dxdt->st, &x->st);
    else sys_rhs<float>(&sys->s, &dxdt->st, &x->st);
    API_END
}

int asat_sys_violations(asat_sys_t* sys, const asat_state_t* x, int* count)
{
    API_BEGIN
    CHECK_PTR(sys); CHECK_PTR(x); CHECK_PTR(count);
    *count = sys->s.prec == ASAT_F64 ? sys_violations<double>(&sys->s, &x->st)
                                     : sys_violations<float>(&sys->s, &x->st);
    API_END
}

int asat_state_create(asat_sys_t* sys, asat_state_t** out)
{
    API_BEGIN
    CHECK_PTR(sys); CHECK_PTR(out);
    std::unique_ptr<State> st(state_new(&sys->s));
    asat_state* w = new asat_state();
    w->st = *st;
    st->d = nullptr;
    *out = w;
    API_END
}

int asat_state_destroy(asat_state_t* st)
{
    API_BEGIN
    if (st) {
        if (st->st.d) (void)hipFree(st->st.d);
        delete st;
    }
    API_END
}

int asat_state_set_public(asat_state_t* st, const void* host_pub)
{
    API_BEGIN
    CHECK_PTR(st); CHECK_PTR(host_pub);
    state_set_public(&st->st, host_pub);
    API_END
}

int asat_state_get_public(const asat_state_t* st, void* host_pub)
{
    API_BEGIN
    CHECK_PTR(st); CHECK_PTR(host_pub);
    state_get_public(&st->st, host_pub);
    API_END
}

int asat_state_get_boolean(const asat_state_t* st, unsigned char* out)
{
    API_BEGIN
    CHECK_PTR(st); CHECK_PTR(out);
    const Sys* s = st->st.sys;
    std::vector<char> buf((size_t)s->esize * (s->n + s->m));
    state_get_public(&st->st, buf.data());
    for (int i = 0; i < s->n; i++)
        out[i] = s->prec == ASAT_F64 ? (((double*)buf.data())[i] > 0) : (((float*)buf.data())[i] > 0);
    API_END
}

int asat_state_copy(asat_state_t* dst, const asat_state_t* src)
{
    API_BEGIN
    CHECK_PTR(dst); CHECK_PTR(src);
    state_check(&dst->st); state_check(&src->st);
    require(dst->st.sys == src->st.sys, ASAT_ERR_INVALID_ARGUMENT, "states belong to different systems");
    Sys* s = dst->st.sys;
    HIP_CHECK(hipMemcpyAsync(dst->st.d, src->st.d, (size_t)s->esize * s->alloc, hipMemcpyDeviceToDevice, s->stream));
    HIP_CHECK(hipStreamSynchronize(s->stream));
    API_END
}

int asat_state_set_zero(asat_state_t* st)
{
    API_BEGIN
    CHECK_PTR(st);
    state_check(&st->st);
    Sys* s = st->st.sys;
    HIP_CHECK(hipMemsetAsync(st->st.d, 0, (size_t)s->esize * s->alloc, s->stream));
    HIP_CHECK(hipStreamSynchronize(s->stream));
    API_END
}

int asat_state_alloc_size(const asat_state_t* st, int* size)
{
    API_BEGIN
    CHECK_PTR(st); CHECK_PTR(size);
    *size = st->st.alloc;
    API_END
}

int asat_state_device_ptr(asat_state_t* st, void** dev_ptr, int* aux_offset)
{
    API_BEGIN
    CHECK_PTR(st); CHECK_PTR(dev_ptr);
    *dev_ptr = st->st.d;
    if (aux_offset) *aux_offset = st->st.sys->aoff;
    API_END
}

int asat_sys_set_family(asat_sys_t* sys, int family)
{
    API_BEGIN
    CHECK_PTR(sys);
    require(family == ASAT_FAMILY_CTDS || family == ASAT_FAMILY_TANH, ASAT_ERR_INVALID_ARGUMENT,
            "unknown right-hand-side family");
    HIP_CHECK(hipStreamSynchronize(sys->s.stream));
    sys->s.family = family;
    API_END
}

int asat_sys_get_family(const asat_sys_t* sys, int* family)
{
    API_BEGIN
    CHECK_PTR(sys); CHECK_PTR(family);
    *family = sys->s.family;
    API_END
}

int asat_sys_set_q(asat_sys_t* sys, double q)
{
    API_BEGIN
    CHECK_PTR(sys);
    require(std::isfinite(q), ASAT_ERR_INVALID_ARGUMENT, "q must be finite");
    HIP_CHECK(hipStreamSynchronize(sys->s.stream));
    sys->s.q = q;
    API_END
}

int asat_sys_get_q(const asat_sys_t* sys, double* q)
{
    API_BEGIN
    CHECK_PTR(sys); CHECK_PTR(q);
    *q = sys->s.q;
    API_END
}

int asat_sys_clause_order(const asat_sys_t* sys, int* order)
{
    API_BEGIN
    CHECK_PTR(sys); CHECK_PTR(order);
    const Sys* s = &sys->s;
    s->require_ready();
    for (int j = 0; j < s->m; j++) order[j] = s->perm.empty() ? j : s->perm[j];
    API_END
}

int asat_sys_layout(const asat_sys_t* sys, asat_layout_t* out)
{
    API_BEGIN
    CHECK_PTR(sys); CHECK_PTR(out);
    const Sys* s = &sys->s;
    s->require_ready();
    out->clause_pass = s->tiled ? ASAT_CLAUSE_TILED : ASAT_CLAUSE_GENERIC;
    out->var_pass = s->lds_var ? ASAT_VAR_REGION : (s->wide_var ? ASAT_VAR_WIDE : ASAT_VAR_GATHER);
    out->groups = s->ngroups;
    out->group_vars = s->vg;
    out->tiles = s->tiled ? (int)ceil_div(s->m, s->tc) : 0;
    out->tile_clauses = s->tiled ? s->tc : 0;
    out->clause_grid = clause_grid(s);
    out->var_grid = var_grid(s);
    out->var_lds_bytes = s->var_lds;
    out->slot_spins = (s->tiled && s->lds_var) ? 1 : 0;
    out->direct_primary = s->split ? 1 : 0;
    out->run_tables = s->d_runs ? 1 : 0;
    out->primary_words = s->d_pw16 ? 1 : 0;
    API_END
}

int asat_sys_set_random_state(asat_sys_t* sys, asat_state_t* x, asat_rng_t* rng)
{
    API_BEGIN
    CHECK_PTR(sys); CHECK_PTR(x); CHECK_PTR(rng);
    Sys* s = &sys->s;
    s->require_ready();
    state_check(&x->st);
    const int len = s->n + s->m;
    void* u = nullptr;
    HIP_CHECK(hipMalloc(&u, (size_t)s->esize * len));
    dim3 g(ceil_div(len, 256)), b(256);
    if (s->prec == ASAT_F64) {
        hipLaunchKernelGGL(k_uniform<double>, g, b, 0, s->stream, (double*)u, len, rng->g.seed, rng->g.offset);
        hipLaunchKernelGGL(k_set_initial<double>, g, b, 0, s->stream, (double*)x->st.d, s->n, s->m, s->aoff,
                           (const double*)u, (const int*)s->d_perm);
    } else {
        hipLaunchKernelGGL(k_uniform<float>, g, b, 0, s->stream, (float*)u, len, rng->g.seed, rng->g.offset);
        hipLaunchKernelGGL(k_set_initial<float>, g, b, 0, s->stream, (float*)x->st.d, s->n, s->m, s->aoff,
                           (const float*)u, (const int*)s->d_perm);
    }
    rng->g.offset += (unsigned long long)len;
    hipError_t e = hipGetLastError();
    hipError_t e2 = hipStreamSynchronize(s->stream);
    (void)hipFree(u);
    HIP_CHECK(e);
    HIP_CHECK(e2);
    API_END
}

int asat_rk_create(asat_rk_t** out)
{
    API_BEGIN
    CHECK_PTR(out);
    *out = new asat_rk();
    API_END
}

int asat_rk_destroy(asat_rk_t* rk)
{
    API_BEGIN
    delete rk;
    API_END
}

int asat_rk_set_eps(asat_rk_t* rk, double eps)
{
    API_BEGIN
    CHECK_PTR(rk);
    // CpuRungeKutta::SetEpsilon (AnalogSAT/cpu_sat/CpuRungeKutta.cpp:126-131)
    require(!(eps < 1e-16 || eps > 1.0), ASAT_ERR_INVALID_ARGUMENT, "epsilon must be between 1e-16 and 1");
    rk->r.eps = eps;
    if (rk->r.configured) {
        hipLaunchKernelGGL(k_ctrl_set_eps, dim3(1), dim3(64), 0, rk->r.sys->stream, rk->r.d_ctrl, eps, rk->r.floor);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipStreamSynchronize(rk->r.sys->stream));
    }
    API_END
}

int asat_rk_set_xscal_floor(asat_rk_t* rk, double floor)
{
    API_BEGIN
    CHECK_PTR(rk);
    require(floor > 0.0 && floor <= 1.0, ASAT_ERR_INVALID_ARGUMENT, "xscal floor must be in (0, 1]");
    rk->r.floor = floor;
    if (rk->r.configured) {
        hipLaunchKernelGGL(k_ctrl_set_eps, dim3(1), dim3(64), 0, rk->r.sys->stream, rk->r.d_ctrl, rk->r.eps, floor);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipStreamSynchronize(rk->r.sys->stream));
    }
    API_END
}

__global__ void k_ctrl_set_log(Ctrl* c, double* hlog, long long len)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        c->hlog = hlog;
        c->hlog_len = len;
    }
}

int asat_rk_set_step_log(asat_rk_t* rk, double* dev_hlog, long long len)
{
    API_BEGIN
    CHECK_PTR(rk);
    require(len >= 0 && (len == 0 || dev_hlog != nullptr), ASAT_ERR_INVALID_ARGUMENT, "bad step log");
    rk->r.hlog = len ? dev_hlog : nullptr;
    rk->r.hlog_len = len;
    if (rk->r.configured) {
        hipLaunchKernelGGL(k_ctrl_set_log, dim3(1), dim3(64), 0, rk->r.sys->stream, rk->r.d_ctrl, rk->r.hlog,
                           rk->r.hlog_len);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipStreamSynchronize(rk->r.sys->stream));
    }
    API_END
}

int asat_rk_get_eps(const asat_rk_t* rk, double* eps)
{
    API_BEGIN
    CHECK_PTR(rk); CHECK_PTR(eps);
    *eps = rk->r.eps;
    API_END
}

int asat_rk_configure(asat_rk_t* rk, asat_sys_t* sys)
{
    API_BEGIN
    CHECK_PTR(rk); CHECK_PTR(sys);
    rk_configure(&rk->r, &sys->s);
    API_END
}

int asat_rk_step(asat_rk_t* rk, asat_state_t* old_state, asat_state_t* new_state, int count)
{
    API_BEGIN
    CHECK_PTR(rk); CHECK_PTR(old_state); CHECK_PTR(new_state);
    int rc = rk_step(&rk->r, &old_state->st, &new_state->st, count);
    if (rc == ASAT_UNDERFLOW) {
        asat::g_last_error = "step size underflow";
        return rc;
    }
    API_END
}

int asat_rk_set_one_workgroup(asat_rk_t* rk, int enable)
{
    API_BEGIN
    CHECK_PTR(rk);
    rk->r.small_mode = enable ? 1 : 0;
    API_END
}

int asat_rk_uses_one_workgroup(const asat_rk_t* rk, int* uses)
{
    API_BEGIN
    CHECK_PTR(rk); CHECK_PTR(uses);
    rk->r.require_configured();
    *uses = small_path(&rk->r) ? 1 : 0;
    API_END
}

int asat_rk_solve(asat_rk_t* rk, asat_state_t* old_state, asat_state_t* new_state, int batch, long long max_steps,
                  double max_time, int time_float, asat_solve_result_t* out)
{
    API_BEGIN
    CHECK_PTR(rk); CHECK_PTR(old_state); CHECK_PTR(new_state); CHECK_PTR(out);
    rk_solve(&rk->r, &old_state->st, &new_state->st, batch, max_steps, max_time, time_float, out);
    API_END
}

int asat_rk_reset(asat_rk_t* rk)
{
    API_BEGIN
    CHECK_PTR(rk);
    // no copy into the host control block may be pending when it is rewritten
    if (rk->r.configured) HIP_CHECK(hipStreamSynchronize(rk->r.sys->stream));
    rk_reset(&rk->r);
    API_END
}

int asat_rk_stats(const asat_rk_t* rk, asat_rk_stats_t* out)
{
    API_BEGIN
    CHECK_PTR(rk); CHECK_PTR(out);
    rk->r.require_configured();
    const Ctrl* c = rk->r.h_ctrl;
    out->time = c->time;
    out->hdid = c->hdid;
    out->hnext = c->hnext;
    out->errmax = c->errmax;
    out->steps = c->steps;
    out->attempts = c->attempts;
    out->rhscount = c->rhscount;
    API_END
}

int asat_rk_last_step_ms(const asat_rk_t* rk, float* ms)
{
    API_BEGIN
    CHECK_PTR(rk); CHECK_PTR(ms);
    *ms = rk->r.last_ms;
    API_END
}

int asat_rk_set_kernel_timer(asat_rk_t* rk, int kind)
{
    API_BEGIN
    CHECK_PTR(rk);
    require(kind >= -1 && kind <= 12, ASAT_ERR_INVALID_ARGUMENT, "kernel kind must be -1 (off) or 0..12");
    rk->r.timer_kind = kind;
    rk->r.timer_ms = 0.0;
    rk->r.timer_launches = 0;
    API_END
}

int asat_rk_kernel_timer(const asat_rk_t* rk, double* total_ms, long long* launches)
{
    API_BEGIN
    CHECK_PTR(rk); CHECK_PTR(total_ms); CHECK_PTR(launches);
    *total_ms = rk->r.timer_ms;
    *launches = rk->r.timer_launches;
    API_END
}

int asat_rk_profile(asat_rk_t* rk, asat_state_t* old_state, asat_state_t* new_state, int reps, float* us_out)
{
    API_BEGIN
    CHECK_PTR(rk); CHECK_PTR(old_state); CHECK_PTR(new_state); CHECK_PTR(us_out);
    RK* r = &rk->r;
    r->require_configured();
    state_check(&old_state->st); state_check(&new_state->st);
    require(reps > 0, ASAT_ERR_INVALID_ARGUMENT, "reps must be positive");
    require(old_state->st.d != new_state->st.d, ASAT_ERR_INVALID_ARGUMENT, "old and new state must differ");
    if (r->prec == ASAT_F64) rk_profile_impl<double>(r, &old_state->st, &new_state->st, reps, us_out);
    else rk_profile_impl<float>(r, &old_state->st, &new_state->st, reps, us_out);
    API_END
}

int asat_util_copy_bench(long long bytes, int reps, float* us_per_copy)
{
    API_BEGIN
    CHECK_PTR(us_per_copy);
    require(bytes >= 16 && bytes % 16 == 0 && reps > 0, ASAT_ERR_INVALID_ARGUMENT,
            "bytes must be a positive multiple of 16 and reps positive");
    void *a = nullptr, *b = nullptr;
    HIP_CHECK(hipMalloc(&a, (size_t)bytes));
    HIP_CHECK(hipMalloc(&b, (size_t)bytes));
    HIP_CHECK(hipMemset(a, 1, (size_t)bytes));
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    const long long n16 = bytes / 16;
    hipLaunchKernelGGL(k_copy16, dim3(4096), dim3(256), 0, 0, (const int4*)a, (int4*)b, n16);
    HIP_CHECK(hipEventRecord(e0, 0));
    for (int r = 0; r < reps; r++)
        hipLaunchKernelGGL(k_copy16, dim3(4096), dim3(256), 0, 0, (const int4*)a, (int4*)b, n16);
    HIP_CHECK(hipEventRecord(e1, 0));
    HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    *us_per_copy = ms * 1000.f / reps;
    (void)hipEventDestroy(e0); (void)hipEventDestroy(e1);
    (void)hipFree(a); (void)hipFree(b);
    API_END
}

int asat_util_stream_bench(long long bytes, int reps, float* us_copy, float* us_read)
{
    API_BEGIN
    CHECK_PTR(us_copy); CHECK_PTR(us_read);
    require(bytes >= 16 && bytes % 16 == 0 && reps > 0, ASAT_ERR_INVALID_ARGUMENT,
            "bytes must be a positive multiple of 16 and reps positive");
    void *a = nullptr, *b = nullptr;
    int* o = nullptr;
    HIP_CHECK(hipMalloc(&a, (size_t)bytes));
    HIP_CHECK(hipMalloc(&b, (size_t)bytes));
    HIP_CHECK(hipMalloc((void**)&o, sizeof(int)));
    HIP_CHECK(hipMemset(a, 0, (size_t)bytes));
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    const long long n16 = bytes / 16;
    const dim3 grid(8192), blk(256);
    float ms = 0.f;
    hipLaunchKernelGGL(k_stream_copy, grid, blk, 0, 0, (const stream_v4*)a, (stream_v4*)b, n16);
    HIP_CHECK(hipEventRecord(e0, 0));
    for (int r = 0; r < reps; r++) hipLaunchKernelGGL(k_stream_copy, grid, blk, 0, 0, (const stream_v4*)a, (stream_v4*)b, n16);
    HIP_CHECK(hipEventRecord(e1, 0));
    HIP_CHECK(hipEventSynchronize(e1));
    HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    *us_copy = ms * 1000.f / reps;
    hipLaunchKernelGGL(k_stream_read, grid, blk, 0, 0, (const stream_v4*)a, o, n16);
    HIP_CHECK(hipEventRecord(e0, 0));
    for (int r = 0; r < reps; r++) hipLaunchKernelGGL(k_stream_read, grid, blk, 0, 0, (const stream_v4*)a, o, n16);
    HIP_CHECK(hipEventRecord(e1, 0));
    HIP_CHECK(hipEventSynchronize(e1));
    HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    *us_read = ms * 1000.f / reps;
    (void)hipEventDestroy(e0); (void)hipEventDestroy(e1);
    (void)hipFree(a); (void)hipFree(b); (void)hipFree(o);
    API_END
}

int asat_util_gpu_name(char* buf, int len)
{
    API_BEGIN
    CHECK_PTR(buf);
    require(len > 0, ASAT_ERR_INVALID_ARGUMENT, "len must be positive");
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIP_CHECK(hipGetDeviceProperties(&prop, dev));
    if (prop.name[0] != '\0') {
        std::snprintf(buf, (size_t)len, "%s", prop.name);
    } else {
        // some drivers leave the marketing name empty: the architecture then
        // names the device (the part of gcnArchName before its feature flags)
        std::string arch(prop.gcnArchName);
        arch = arch.substr(0, arch.find(':'));
        std::snprintf(buf, (size_t)len, "AMD GPU %s", arch.empty() ? "(unknown)" : arch.c_str());
    }
    API_END
}

int asat_rng_create(unsigned long long seed, asat_rng_t** out)
{
    API_BEGIN
    CHECK_PTR(out);
    asat_rng* r = new asat_rng();
    r->g.seed = seed;
    r->g.offset = 0;
    *out = r;
    API_END
}

int asat_rng_destroy(asat_rng_t* rng)
{
    API_BEGIN
    delete rng;
    API_END
}

int asat_rng_uniform(asat_rng_t* rng, void* dev_ptr, int len, int precision)
{
    API_BEGIN
    CHECK_PTR(rng); CHECK_PTR(dev_ptr);
    require(len >= 0, ASAT_ERR_INVALID_ARGUMENT, "negative length");
    require(precision == ASAT_F64 || precision == ASAT_F32, ASAT_ERR_INVALID_ARGUMENT, "unknown precision");
    if (len == 0) return ASAT_OK;
    dim3 g(ceil_div(len, 256)), b(256);
    if (precision == ASAT_F64)
        hipLaunchKernelGGL(k_uniform<double>, g, b, 0, 0, (double*)dev_ptr, len, rng->g.seed, rng->g.offset);
    else
        hipLaunchKernelGGL(k_uniform<float>, g, b, 0, 0, (float*)dev_ptr, len, rng->g.seed, rng->g.offset);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipDeviceSynchronize());
    rng->g.offset += (unsigned long long)len;
    API_END
}

}  // extern "C"
