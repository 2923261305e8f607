/*
 * Type-generic body of the CPU oracle (test infrastructure only).
 *
 * Included by asat_oracle.c once per precision mode with these macros set:
 *   ST   state / right-hand-side arithmetic type (double, float or long double)
 *   CT   step-control type: time, h, error norm, accept/reject (double, or
 *        long double with ST = long double)
 *   SFX  symbol suffix (_f64, _mix for ST=float, _ld for long double)
 *   ST_SIN, ST_TANH, CT_FABS, CT_FMAX, CT_FMIN: the math of those types;
 *   CT_POW: the controller's pow (glibc or correctly rounded in double,
 *   powl in long double as CpuRungeKutta<long double> calls it)
 *
 * Every routine restates a routine of the reference CPU path and follows its
 * floating-point operation order, so that with -ffp-contract=off the FP64 mode
 * reproduces the reference's results bit for bit:
 *   rhs        AnalogSAT/cpu_sat/CpuSat.cpp:64-94 (GetDerivatives) and
 *              :197-241 (CalculateAllEx<K>), products as in
 *              AnalogSAT/cpu_sat/cpu_helper.h:23-113 (CpuProdWrapper/CpuKmiWrapper)
 *   violations AnalogSAT/cpu_sat/CpuSat.cpp:135-174
 *   rk step    AnalogSAT/cpu_sat/CpuRungeKutta.cpp:65-100 (Step),
 *              :180-215 (rkck), :217-267 (rkqs), :168-177 (Reset);
 *              tableau AnalogSAT/cpu_sat/CpuRungeKutta.h:72-112
 *   solve      AnalogSAT/solver/SatSolver.h:273-355 (DoSolve, no callback)
 *
 * In the mixed mode (ST=float, CT=double) the state arithmetic stays in float
 * exactly as CpuSat<float> does it, while xscal, the error estimate, the max
 * norm and the step controller run in double (h enters the float stage sums
 * as (float)h).  This mode has no counterpart in the reference; it is the
 * definition the GPU FP32 path is checked against.
 */

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SFX)

/* one Cash-Karp integrator bound to a problem; state layout [-1, s_1..s_N, a_1..a_M] */
typedef struct {
    int n, m, k;
    const int *C;        /* clause matrix, clause-major: C[j*k + q], 0 = padding */
    ST knorm;
    ST b;                /* bias prefactor; 0 disables the sine term */
    int family;          /* 0: CpuSat, 1: CpuSatTanh */
    ST q;                /* CpuSatTanh's q (default 1.1) */
    int len;             /* n + m + 1 */
    ST *x, *xnew, *temp, *dxdt, *ak2, *ak3, *ak4, *ak5, *ak6;
    CT *xscal, *xerr;
    CT time, hdid, hnext, eps;
    CT floor;            /* constant term of xscal: TINY = 1e-20 in the reference */
    long steps, attempts, rhscount;
    CT last_errmax;
} FN(orc_rk_t);

/* ---- right-hand side (CpuSat::GetDerivatives / CalculateAllEx<K>) ---- */

/* product of s[0..kk-1] skipping position skip, left to right
   (CpuProdWrapper<TFloat,K,Q>::KmiProd) */
static inline ST FN(kmi_prod)(const ST *s, int kk, int skip)
{
    ST p = 0;
    int first = 1;
    for (int q = 0; q < kk; q++) {
        if (q == skip) continue;
        if (first) { p = s[q]; first = 0; }
        else p = p * s[q];
    }
    if (first) p = (ST)1;   /* K == 1 */
    return p;
}

void FN(orc_rhs)(int n, int m, int k, const int *C, ST b, const ST *state, ST *rhs)
{
    const int D = n + m + 1;
    const ST one = (ST)1.0, zero = (ST)0.0;
    const ST knorm = (ST)pow(2.0, -k);   /* CpuSat.cpp:355 */
    const ST bb = b * knorm * knorm;      /* CpuSat.cpp:76 */
    const ST alpha = (ST)m / (ST)n;       /* CpuSat.cpp:55 */
    ST s[64];
    int index[64];
    ST sum = 0;
    double sumd = 0;

    memset(rhs, 0, sizeof(ST) * (size_t)D);             /* dxdt.SetZero() */
    for (int j = 0; j < m; j++) {
        ST km = knorm;
        for (int q = 0; q < k; q++) index[q] = C[(size_t)j * k + q];
        for (int q = 0; q < k; q++) {
            s[q] = index[q] != 0 ? state[abs(index[q])] : (ST)-1.0;
            s[q] = index[q] < 0 ? one + s[q] : one - s[q];
            km *= s[q];
        }
        int p = n + 1 + j;
        ST am = state[p];
        rhs[p] = am * km * km;
        sum += am;
        sumd += (double)am;
        if (k <= 10) {
            ST r = km * knorm * am;
            /* CpuKmiWrapper::KmiFunc: stops at the first zero (padding) index */
            for (int q = 0; q < k; q++) {
                if (index[q] == 0) break;
                ST rr = r * FN(kmi_prod)(s, k, q);
                rr *= index[q] < 0 ? (ST)-2.0 : (ST)2.0;
                rhs[abs(index[q])] += rr;
            }
        } else {
            /* CpuSat::CalculateAll (CpuSat.cpp:290-349), used for K > 10:
               temp = km * prod_{q' != q} f_q', then ((2*temp)*knorm)*am*sign;
               padding slots add into rhs[0], which is cleared afterwards */
            for (int q = 0; q < k; q++) {
                ST temp = km;
                for (int qq = 0; qq < k; qq++) temp *= (qq == q) ? one : s[qq];
                ST sign = index[q] < 0 ? (ST)-1.0 : one;
                rhs[abs(index[q])] += (ST)2.0 * temp * knorm * am * sign;
            }
        }
    }
    if (bb > zero) {
        ST a = (g_mean_mode ? (ST)sumd : sum) / (ST)m;
        for (int i = 1; i < n + 1; i++)
            rhs[i] += a * alpha * bb * (ST)1.570796326794897 * ST_SIN(state[i] * (ST)3.141592653589793);
    }
    rhs[0] = zero;
}

/* ---- tanh family right-hand side (CpuSatTanh::GetDerivatives) ----
   AnalogSAT/cpu_sat/CpuSatTanh.cpp:64-94, CalculateAllEx<K> :192-232 (K = 3..10),
   CalculateAll :235-286 (other K), KmiFuncTanh AnalogSAT/cpu_sat/cpu_helper.h:115-125.
   Per slot: knorm * a * |prod of the other factors| * tanh(q * f) * c; the aux
   derivative is a K^2 as in CpuSat.  The two reference branches multiply the
   sign (+-1) in at different places, which gives identical results; they
   differ in the tanh they call (see below).  The bias term is added
   unconditionally as there. */
void FN(orc_rhs_tanh)(int n, int m, int k, const int *C, ST b, ST qq, const ST *state, ST *rhs)
{
    const int D = n + m + 1;
    const ST one = (ST)1.0, zero = (ST)0.0;
    const ST knorm = (ST)pow(2.0, -k);   /* CpuSatTanh.cpp:318 */
    const ST bb = b * knorm * knorm;      /* CpuSatTanh.cpp:77 */
    const ST alpha = (ST)m / (ST)n;       /* CpuSatTanh.cpp:56 */
    ST s[64];
    int index[64];
    ST sum = 0;
    double sumd = 0;

    memset(rhs, 0, sizeof(ST) * (size_t)D);
    for (int j = 0; j < m; j++) {
        ST km = knorm;
        for (int q = 0; q < k; q++) index[q] = C[(size_t)j * k + q];
        for (int q = 0; q < k; q++) {
            s[q] = index[q] != 0 ? state[abs(index[q])] : (ST)-1.0;
            s[q] = index[q] < 0 ? one + s[q] : one - s[q];
            km *= s[q];
        }
        int p = n + 1 + j;
        ST am = state[p];
        rhs[p] = am * km * km;
        sum += am;
        sumd += (double)am;
        ST r = knorm * am;
        for (int q = 0; q < k; q++) {
            if (index[q] == 0) break;   /* padding: the generic branch adds into rhs[0] only */
            ST kmi = FN(kmi_prod)(s, k, q);
            ST rr;
            if (k >= 3 && k <= 10)
                /* KmiFuncTanh calls the C library's double tanh (cpu_helper.h is
                   included before CpuSatTanh.cpp's using-directive); the usual
                   conversions then form the product in double (float state,
                   rounded once) or in long double */
                rr = (ST)((r * (kmi < zero ? -kmi : kmi)) * tanh((double)(qq * s[q])));
            else
                rr = r * (kmi < zero ? -kmi : kmi) * ST_TANH(qq * s[q]);
            rr *= index[q] < 0 ? (ST)-1.0 : one;
            rhs[abs(index[q])] += rr;
        }
    }
    {
        ST a = (g_mean_mode ? (ST)sumd : sum) / (ST)m;
        for (int i = 1; i < n + 1; i++)
            rhs[i] += a * alpha * bb * (ST)1.570796326794897 * ST_SIN(state[i] * (ST)3.141592653589793);
    }
    rhs[0] = zero;
}

/* ---- violated-clause count (CpuSat::CalculateClauses) ---- */
int FN(orc_violations)(int n, int m, int k, const int *C, const ST *state)
{
    (void)n;
    int count = 0;
    for (int j = 0; j < m; j++) {
        int clause = 0;
        for (int i = 0; i < k; i++) {
            int item = C[(size_t)j * k + i];
            int index = item < 0 ? -item : item;
            int lit = state[index] > (ST)0;
            if (item < 0) lit = !lit;
            clause |= lit;
            if (clause) break;
        }
        count += clause ? 0 : 1;
    }
    return count;
}

/* ---- Cash-Karp RK45 (CpuRungeKutta) ---- */

FN(orc_rk_t) *FN(orc_rk_create)(int n, int m, int k, const int *C, CT eps)
{
    FN(orc_rk_t) *r = (FN(orc_rk_t) *)calloc(1, sizeof(*r));
    r->n = n; r->m = m; r->k = k; r->C = C;
    r->knorm = (ST)pow(2.0, -k);
    r->len = n + m + 1;
    size_t L = (size_t)r->len;
    r->x = (ST *)calloc(L, sizeof(ST));    r->xnew = (ST *)calloc(L, sizeof(ST));
    r->temp = (ST *)calloc(L, sizeof(ST)); r->dxdt = (ST *)calloc(L, sizeof(ST));
    r->ak2 = (ST *)calloc(L, sizeof(ST));  r->ak3 = (ST *)calloc(L, sizeof(ST));
    r->ak4 = (ST *)calloc(L, sizeof(ST));  r->ak5 = (ST *)calloc(L, sizeof(ST));
    r->ak6 = (ST *)calloc(L, sizeof(ST));
    r->xscal = (CT *)calloc(L, sizeof(CT)); r->xerr = (CT *)calloc(L, sizeof(CT));
    r->x[0] = (ST)-1.0; r->xnew[0] = (ST)-1.0;
    r->eps = eps;
    r->floor = (CT)1e-20;
    r->b = 0;
    r->family = 0;
    r->q = (ST)1.1;       /* CpuSatTanh.cpp:39 */
    /* Reset(): CpuRungeKutta.cpp:168-177 */
    r->hnext = (CT)1e-9; r->steps = 0; r->attempts = 0; r->time = 0; r->rhscount = 0;
    r->hdid = 0;
    return r;
}

void FN(orc_rk_destroy)(FN(orc_rk_t) *r)
{
    if (!r) return;
    free(r->x); free(r->xnew); free(r->temp); free(r->dxdt);
    free(r->ak2); free(r->ak3); free(r->ak4); free(r->ak5); free(r->ak6);
    free(r->xscal); free(r->xerr);
    free(r);
}

void FN(orc_rk_reset)(FN(orc_rk_t) *r)
{
    r->hnext = (CT)1e-9; r->steps = 0; r->attempts = 0; r->time = 0; r->rhscount = 0;
}

void FN(orc_rk_set_b)(FN(orc_rk_t) *r, double b) { r->b = (ST)b; }

/* right-hand side family: 0 = CpuSat, 1 = CpuSatTanh with parameter q */
void FN(orc_rk_set_family)(FN(orc_rk_t) *r, int family, double q)
{
    r->family = family;
    r->q = (ST)q;
}

/* xscal = |x| + |h k1| + floor; floor = TINY (1e-20) reproduces the reference */
void FN(orc_rk_set_floor)(FN(orc_rk_t) *r, double floor) { r->floor = (CT)floor; }

/* public [s, a] -> internal [-1, s, a] (CpuSatStateImpl::CopyPartsFrom) */
void FN(orc_rk_set_state)(FN(orc_rk_t) *r, const ST *pub)
{
    r->x[0] = (ST)-1.0;
    memcpy(r->x + 1, pub, sizeof(ST) * (size_t)(r->n + r->m));
}

void FN(orc_rk_get_state)(const FN(orc_rk_t) *r, ST *pub)
{
    memcpy(pub, r->x + 1, sizeof(ST) * (size_t)(r->n + r->m));
}

static void FN(rhs_call)(FN(orc_rk_t) *r, ST *out, const ST *in)
{
    if (r->family == 1) FN(orc_rhs_tanh)(r->n, r->m, r->k, r->C, r->b, r->q, in, out);
    else FN(orc_rhs)(r->n, r->m, r->k, r->C, r->b, in, out);
    r->rhscount++;
}

/* rkck: CpuRungeKutta.cpp:180-215.  h is the control-type step; the state sums
   use hs = (ST)h, which is h itself in the FP64 mode. */
static void FN(rkck)(FN(orc_rk_t) *r, CT h)
{
    const int n = r->len;
    const ST hs = (ST)h;
    const ST b21 = (ST)0.2, b31 = (ST)(3.0 / 40.0), b32 = (ST)(9.0 / 40.0),
             b41 = (ST)0.3, b42 = (ST)-0.9, b43 = (ST)1.2,
             b51 = (ST)(-11.0 / 54.0), b52 = (ST)2.5, b53 = (ST)(-70.0 / 27.0), b54 = (ST)(35.0 / 27.0),
             b61 = (ST)(1631.0 / 55296.0), b62 = (ST)(175.0 / 512.0), b63 = (ST)(575.0 / 13824.0),
             b64 = (ST)(44275.0 / 110592.0), b65 = (ST)(253.0 / 4096.0),
             c1 = (ST)(37.0 / 378.0), c3 = (ST)(250.0 / 621.0), c4 = (ST)(125.0 / 594.0),
             c6 = (ST)(512.0 / 1771.0);
    /* the error coefficients live in the control type (CpuRungeKutta.h:107-112) */
    const CT C1 = (CT)(37.0 / 378.0), C3 = (CT)(250.0 / 621.0), C4 = (CT)(125.0 / 594.0),
             C6 = (CT)(512.0 / 1771.0);
    const CT dc5 = (CT)(-277.0 / 14336.0);
    const CT dc1 = C1 - (CT)(2825.0 / 27648.0), dc3 = C3 - (CT)(18575.0 / 48384.0),
             dc4 = C4 - (CT)(13525.0 / 55296.0), dc6 = C6 - (CT)0.25;
    const ST *x = r->x;
    ST *t = r->temp, *k1 = r->dxdt, *k2 = r->ak2, *k3 = r->ak3, *k4 = r->ak4, *k5 = r->ak5, *k6 = r->ak6;

    for (int i = 0; i < n; i++) t[i] = x[i] + b21 * hs * k1[i];
    FN(rhs_call)(r, k2, t);
    for (int i = 0; i < n; i++) t[i] = x[i] + hs * (b31 * k1[i] + b32 * k2[i]);
    FN(rhs_call)(r, k3, t);
    for (int i = 0; i < n; i++) t[i] = x[i] + hs * (b41 * k1[i] + b42 * k2[i] + b43 * k3[i]);
    FN(rhs_call)(r, k4, t);
    for (int i = 0; i < n; i++) t[i] = x[i] + hs * (b51 * k1[i] + b52 * k2[i] + b53 * k3[i] + b54 * k4[i]);
    FN(rhs_call)(r, k5, t);
    for (int i = 0; i < n; i++) t[i] = x[i] + hs * (b61 * k1[i] + b62 * k2[i] + b63 * k3[i] + b64 * k4[i] + b65 * k5[i]);
    FN(rhs_call)(r, k6, t);
    for (int i = 0; i < n; i++) r->xnew[i] = x[i] + hs * (c1 * k1[i] + c3 * k3[i] + c4 * k4[i] + c6 * k6[i]);
    for (int i = 0; i < n; i++)
        r->xerr[i] = h * (dc1 * (CT)k1[i] + dc3 * (CT)k3[i] + dc4 * (CT)k4[i] + dc5 * (CT)k5[i] + dc6 * (CT)k6[i]);
}

/* rkqs: CpuRungeKutta.cpp:217-267.  Returns 1 on success, 0 on step underflow. */
static int FN(rkqs)(FN(orc_rk_t) *r)
{
    const CT SAFETY = (CT)0.9, PGROW = (CT)-0.2, PSHRNK = (CT)-0.25, ERRCON = (CT)1.89e-4,
             HFACTOR = (CT)0.1;
    CT h = r->hnext;
    CT ermax = 0;
    const int n = r->len;
    for (;;) {
        FN(rkck)(r, h);
        r->attempts++;
        for (int i = 0; i < n; i++) r->xerr[i] = CT_FABS(r->xerr[i] / r->xscal[i]);
        ermax = 0;
        for (int i = 0; i < n; i++) ermax = CT_FMAX(ermax, r->xerr[i]);
        ermax /= r->eps;
        r->last_errmax = ermax;
        if (ermax <= (CT)1.0) break;
        CT htemp = SAFETY * h * CT_POW(ermax, PSHRNK);
        h = h >= (CT)0 ? CT_FMAX(htemp, HFACTOR * h) : CT_FMIN(htemp, HFACTOR * h);
        CT tnew = r->time + h;
        if (tnew == r->time) return 0;
    }
    if (ermax > ERRCON) r->hnext = SAFETY * h * CT_POW(ermax, PGROW);
    else r->hnext = (CT)5.0 * h;
    r->hdid = h;
    r->time += r->hdid;
    return 1;
}

/* Step(count): CpuRungeKutta.cpp:65-100.  Returns 1, or 0 on underflow. */
int FN(orc_rk_step)(FN(orc_rk_t) *r, int count)
{
    const CT TINY = r->floor;
    const int n = r->len;
    for (int c = 0; c < count; c++) {
        FN(rhs_call)(r, r->dxdt, r->x);
        for (int i = 0; i < n; i++)
            r->xscal[i] = CT_FABS((CT)r->x[i]) + CT_FABS((CT)r->dxdt[i] * r->hnext) + TINY;
        if (FN(rkqs)(r)) {
            r->steps++;
            ST *tmp = r->x; r->x = r->xnew; r->xnew = tmp;
        } else return 0;
    }
    return 1;
}

int FN(orc_rk_violations)(const FN(orc_rk_t) *r)
{
    return FN(orc_violations)(r->n, r->m, r->k, r->C, r->x);
}

/* SatSolver::DoSolve without a callback (SatSolver.h:317-339).
   result codes follow AnalogSAT/solver/SatResult.h:24-32:
   0 solution, 1 underflow, 2 interrupted, 3 max time, 4 max steps, 5 unknown.
   out_batches (optional, size >= max_batches): violation count after each batch. */
int FN(orc_solve)(FN(orc_rk_t) *r, CT max_time, int max_steps, int batch,
                  int *out_batches, int max_batches, int *n_batches)
{
    int result = 5;
    long steps = 0;
    CT time = 0;
    int nb = 0;
    FN(orc_rk_reset)(r);
    while (time < max_time && steps < max_steps) {
        long before = r->steps;
        if (!FN(orc_rk_step)(r, batch)) { result = 1; break; }
        steps += r->steps - before;
        time = r->time;
        int viol = FN(orc_rk_violations)(r);
        if (out_batches && nb < max_batches) out_batches[nb] = viol;
        nb++;
        if (viol == 0) { result = 0; break; }
    }
    if (result == 5) {
        if (time >= max_time) result = 3;
        else if (steps >= max_steps) result = 4;
    }
    if (n_batches) *n_batches = nb;
    return result;
}

/* per-element |xerr/xscal| of the last attempt (length n+m+1, internal layout) */
void FN(orc_rk_err_vector)(const FN(orc_rk_t) *r, double *out)
{
    for (int i = 0; i < r->len; i++) out[i] = (double)r->xerr[i];
}

/* accessors for the Python side */
double FN(orc_rk_time)(const FN(orc_rk_t) *r) { return (double)r->time; }
double FN(orc_rk_hdid)(const FN(orc_rk_t) *r) { return (double)r->hdid; }
double FN(orc_rk_hnext)(const FN(orc_rk_t) *r) { return (double)r->hnext; }
double FN(orc_rk_errmax)(const FN(orc_rk_t) *r) { return (double)r->last_errmax; }
long FN(orc_rk_steps)(const FN(orc_rk_t) *r) { return r->steps; }
long FN(orc_rk_attempts)(const FN(orc_rk_t) *r) { return r->attempts; }
long FN(orc_rk_rhscount)(const FN(orc_rk_t) *r) { return r->rhscount; }


/* ===========================================================================
 * Adjoint (no reference counterpart; the derivation is this build's):
 * gradient of V(T) = sum_m a_m K_m^2 at the state after T accepted steps,
 * with respect to the start state, through the realised Cash-Karp steps
 * (step sizes frozen).  Per step x' = x + h sum_l c_l k_l, k_l = f(Y_l),
 * Y_l = x + h sum_{j<l} b_lj k_j.  Reverse: mu_l = h c_l lam' +
 * sum_{j>l} h b_jl nu_j, nu_l = J(Y_l)^T mu_l, lam = lam' + sum_l nu_l.
 * VJP of one clause (f_q = 1 - c_q s_q, K = 2^-k prod f, K_q / K_qp the
 * products without q / without q,p; padding factors 2 are constants):
 *   u_a    = w_a K^2 + sum_q w_q 2 c_q K K_q
 *   u_s[p] = w_a (-2 a c_p K K_p) + w_p (-2 a K_p^2) + sum_{q!=p} w_q (-4 a c_q c_p K K_qp)
 * All adjoint arithmetic is double; the state type ST only sets how the
 * forward states (and so the stage inputs) are rounded.
 * =========================================================================*/

/* product over q' of fd[q'] skipping up to two positions, left to right, times kn */
static double FN(adj_prod)(const double *fd, int k, int skip1, int skip2, double kn)
{
    double p = kn;
    for (int q = 0; q < k; q++) {
        if (q == skip1 || q == skip2) continue;
        p = p * fd[q];
    }
    return p;
}

/* u = J(Y)^T w on the internal layout [0, s_1..s_N, a_1..a_M] (index 0 unused) */
void FN(orc_vjp)(int n, int m, int k, const int *C, const ST *Y, const double *w, double *u)
{
    const int D = n + m + 1;
    const double kn = pow(2.0, -k);
    double fd[64], cs[64];
    int v[64];
    for (int i = 0; i < D; i++) u[i] = 0.0;
    for (int j = 0; j < m; j++) {
        for (int q = 0; q < k; q++) {
            int l = C[(size_t)j * k + q];
            v[q] = l < 0 ? -l : l;
            cs[q] = l < 0 ? -1.0 : 1.0;
            ST s = l != 0 ? Y[v[q]] : (ST)-1.0;
            ST f = l < 0 ? (ST)1.0 + s : (ST)1.0 - s;
            fd[q] = (double)f;
        }
        const double K = FN(adj_prod)(fd, k, -1, -1, kn);
        const double a = (double)Y[n + 1 + j];
        const double wa = w[n + 1 + j];
        double ua = wa * K * K;
        for (int q = 0; q < k; q++) {
            if (v[q] == 0) continue;
            ua += w[v[q]] * (2.0 * cs[q] * K * FN(adj_prod)(fd, k, q, -1, kn));
        }
        u[n + 1 + j] = ua;
        for (int p = 0; p < k; p++) {
            if (v[p] == 0) continue;
            const double Kp = FN(adj_prod)(fd, k, p, -1, kn);
            double t = wa * (-2.0 * a * cs[p] * K * Kp) + w[v[p]] * (-2.0 * a * Kp * Kp);
            for (int q = 0; q < k; q++) {
                if (q == p || v[q] == 0) continue;
                t += w[v[q]] * (-4.0 * a * cs[q] * cs[p] * K * FN(adj_prod)(fd, k, q, p, kn));
            }
            u[v[p]] += t;
        }
    }
}

/* grad V at x: dV/da_m = K_m^2, dV/ds_i = sum_m -2 a_m c_mi K_m K_mi (double) */
double FN(orc_energy_grad)(int n, int m, int k, const int *C, const ST *x, double *g)
{
    const int D = n + m + 1;
    const double kn = pow(2.0, -k);
    double fd[64], cs[64];
    int v[64];
    double V = 0.0;
    for (int i = 0; i < D; i++) g[i] = 0.0;
    for (int j = 0; j < m; j++) {
        for (int q = 0; q < k; q++) {
            int l = C[(size_t)j * k + q];
            v[q] = l < 0 ? -l : l;
            cs[q] = l < 0 ? -1.0 : 1.0;
            ST s = l != 0 ? x[v[q]] : (ST)-1.0;
            ST f = l < 0 ? (ST)1.0 + s : (ST)1.0 - s;
            fd[q] = (double)f;
        }
        const double K = FN(adj_prod)(fd, k, -1, -1, kn);
        const double a = (double)x[n + 1 + j];
        V += a * K * K;
        g[n + 1 + j] = K * K;
        for (int p = 0; p < k; p++) {
            if (v[p] == 0) continue;
            g[v[p]] += -2.0 * a * cs[p] * K * FN(adj_prod)(fd, k, p, -1, kn);
        }
    }
    return V;
}

/* one Cash-Karp step with a given h (no control), the realised map; also
   returns the stage inputs Y[0..5] (each length len) when Ys != NULL */
static void FN(frozen_step)(FN(orc_rk_t) *r, const ST *x, ST *xout, double h, ST *Ys)
{
    const int n = r->len;
    const ST hs = (ST)h;
    const ST b21 = (ST)0.2, b31 = (ST)(3.0 / 40.0), b32 = (ST)(9.0 / 40.0),
             b41 = (ST)0.3, b42 = (ST)-0.9, b43 = (ST)1.2,
             b51 = (ST)(-11.0 / 54.0), b52 = (ST)2.5, b53 = (ST)(-70.0 / 27.0), b54 = (ST)(35.0 / 27.0),
             b61 = (ST)(1631.0 / 55296.0), b62 = (ST)(175.0 / 512.0), b63 = (ST)(575.0 / 13824.0),
             b64 = (ST)(44275.0 / 110592.0), b65 = (ST)(253.0 / 4096.0),
             c1 = (ST)(37.0 / 378.0), c3 = (ST)(250.0 / 621.0), c4 = (ST)(125.0 / 594.0),
             c6 = (ST)(512.0 / 1771.0);
    ST *k1 = r->dxdt, *k2 = r->ak2, *k3 = r->ak3, *k4 = r->ak4, *k5 = r->ak5, *k6 = r->ak6, *t = r->temp;
    FN(orc_rhs)(r->n, r->m, r->k, r->C, r->b, x, k1);
    if (Ys) memcpy(Ys, x, sizeof(ST) * (size_t)n);
    for (int i = 0; i < n; i++) t[i] = x[i] + b21 * hs * k1[i];
    if (Ys) memcpy(Ys + (size_t)n, t, sizeof(ST) * (size_t)n);
    FN(orc_rhs)(r->n, r->m, r->k, r->C, r->b, t, k2);
    for (int i = 0; i < n; i++) t[i] = x[i] + hs * (b31 * k1[i] + b32 * k2[i]);
    if (Ys) memcpy(Ys + 2 * (size_t)n, t, sizeof(ST) * (size_t)n);
    FN(orc_rhs)(r->n, r->m, r->k, r->C, r->b, t, k3);
    for (int i = 0; i < n; i++) t[i] = x[i] + hs * (b41 * k1[i] + b42 * k2[i] + b43 * k3[i]);
    if (Ys) memcpy(Ys + 3 * (size_t)n, t, sizeof(ST) * (size_t)n);
    FN(orc_rhs)(r->n, r->m, r->k, r->C, r->b, t, k4);
    for (int i = 0; i < n; i++) t[i] = x[i] + hs * (b51 * k1[i] + b52 * k2[i] + b53 * k3[i] + b54 * k4[i]);
    if (Ys) memcpy(Ys + 4 * (size_t)n, t, sizeof(ST) * (size_t)n);
    FN(orc_rhs)(r->n, r->m, r->k, r->C, r->b, t, k5);
    for (int i = 0; i < n; i++) t[i] = x[i] + hs * (b61 * k1[i] + b62 * k2[i] + b63 * k3[i] + b64 * k4[i] + b65 * k5[i]);
    if (Ys) memcpy(Ys + 5 * (size_t)n, t, sizeof(ST) * (size_t)n);
    FN(orc_rhs)(r->n, r->m, r->k, r->C, r->b, t, k6);
    for (int i = 0; i < n; i++) xout[i] = x[i] + hs * (c1 * k1[i] + c3 * k3[i] + c4 * k4[i] + c6 * k6[i]);
    xout[0] = (ST)-1.0;
}

/* V after T frozen steps from the public start pub (h[0..T-1] given) */
double FN(orc_frozen_energy)(FN(orc_rk_t) *r, const ST *pub, const double *h, int T)
{
    const int n = r->len;
    ST *x = (ST *)malloc(sizeof(ST) * (size_t)n), *y = (ST *)malloc(sizeof(ST) * (size_t)n);
    double *g = (double *)malloc(sizeof(double) * (size_t)n);
    x[0] = (ST)-1.0;
    memcpy(x + 1, pub, sizeof(ST) * (size_t)(n - 1));
    for (int s = 0; s < T; s++) {
        FN(frozen_step)(r, x, y, h[s], NULL);
        ST *t = x; x = y; y = t;
    }
    double V = FN(orc_energy_grad)(r->n, r->m, r->k, r->C, x, g);
    free(x); free(y); free(g);
    return V;
}

/* Discrete adjoint of T Cash-Karp steps with the given step sizes h[0..T-1]
   (the realised map of an adaptive run, or any frozen sequence): returns V(T)
   and its gradient (public layout [s, a], double) with respect to pub. */
double FN(orc_adjoint_frozen)(FN(orc_rk_t) *r, const ST *pub, const double *h, int T, double *grad)
{
    const int n = r->len;
    ST *X = (ST *)malloc(sizeof(ST) * (size_t)n * (size_t)(T + 1));
    ST *Ys = (ST *)malloc(sizeof(ST) * (size_t)n * 6);
    ST *y = (ST *)malloc(sizeof(ST) * (size_t)n);
    double *lam = (double *)malloc(sizeof(double) * (size_t)n);
    double *mu = (double *)malloc(sizeof(double) * (size_t)n);
    double *nu = (double *)malloc(sizeof(double) * (size_t)n * 6);
    X[0] = (ST)-1.0;
    memcpy(X + 1, pub, sizeof(ST) * (size_t)(n - 1));
    for (int s = 0; s < T; s++) FN(frozen_step)(r, X + (size_t)s * n, X + (size_t)(s + 1) * n, h[s], NULL);
    const double V = FN(orc_energy_grad)(r->n, r->m, r->k, r->C, X + (size_t)T * n, lam);
    lam[0] = 0.0;
    /* Butcher coefficients as doubles of the state-type values used forward */
    const double B[7][6] = {
        {0}, {0},
        {(double)(ST)0.2},
        {(double)(ST)(3.0 / 40.0), (double)(ST)(9.0 / 40.0)},
        {(double)(ST)0.3, (double)(ST)-0.9, (double)(ST)1.2},
        {(double)(ST)(-11.0 / 54.0), (double)(ST)2.5, (double)(ST)(-70.0 / 27.0), (double)(ST)(35.0 / 27.0)},
        {(double)(ST)(1631.0 / 55296.0), (double)(ST)(175.0 / 512.0), (double)(ST)(575.0 / 13824.0),
         (double)(ST)(44275.0 / 110592.0), (double)(ST)(253.0 / 4096.0)}};
    const double Cw[7] = {0, (double)(ST)(37.0 / 378.0), 0, (double)(ST)(250.0 / 621.0),
                          (double)(ST)(125.0 / 594.0), 0, (double)(ST)(512.0 / 1771.0)};
    for (int s = T - 1; s >= 0; s--) {
        const double hh = (double)(ST)h[s];
        FN(frozen_step)(r, X + (size_t)s * n, y, h[s], Ys);
        for (int l = 6; l >= 1; l--) {
            for (int i = 0; i < n; i++) {
                double acc = Cw[l] * lam[i];
                for (int j = l + 1; j <= 6; j++) acc += B[j][l - 1] * nu[(size_t)(j - 1) * n + i];
                mu[i] = hh * acc;
            }
            mu[0] = 0.0;
            FN(orc_vjp)(r->n, r->m, r->k, r->C, Ys + (size_t)(l - 1) * n, mu, nu + (size_t)(l - 1) * n);
        }
        for (int i = 1; i < n; i++)
            lam[i] = lam[i] + nu[i] + nu[(size_t)n + i] + nu[2 * (size_t)n + i] + nu[3 * (size_t)n + i] +
                     nu[4 * (size_t)n + i] + nu[5 * (size_t)n + i];
    }
    memcpy(grad, lam + 1, sizeof(double) * (size_t)(n - 1));
    free(X); free(Ys); free(y); free(lam); free(mu); free(nu);
    return V;
}

/* Adaptive forward of T accepted steps from pub (records h[0..T-1]), then the
   discrete adjoint of those realised steps.  Returns V(T), or NaN on underflow. */
double FN(orc_adjoint)(FN(orc_rk_t) *r, const ST *pub, int T, double *h_out, double *grad)
{
    FN(orc_rk_set_state)(r, pub);
    FN(orc_rk_reset)(r);
    for (int s = 0; s < T; s++) {
        if (!FN(orc_rk_step)(r, 1)) return NAN;
        h_out[s] = (double)r->hdid;
    }
    return FN(orc_adjoint_frozen)(r, pub, h_out, T, grad);
}

#undef FN
#undef CAT
#undef CAT2
