/*
 * analogsat_hip.h -- C ABI of the MI355X (gfx950) AnalogSat CTDS integrator.
 *
 * Plain pointers and sizes only.  Every call returns an int status
 * (ASAT_OK == 0); on failure asat_last_error() returns a thread-local message.
 * The C++ headers under include/asat/ wrap these calls in classes with the
 * reference's method names (ISat / IODEInt / SatSolver) and re-throw the
 * reference's exception types.
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   asat_sys_*     ISat<TFloat,TState> : IODE  (AnalogSAT/solver/ISat.h:31-87,
 *                  AnalogSAT/solver/IODE.h:26-41), implemented by CudaSat1/2/3
 *                  (AnalogSAT/cuda_sat/CudaSat3.h:33-118)
 *   asat_state_*   ISatState / IODEState  (AnalogSAT/solver/ISatState.h:31-161,
 *                  AnalogSAT/solver/IODEState.h:26-66), implemented by
 *                  CudaSatState2 (AnalogSAT/cuda_base/CudaSatState2.h:31-106)
 *   asat_rk_*      IODEInt<TFloat,TState> (AnalogSAT/solver/IODEInt.h:28-60),
 *                  implemented by CudaRungeKutta (AnalogSAT/cuda_sat/CudaRungeKutta.h:30-136)
 *   asat_rng_*     ISatRandom<TFloat> (AnalogSAT/solver/ISatRandom.h:25-33),
 *                  implemented by CudaRandom (AnalogSAT/cuda_base/CudaRandom.h:29-46)
 *   asat_problem_* FastSatMaker (AnalogSAT/problem/FastSatMaker.h:29-61),
 *                  RandomSatMaker (AnalogSAT/problem/RandomSatMaker.h:29-63) and
 *                  SatProblem::AddClauses (AnalogSAT/problem/SatProblem.h:49)
 *   asat_batch_*   no reference counterpart: batched ragged solver (many
 *                  trajectories / instances in one persistent launch)
 *   asat_adj_*     no reference counterpart: discrete adjoint of RK45
 */
#ifndef ANALOGSAT_HIP_H
#define ANALOGSAT_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define ASAT_OK 0
#define ASAT_ERR_INVALID_ARGUMENT 1 /* reference: std::invalid_argument */
#define ASAT_ERR_NOT_CONFIGURED 2   /* reference: runtime_error("SAT problem has not been set") */
#define ASAT_ERR_HIP 3              /* reference: CudaSafe -> runtime_error (AnalogSAT/cuda_base/CudaUtils.cu:27-44) */
#define ASAT_UNDERFLOW 4            /* reference: IODEInt::Step returns false (step-size underflow) */
#define ASAT_ERR_UNSUPPORTED 5

/* precision of the state vectors: FP64 state (reference's <double>) or FP32
   state with FP64 step control, error norm and time */
#define ASAT_F64 0
#define ASAT_F32 1

/* right-hand-side families */
#define ASAT_FAMILY_CTDS 0
#define ASAT_FAMILY_TANH 1

typedef struct asat_sys asat_sys_t;
typedef struct asat_state asat_state_t;
typedef struct asat_rk asat_rk_t;
typedef struct asat_rng asat_rng_t;
typedef struct asat_batch asat_batch_t;

/* last error message of the calling thread ("" if none) */
const char* asat_last_error(void);
/* library version string */
const char* asat_version(void);
/* number of visible GPUs (0 without a GPU); never fails */
int asat_device_count(void);

/* ---------------- problem helpers (host only) ------------------------------ */

/* FastSatMaker(seed) with Set_N(n), Set_M(m), Set_K(k), MakeSatProblem:
   writes m*k literals, clause-major (AnalogSAT/problem/FastSatMaker.cpp:59-107).
   Same libstdc++ minstd_rand0 stream, so the same seed gives the same instance. */
int asat_problem_fastsat(int n, int m, int k, unsigned long seed, int* literals_out);
/* The same generator as a continuing stream (one FastSatMaker object making
   several problems): seed the engine once, then each call advances *engine. */
int asat_problem_seed_engine(unsigned long seed, unsigned long long* engine);
int asat_problem_fastsat_stream(int n, int m, int k, unsigned long long* engine, int* literals_out);
/* RandomSatMaker(seed) (AnalogSAT/problem/RandomSatMaker.cpp:59-125): the
   FastSatMaker draw, with a clause equal to one already made (same signed
   literals, sorted) drawn again, so no clause repeats.  Same stream as
   FastSatMaker.  Asking for more than C(n,k) 2^k clauses, where the
   reference would draw forever, is an ASAT_ERR_INVALID_ARGUMENT. */
int asat_problem_randomsat(int n, int m, int k, unsigned long seed, int* literals_out);
int asat_problem_randomsat_stream(int n, int m, int k, unsigned long long* engine, int* literals_out);

/* ---------------- CTDS system (ISat) --------------------------------------- */

/* new HipSat<T>; precision = ASAT_F64 or ASAT_F32 (CudaSat3<TFloat>() ctor) */
int asat_sys_create(int precision, asat_sys_t** out);
int asat_sys_destroy(asat_sys_t* sys);
/* ISat::SetProblem(SatProblem): literals/lengths as SatProblem::AddClauses
   (zero literals and empty clauses dropped; N = max |lit|, K = max length) */
int asat_sys_set_problem(asat_sys_t* sys, const int* literals, const int* lengths, int nclauses);
/* ISat::Get_N / Get_M / Get_K */
int asat_sys_get_nmk(const asat_sys_t* sys, int* n, int* m, int* k);
/* ISat::Get_B / Set_B: sine bias prefactor b (AnalogSAT/cpu_sat/CpuSat.cpp:76,232-239);
   b > 0 adds mean(a) (M/N) b 2^-2k (pi/2) sin(pi s_i) to ds_i.  The mean is a
   deterministic tree sum (as CudaSat3's CudaReduceSum), not the CPU's serial
   sum, so with b > 0 results match the CPU reference to rounding, not bits.
   Plain solver only; the batched solver and the adjoint use b = 0. */
int asat_sys_set_b(asat_sys_t* sys, double b);
int asat_sys_get_b(const asat_sys_t* sys, double* b);
/* IODE::GetDerivatives(dxdt, x, t) */
int asat_sys_rhs(asat_sys_t* sys, asat_state_t* dxdt, const asat_state_t* x, double t);
/* ISat::GetClauseViolationCount(state) */
int asat_sys_violations(asat_sys_t* sys, const asat_state_t* x, int* count);
/* ISat::SetRandomState(state, random): s = 2u-1, a = u+1 */
int asat_sys_set_random_state(asat_sys_t* sys, asat_state_t* x, asat_rng_t* rng);
/* Right-hand-side family: ASAT_FAMILY_CTDS is CpuSat / CudaSat1-3
   (AnalogSAT/cpu_sat/CpuSat.cpp:197-241); ASAT_FAMILY_TANH is CpuSatTanh /
   CudaSatTanh1-3 (AnalogSAT/cpu_sat/CpuSatTanh.cpp:64-94,192-286), where a
   slot adds knorm a |K_mi| tanh(q f_mi) c_mi.  The reference selects the
   family by class; here it is a setting of the system (plain solver only). */
int asat_sys_set_family(asat_sys_t* sys, int family);
int asat_sys_get_family(const asat_sys_t* sys, int* family);
/* CpuSatTanh::Set_Q / Get_Q (AnalogSAT/cpu_sat/CpuSatTanh.cpp:111-117), default 1.1 */
int asat_sys_set_q(asat_sys_t* sys, double q);
int asat_sys_get_q(const asat_sys_t* sys, double* q);
/* No reference counterpart (the CUDA classes keep clause order on the
   device): order[j] (M ints) = the clause whose aux value sits at position j
   of a state's device aux block.  Copies to and from host memory
   (asat_state_set_public / get_public) already use the public clause order. */
int asat_sys_clause_order(const asat_sys_t* sys, int* order);

/* No reference counterpart (the CUDA classes pick their kernels by class):
   which passes and launch shapes SetProblem chose for this problem. */
#define ASAT_CLAUSE_TILED 0    /* tiles of 2048 clauses, K = 3 or 4 */
#define ASAT_CLAUSE_GENERIC 1  /* thread per clause, any K <= 32 */
#define ASAT_VAR_REGION 0      /* one workgroup per variable group, region in LDS */
#define ASAT_VAR_GATHER 1      /* gather from global memory, many variables per workgroup */
#define ASAT_VAR_WIDE 2        /* gather, one workgroup per variable (long rows) */
typedef struct {
    int clause_pass;     /* ASAT_CLAUSE_* */
    int var_pass;        /* ASAT_VAR_* */
    int groups;          /* variable groups (contribution regions) */
    int group_vars;      /* variables per group */
    int tiles;           /* tiles of the tiled clause pass (0 otherwise) */
    int clause_grid;     /* workgroups of one clause pass */
    int var_grid;        /* workgroups of one var pass */
    int var_lds_bytes;   /* dynamic LDS of the region var pass */
    int slot_spins;      /* 1: the integrator carries the stage spins in contribution layout */
    int direct_primary;  /* 1: each clause's primary slot bypasses the regions (own spin read and term store) */
    int tile_clauses;    /* clauses per tile of the tiled clause pass (0 otherwise) */
    int run_tables;      /* 1: the tiles' sorted positions are rebuilt from per-tile run tables */
    int primary_words;   /* 1: a 16-bit word per clause names its primary variable, literal and sign */
} asat_layout_t;
int asat_sys_layout(const asat_sys_t* sys, asat_layout_t* out);

/* ---------------- states (ISatState) -------------------------------------- */

/* ISat::MakeState() */
int asat_state_create(asat_sys_t* sys, asat_state_t** out);
int asat_state_destroy(asat_state_t* st);
/* ISatState::CopyFrom(vector<TFloat>) / CopyTo(vector<TFloat>&):
   host array [s_1..s_N, a_1..a_M] of the system's precision (double or float) */
int asat_state_set_public(asat_state_t* st, const void* host_pub);
int asat_state_get_public(const asat_state_t* st, void* host_pub);
/* ISatState::CopyBooleanTo: out[i] = s_i > 0 */
int asat_state_get_boolean(const asat_state_t* st, unsigned char* out);
/* IODEState::CopyFrom(const IODEState&) */
int asat_state_copy(asat_state_t* dst, const asat_state_t* src);
/* IODEState::SetZero */
int asat_state_set_zero(asat_state_t* st);
/* IODEState::GetAllocSize */
int asat_state_alloc_size(const asat_state_t* st, int* size);
/* IODEState::GetState(): device pointer of the internal layout
   [-1, s_1..s_N, pad, aux block, pad]; the aux block starts at *aux_offset and
   holds a_c for clause c = order[j] at position j (asat_sys_clause_order) */
int asat_state_device_ptr(asat_state_t* st, void** dev_ptr, int* aux_offset);

/* ---------------- integrator (IODEInt / CudaRungeKutta) -------------------- */

typedef struct {
    double time;      /* IODEInt::GetTime */
    double hdid;      /* IODEInt::GetLastStepSize */
    double hnext;     /* proposed next step */
    double errmax;    /* last error norm divided by eps */
    long long steps;     /* IODEInt::GetStepCount */
    long long attempts;  /* IODEInt::GetAttemptCount */
    long long rhscount;  /* IODEInt::GetRHSCount */
} asat_rk_stats_t;

int asat_rk_create(asat_rk_t** out);
int asat_rk_destroy(asat_rk_t* rk);
/* CudaRungeKutta::SetEpsilon (range [1e-16, 1]) / GetEpsilon */
int asat_rk_set_eps(asat_rk_t* rk, double eps);
int asat_rk_get_eps(const asat_rk_t* rk, double* eps);
/* Constant term of the error scale xscal = |x| + |h dxdt| + floor.  The
   reference fixes it at TINY = 1e-20 (AnalogSAT/cpu_sat/CpuRungeKutta.cpp:84-85),
   which is the default here and reproduces the reference exactly.  A larger
   floor (e.g. 1e-3) adds an absolute-error floor near zero crossings; the FP32
   state needs it at eps ~ 1e-6, where float rounding of the stage inputs
   otherwise drives h towards zero whenever a spin passes through s = 0. */
int asat_rk_set_xscal_floor(asat_rk_t* rk, double floor);
/* Record accepted step sizes on the device: after the k-th accepted step
   since Reset, dev_hlog[k-1] = hdid (for k <= len).  len = 0 switches the
   log off.  No reference counterpart (for callers that replay a run with
   its step sizes frozen, as asat_adj_replay does). */
int asat_rk_set_step_log(asat_rk_t* rk, double* dev_hlog, long long len);
/* IODEInt::Configure(ode): allocates scratch, resets h = 1e-9 */
int asat_rk_configure(asat_rk_t* rk, asat_sys_t* sys);
/* IODEInt::Step(old, new, count): count accepted steps; on return `old`
   holds the newest state (the two device buffers are swapped as needed).
   Returns ASAT_UNDERFLOW if the step size underflowed. */
int asat_rk_step(asat_rk_t* rk, asat_state_t* old_state, asat_state_t* new_state, int count);
/* No reference counterpart.  Problems whose whole integrator fits in one
   workgroup's LDS (about 150 KB: 7 state vectors, the spin stage input, the
   slot contributions, literals and term positions) and that use no bias term run
   Step and asat_rk_solve in one persistent single-workgroup kernel, with the
   same arithmetic as the multi-kernel path (bit-identical results).
   enable = 0 forces the multi-kernel path (13 launches per attempt);
   asat_rk_uses_one_workgroup reports which path the configured problem takes. */
int asat_rk_set_one_workgroup(asat_rk_t* rk, int enable);
int asat_rk_uses_one_workgroup(const asat_rk_t* rk, int* uses);
/* SatSolver::DoSolve's loop without a callback (AnalogSAT/solver/SatSolver.h:313-331):
   while (time < max_time && steps < max_steps) { Step(batch); count the
   violated clauses of old_state; stop on 0 }, ending with SAT_UNDERFLOW on
   step-size underflow.  `time` is the integrator's time, compared as a float
   when time_float != 0 (TFloat = float); the solver's step count starts at 0,
   the integrator's counters continue (SatSolver::Solve configures the
   integrator first).  Problems whose whole integrator fits in one
   workgroup's LDS (and use no bias term) run the loop inside one persistent
   kernel with no host round trip per batch; others run it on the host over
   asat_rk_step and asat_sys_violations.  Same results either way.  No
   reference counterpart as one call (the reference loops on the host). */
typedef struct {
    int result;              /* SatResult code (AnalogSAT/solver/SatResult.h:24-32) */
    int last_violations;     /* SatSolver::lastViolationCount: after the last completed batch, 0 before */
    long long steps;         /* SatSolver::steps: accepted steps of the completed batches */
    int on_device;           /* 1: the loop ran inside the persistent kernel */
    int pad_;
} asat_solve_result_t;
int asat_rk_solve(asat_rk_t* rk, asat_state_t* old_state, asat_state_t* new_state, int batch, long long max_steps,
                  double max_time, int time_float, asat_solve_result_t* out);
/* No reference counterpart.  For callers that cannot allocate device memory
   (code under HIP graph capture, tight serving loops, harnesses that forbid
   hipMalloc): asat_rk_solve for one problem given as host arrays (clauses as
   SatProblem::AddClauses, the public start [s, a] of the given precision),
   with no handles and no allocation.  eps as CudaRungeKutta::SetEpsilon;
   batch, max_steps, max_time and time_float as asat_rk_solve; the integrator
   starts from CudaRungeKutta's reset state (h = 1e-9, error floor 1e-20).
   Writes the newest state to pub_out, the loop's outcome to *out and, if
   stats is not null, the integrator's counters.  Every device buffer (control
   and solver blocks, state, literals, term positions) lives in one
   static array of ASAT_NOALLOC_SCRATCH_BYTES per GPU, held until the process
   ends; the problem must fit it and the one-workgroup path's LDS (about
   150 KB), or the call fails with ASAT_ERR_INVALID_ARGUMENT.  Calls on one
   GPU share that array: calls that overlap on one GPU race.  Results equal
   asat_rk_solve's on the same problem and start, bit for bit. */
#define ASAT_NOALLOC_SCRATCH_BYTES (512 * 1024)
int asat_solve_noalloc(int precision, const int* literals, const int* lengths, int nclauses, const void* pub0,
                       double eps, int batch, long long max_steps, double max_time, int time_float, void* pub_out,
                       asat_solve_result_t* out, asat_rk_stats_t* stats);
/* IODEInt::Reset */
int asat_rk_reset(asat_rk_t* rk);
int asat_rk_stats(const asat_rk_t* rk, asat_rk_stats_t* out);
/* Device time of the last asat_rk_step call in milliseconds (HIP events on
   the integrator's stream); *ms = -1 if none */
int asat_rk_last_step_ms(const asat_rk_t* rk, float* ms);
/* Average device time in microseconds of each kernel of one attempt, measured
   over `reps` launches with HIP events on the integrator's stream.
   us_out[13]: clause pass J=1..6, var pass J=1..6, control.  No reference
   counterpart (measurement hook); the integrator state is left unchanged. */
int asat_rk_profile(asat_rk_t* rk, asat_state_t* old_state, asat_state_t* new_state, int reps, float* us_out);
/* Measurement hook, no reference counterpart: while kind is 0..12 (the
   order of asat_rk_profile), every attempt that Step enqueues records a HIP
   event pair on the integrator's stream just before and after the kernel of
   that kind, and the elapsed times are added up; -1 turns the timer off.
   Setting it clears the totals.  A retry's J = 1 clause pass returns at once
   and is counted as it runs.  asat_rk_kernel_timer reads the total
   milliseconds and the number of timed launches. */
int asat_rk_set_kernel_timer(asat_rk_t* rk, int kind);
int asat_rk_kernel_timer(const asat_rk_t* rk, double* total_ms, long long* launches);

/* Measurement hook, no reference counterpart: copies `bytes` between two
   device buffers `reps` times with a plain 16-byte-per-lane copy kernel
   (k_copy16) and reports the average microseconds per copy.  Used to
   calibrate the HBM byte counters of rocprofv3 against a known traffic. */
int asat_util_copy_bench(long long bytes, int reps, float* us_per_copy);

/* Measurement hook, no reference counterpart: the device's practical HBM
   rates for well-formed streams (16 bytes per lane, four non-temporal loads
   in flight per thread): average microseconds of a copy of `bytes` and of a
   read-only pass over `bytes`, over `reps` launches each. */
int asat_util_stream_bench(long long bytes, int reps, float* us_copy, float* us_read);

/* GetGpuName (AnalogSAT/util/gpuidtool.cpp:28-40): the current device's
   name, NUL-terminated and truncated to len - 1 bytes. */
int asat_util_gpu_name(char* buf, int len);

/* ---------------- random numbers (ISatRandom / CudaRandom) ----------------- */

/* own counter-based generator (Philox4x32-10, Salmon et al., SC'11) */
int asat_rng_create(unsigned long long seed, asat_rng_t** out);
int asat_rng_destroy(asat_rng_t* rng);
/* ISatRandom::GenerateUniform(devptr, len): len uniforms in [0,1) written to
   device memory of the given precision (ASAT_F64 double / ASAT_F32 float) */
int asat_rng_uniform(asat_rng_t* rng, void* dev_ptr, int len, int precision);

/* ---------------- batched ragged solver (no reference counterpart) ---------
 * Many independent trajectories (each its own instance with N <= 8191,
 * clauses of <= 4 literals, K*M <= 262143; own step size, own early exit on
 * SAT) advanced together by persistent launches, one workgroup per
 * trajectory.  Each trajectory follows the same FP32-state / FP64-control
 * Cash-Karp algorithm as asat_rk_step on an ASAT_F32 system, with SAT checked
 * after every `check_every` accepted steps (SatSolver batch size). */

#define ASAT_BATCH_RUNNING 0
#define ASAT_BATCH_SOLVED 1    /* reference: SAT_SOLUTION_FOUND */
#define ASAT_BATCH_MAXSTEPS 2  /* reference: SAT_MAXITER_REACHED */
#define ASAT_BATCH_UNDERFLOW 3 /* reference: SAT_UNDERFLOW */
#define ASAT_BATCH_MAXTIME 4   /* reference: SAT_MAXTIME_REACHED */

typedef struct {
    int status;              /* ASAT_BATCH_* */
    long long steps, attempts, rhscount;
    double time, hdid, hnext;
    int min_violations;      /* minimum violated-clause count over the checked steps (MAXSAT) */
    int violations;          /* count at the last check */
    int sat_step;            /* accepted step at which SAT was found, -1 if not */
    double sat_time;         /* analog time at SAT, -1 if not */
    double solved_wall_ms;   /* host wall time since asat_batch_run started when SAT was seen */
} asat_batch_result_t;

typedef struct {
    int trajectories, max_n, max_m, max_slots;
    int block;               /* workgroup size chosen (256 / 512 / 1024) */
    int lds;                 /* 1: contributions in LDS, 0: in global scratch, 2: register-resident kernel,
                                3: clause-chunked kernel (no contribution array) */
    int grid;                /* resident workgroups */
    long long lds_bytes;
    double tuned_us_per_attempt;   /* autotuner's measurement, -1 if not tuned */
} asat_batch_bucket_t;

int asat_batch_create(asat_batch_t** out);
int asat_batch_destroy(asat_batch_t* b);
/* add one trajectory: clauses as SatProblem::AddClauses, FP32 start [s, a] */
int asat_batch_add(asat_batch_t* b, const int* literals, const int* lengths, int nclauses, const float* pub0,
                   int* index_out);
/* eps as CudaRungeKutta::SetEpsilon; max_steps / max_time as SatSolver;
   chunk_steps: accepted steps per trajectory per launch; check_every: SAT check cadence */
int asat_batch_set_params(asat_batch_t* b, double eps, long long max_steps, double max_time, int chunk_steps,
                          int check_every);
/* xscal floor as asat_rk_set_xscal_floor (default 1e-20) */
int asat_batch_set_xscal_floor(asat_batch_t* b, double floor);
/* block: 0 autotune or 256/512/1024; lds_mode: -1 auto, 0 contributions in
   global scratch, 1 in LDS, 2 register-resident kernel (state and stage
   derivatives in registers, contributions in LDS; its workgroup size is
   fixed by the instance size), 3 clause-chunked kernel (per chunk of
   clauses, their r_m in LDS and each variable's terms of that chunk formed
   from 8-byte var-major records; for instances whose contributions do not
   fit in LDS); nbuckets: size buckets (trajectories split by work, one
   launch each) */
int asat_batch_set_tiling(asat_batch_t* b, int block, int lds_mode, int nbuckets);
/* clause-chunked kernel: at most `clauses` clauses per chunk (0: as many as
   the LDS holds); results do not depend on it */
int asat_batch_set_chunk(asat_batch_t* b, int clauses);
/* upload and (autotune != 0) pick the tile mapping per size bucket */
int asat_batch_prepare(asat_batch_t* b, int autotune);
/* launch graphs of graph_chunks chunks until all trajectories finished or
   max_chunks chunks were launched */
int asat_batch_run(asat_batch_t* b, int max_chunks, int graph_chunks);
/* restore every trajectory to its start */
int asat_batch_reset(asat_batch_t* b);
int asat_batch_size(const asat_batch_t* b, int* count);
int asat_batch_result(const asat_batch_t* b, int index, asat_batch_result_t* out);
int asat_batch_get_state(const asat_batch_t* b, int index, float* pub);
int asat_batch_stats(const asat_batch_t* b, double* run_ms, int* chunks, int* nbuckets);
int asat_batch_bucket_info(const asat_batch_t* b, int bucket, asat_batch_bucket_t* out);

/* ---------------- discrete adjoint (no reference counterpart) --------------
 * Gradient of V(T) = sum_m a_m K_m^2 after T accepted steps of the FP32-state
 * integrator, with respect to the start state [s(0), a(0)], for a batch of
 * trajectories of one instance (k = 3 or 4).  Forward: adaptive, step sizes
 * logged, checkpoint every C steps.  Backward: recompute from checkpoints,
 * discrete adjoint of the realised steps in FP64. */
typedef struct asat_adj asat_adj_t;
int asat_adj_create(asat_adj_t** out);
int asat_adj_destroy(asat_adj_t* a);
int asat_adj_set_problem(asat_adj_t* a, const int* literals, const int* lengths, int nclauses);
/* eps as CudaRungeKutta::SetEpsilon; steps T; checkpoint interval C */
int asat_adj_set_params(asat_adj_t* a, double eps, int steps, int checkpoint_every);
/* adaptive forward of ntraj trajectories (pub0: ntraj x (n+m) floats [s, a]);
   V_out[ntraj] = V(T) */
int asat_adj_forward(asat_adj_t* a, int ntraj, const float* pub0, double* V_out);
/* accepted step sizes of the last forward run, ntraj x T */
int asat_adj_get_steps(const asat_adj_t* a, double* h_out);
/* grad_out[ntraj x (n+m)] = dV(T)/d[s(0), a(0)] of the last forward run (FP64) */
int asat_adj_backward(asat_adj_t* a, double* grad_out);
/* replay T steps with given step sizes h (ntraj x T), FP32 like the forward:
   V_out[ntraj] and optionally the end states xT_out[ntraj x (n+m)] */
int asat_adj_replay(asat_adj_t* a, int ntraj, const float* pub0, const double* h, double* V_out, float* xT_out);
/* the same frozen-step map with FP64 state (the reference's <double>
   arithmetic): V_out[ntraj] = V(T) from pub0 (ntraj x (n+m) doubles); the
   FP64 reference for finite-difference checks of the gradient */
int asat_adj_replay_f64(asat_adj_t* a, int ntraj, const double* pub0, const double* h, double* V_out);
/* device milliseconds of the last forward / backward */
int asat_adj_timing(const asat_adj_t* a, double* forward_ms, double* backward_ms);
/* Measurement hook, no counterpart in the reference: attempts of the last
   forward run, summed over the trajectories, and lockstep attempts launched
   (every launch covers all trajectories; finished ones return at once). */
int asat_adj_attempts(const asat_adj_t* a, long long* trajectory_attempts, long long* lockstep_attempts);

#ifdef __cplusplus
}
#endif

#endif /* ANALOGSAT_HIP_H */
