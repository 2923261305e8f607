// extern "C" binding of the reference's CPU solver classes -- TEST INFRASTRUCTURE ONLY.
//
// Compiled by oracle/Makefile together with the reference's own CPU sources
// (read in place under /root/reference, never copied) into
// oracle/_ref/libasat_ref.so.  The tests use it to write golden fixtures and
// to pin the C restatement in asat_oracle.c bit for bit.  The classes used are
// the reference's: CpuSat<double> and CpuSat<float> (AnalogSAT/cpu_sat/CpuSat.h),
// CpuRungeKutta<double> (AnalogSAT/cpu_sat/CpuRungeKutta.h), SatSolver
// (AnalogSAT/solver/SatSolver.h), SatProblem, FastSatMaker, RandomSatMaker and
// CpuRandom.
#include <memory>
#include <stdexcept>
#include <vector>

#include "cpu_sat/CpuSat.h"
#include "cpu_sat/CpuSatTanh.h"
#include "cpu_sat/CpuRungeKutta.h"
#include "cpu_sat/CpuRandom.h"
#include "solver/SatSolver.h"
#include "problem/SatProblem.h"
#include "problem/FastSatMaker.h"
#include "problem/RandomSatMaker.h"

using namespace analogsat;

namespace {

SatProblem make_problem(const int* lits, const int* lens, int nclauses)
{
    size_t total = 0;
    for (int j = 0; j < nclauses; j++) total += (size_t)lens[j];
    std::vector<int> l(lits, lits + total), n(lens, lens + nclauses);
    SatProblem p;
    p.AddClauses(l, n);
    return p;
}

struct RefTraj {
    SatProblem problem;
    std::shared_ptr<CpuSat<double>> sat;
    CpuRungeKutta<double> rk;
    std::unique_ptr<ISatState<double, std::vector<double>>> oldS, newS;
};

}  // namespace

// CpuSatTanh<T>::GetDerivatives on the internal layout [-1, s, a]
template <typename T>
static int rhs_tanh(const int* lits, const int* lens, int nclauses, double b, double q, const T* x, T* dxdt)
{
    SatProblem p = make_problem(lits, lens, nclauses);
    CpuSatTanh<T> sat;
    sat.Set_B((T)b);
    sat.Set_Q((T)q);
    sat.SetProblem(p);
    std::unique_ptr<CpuSatStateImpl<T>> xs(sat.MakeState()), ds(sat.MakeState());
    int D = p.Get_N() + p.Get_M() + 1;
    xs->GetState().assign(x, x + D);
    sat.GetDerivatives(*ds, *xs, (T)0);
    for (int i = 0; i < D; i++) dxdt[i] = ds->GetState()[i];
    return 0;
}

// CpuSat<long double> + CpuRungeKutta<long double>: `steps` accepted steps one
// at a time from pub0; the last state and every step size
static int traj_ld(const int* lits, const int* lens, int nclauses, double eps, const long double* pub0, int steps,
                   long double* pub_out, long double* h_out)
{
    SatProblem p = make_problem(lits, lens, nclauses);
    auto sat = std::make_shared<CpuSat<long double>>();
    sat->SetProblem(p);
    CpuRungeKutta<long double> rk;
    rk.SetEpsilon((long double)eps);
    rk.Configure(sat);
    std::unique_ptr<ISatState<long double, std::vector<long double>>> a(sat->MakeState()), b(sat->MakeState());
    const int D = p.Get_N() + p.Get_M();
    std::vector<long double> v(pub0, pub0 + D);
    a->CopyFrom(v);
    for (int s = 0; s < steps; s++) {
        if (!rk.Step(*a, *b, 1)) return s;
        h_out[s] = rk.GetLastStepSize();
    }
    a->CopyTo(v);
    for (int i = 0; i < D; i++) pub_out[i] = v[i];
    return steps;
}

extern "C" {

int ref_traj_ld(const int* lits, const int* lens, int nclauses, double eps, const long double* pub0, int steps,
                long double* pub_out, long double* h_out)
{
    return traj_ld(lits, lens, nclauses, eps, pub0, steps, pub_out, h_out);
}


int ref_problem_sizes(const int* lits, const int* lens, int nclauses, int* n, int* m, int* k)
{
    SatProblem p = make_problem(lits, lens, nclauses);
    *n = p.Get_N(); *m = p.Get_M(); *k = p.Get_K();
    return 0;
}

}  // extern "C"

// CpuSat<T>::GetDerivatives on the internal layout [-1, s, a]
template <typename T>
static int rhs_ctds(const int* lits, const int* lens, int nclauses, double b, const T* x, T* dxdt)
{
    SatProblem p = make_problem(lits, lens, nclauses);
    CpuSat<T> sat;
    sat.Set_B((T)b);
    sat.SetProblem(p);
    std::unique_ptr<CpuSatStateImpl<T>> xs(sat.MakeState()), ds(sat.MakeState());
    int D = p.Get_N() + p.Get_M() + 1;
    xs->GetState().assign(x, x + D);
    sat.GetDerivatives(*ds, *xs, (T)0);
    for (int i = 0; i < D; i++) dxdt[i] = ds->GetState()[i];
    return 0;
}

extern "C" {

// one RHS evaluation on the internal layout [-1, s, a]
int ref_rhs_f64(const int* lits, const int* lens, int nclauses, double b, const double* x, double* dxdt)
{
    return rhs_ctds<double>(lits, lens, nclauses, b, x, dxdt);
}

// the same with the reference's float instantiation (AnalogSAT/cpu_sat/CpuSat.cpp:391)
int ref_rhs_f32(const int* lits, const int* lens, int nclauses, double b, const float* x, float* dxdt)
{
    return rhs_ctds<float>(lits, lens, nclauses, b, x, dxdt);
}

int ref_rhs_tanh_f64(const int* lits, const int* lens, int nclauses, double b, double q, const double* x, double* dxdt)
{
    return rhs_tanh<double>(lits, lens, nclauses, b, q, x, dxdt);
}

int ref_rhs_tanh_f32(const int* lits, const int* lens, int nclauses, double b, double q, const float* x, float* dxdt)
{
    return rhs_tanh<float>(lits, lens, nclauses, b, q, x, dxdt);
}

int ref_violations_f64(const int* lits, const int* lens, int nclauses, const double* pub)
{
    SatProblem p = make_problem(lits, lens, nclauses);
    CpuSat<double> sat;
    sat.SetProblem(p);
    std::vector<double> v(pub, pub + p.Get_N() + p.Get_M());
    return sat.GetClauseViolationCount(v);
}

void* ref_rk_create(const int* lits, const int* lens, int nclauses, double eps)
{
    RefTraj* t = new RefTraj();
    t->problem = make_problem(lits, lens, nclauses);
    t->sat = std::make_shared<CpuSat<double>>();
    t->sat->SetProblem(t->problem);
    t->rk.SetEpsilon(eps);
    t->rk.Configure(t->sat);
    t->oldS.reset(t->sat->MakeState());
    t->newS.reset(t->sat->MakeState());
    return t;
}

void ref_rk_destroy(void* h) { delete (RefTraj*)h; }

void ref_rk_set_state(void* h, const double* pub)
{
    RefTraj* t = (RefTraj*)h;
    std::vector<double> v(pub, pub + t->problem.Get_N() + t->problem.Get_M());
    t->oldS->CopyFrom(v);
}

void ref_rk_get_state(void* h, double* pub)
{
    RefTraj* t = (RefTraj*)h;
    std::vector<double> v;
    t->oldS->CopyTo(v);
    for (size_t i = 0; i < v.size(); i++) pub[i] = v[i];
}

int ref_rk_step(void* h, int count)
{
    RefTraj* t = (RefTraj*)h;
    return t->rk.Step(*t->oldS, *t->newS, count) ? 1 : 0;
}

double ref_rk_time(void* h) { return ((RefTraj*)h)->rk.GetTime(); }
double ref_rk_hdid(void* h) { return ((RefTraj*)h)->rk.GetLastStepSize(); }
long ref_rk_steps(void* h) { return ((RefTraj*)h)->rk.GetStepCount(); }
long ref_rk_attempts(void* h) { return ((RefTraj*)h)->rk.GetAttemptCount(); }
long ref_rk_rhscount(void* h) { return ((RefTraj*)h)->rk.GetRHSCount(); }
int ref_rk_violations(void* h)
{
    RefTraj* t = (RefTraj*)h;
    return t->sat->GetClauseViolationCount(*t->oldS);
}

// SatSolver<double>::Solve with a callback that records the violation count
// after every batch.  stats: [steps, rhscount, result]; time_out: analog time.
int ref_solve(const int* lits, const int* lens, int nclauses, const double* pub0,
              double eps, int batch, int max_steps, double max_time,
              int* out_viol, int max_batches, int* n_batches,
              double* out_pub, long* stats, double* time_out)
{
    SatProblem p = make_problem(lits, lens, nclauses);
    CpuRungeKutta<double>* rk = new CpuRungeKutta<double>();
    rk->SetEpsilon(eps);
    CpuSatSolver<double> solver(new CpuSat<double>(), rk);
    solver.SetProblem(p);
    solver.SetMaxTime(max_time);
    solver.SetMaxSteps(max_steps);
    solver.SetBatchSize(batch);
    std::vector<double> v(pub0, pub0 + p.Get_N() + p.Get_M());
    solver.SetSatState(v);
    int nb = 0;
    solver.SetCallback([&](CpuSatSolver<double>::Interface& s) -> bool {
        if (nb < max_batches) out_viol[nb] = s.GetLastClauseViolationCount();
        nb++;
        return true;
    });
    SatResult r = solver.Solve();
    *n_batches = nb;
    std::vector<double> outv;
    solver.GetSatState().CopyTo(outv);
    for (size_t i = 0; i < outv.size(); i++) out_pub[i] = outv[i];
    stats[0] = solver.GetStepCount();
    stats[1] = solver.GetRhsCount();
    stats[2] = (long)r;
    *time_out = solver.GetElapsedTime();
    return (int)r;
}

// FastSatMaker(seed): k-SAT with n variables and m clauses, flat clause-major literals
int ref_fastsat(int n, int m, int k, unsigned long seed, int* out)
{
    FastSatMaker maker(seed);
    maker.Set_K(k);
    maker.Set_N(n);
    maker.Set_M(m);
    SatProblem p;
    maker.MakeSatProblem(p);
    std::vector<int> lits, lens;
    p.GetAllClauses(lits, lens);
    for (size_t i = 0; i < lits.size(); i++) out[i] = lits[i];
    return (int)lits.size();
}

// RandomSatMaker(seed): as FastSatMaker, but a clause equal to one already made
// (same signed literals, in sorted order) is drawn again
int ref_randomsat(int n, int m, int k, unsigned long seed, int* out)
{
    RandomSatMaker maker(seed);
    maker.Set_K(k);
    maker.Set_N(n);
    maker.Set_M(m);
    SatProblem p;
    maker.MakeSatProblem(p);
    std::vector<int> lits, lens;
    p.GetAllClauses(lits, lens);
    for (size_t i = 0; i < lits.size(); i++) out[i] = lits[i];
    return (int)lits.size();
}

// CpuRandom<double>(seed).GenerateUniform
void ref_uniform_f64(unsigned int seed, double* out, int len)
{
    CpuRandom<double> r(seed);
    r.GenerateUniform(out, len);
}

}  // extern "C"
