// MI355X implementations of the plugin interfaces, thin RAII wrappers over the
// C ABI in analogsat_hip.h.  They replace the reference's CUDA classes:
//   HipSat<T>         for CudaSat1/2/3<T>    (AnalogSAT/cuda_sat/CudaSat3.h:33-118)
//   HipSatTanh<T>     for CudaSatTanh1/2/3<T> (AnalogSAT/cuda_sat/CudaSatTanh3.cu)
//   HipSatState<T>    for CudaSatState2<T>   (AnalogSAT/cuda_base/CudaSatState2.h:31-106)
//   HipRungeKutta<T>  for CudaRungeKutta<T>  (AnalogSAT/cuda_sat/CudaRungeKutta.h:30-136)
//   HipRandom<T>      for CudaRandom<T>      (AnalogSAT/cuda_base/CudaRandom.h:29-46)
//   HipSatSolver<T>   for CudaSatSolver<T>   (AnalogSAT/solver/SatSolver.h:362-364)
// T = double: FP64 state (the reference's <double>); T = float: FP32 state
// with FP64 step control, error norm and time.  Errors are rethrown as the
// reference's exception types (std::invalid_argument / std::runtime_error);
// Step() returns false on step-size underflow.  Link with -lanalogsat_hip.
#ifndef ASAT_HIP_SAT_H
#define ASAT_HIP_SAT_H

#include <chrono>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../analogsat_hip.h"
#include "interfaces.h"
#include "sat_problem.h"
#include "sat_solver.h"

namespace analogsat {

inline void HipSafe(int rc)
{
    if (rc == ASAT_OK) return;
    if (rc == ASAT_ERR_INVALID_ARGUMENT) throw std::invalid_argument(asat_last_error());
    throw std::runtime_error(asat_last_error());
}

template <typename TFloat>
struct HipPrecision;
template <>
struct HipPrecision<double> {
    static constexpr int value = ASAT_F64;
};
template <>
struct HipPrecision<float> {
    static constexpr int value = ASAT_F32;
};

template <typename TFloat>
class HipSat;

template <typename TFloat>
class HipSatState : public ISatState<TFloat, HipODEState<TFloat>> {
public:
    typedef HipODEState<TFloat> TState;
    explicit HipSatState(asat_sys_t* sys, int n, int m) : ISatState<TFloat, TState>(n, m)
    {
        HipSafe(asat_state_create(sys, &h));
        Refresh();
    }
    ~HipSatState() override { asat_state_destroy(h); }
    HipSatState(const HipSatState&) = delete;
    HipSatState& operator=(const HipSatState&) = delete;

    TState& GetState() override { return Refresh(); }
    const TState& GetState() const override { return const_cast<HipSatState*>(this)->Refresh(); }
    operator TState&() override { return Refresh(); }
    operator const TState&() const override { return const_cast<HipSatState*>(this)->Refresh(); }
    int GetAllocSize() const override
    {
        int s = 0;
        HipSafe(asat_state_alloc_size(h, &s));
        return s;
    }
    void CopyTo(std::vector<TFloat>& target) const override
    {
        target.resize((size_t)this->N + this->M);
        HipSafe(asat_state_get_public(h, target.data()));
    }
    void CopyFrom(const std::vector<TFloat>& source) override
    {
        if (source.size() < (size_t)this->N + this->M) throw std::invalid_argument("source vector is too short");
        HipSafe(asat_state_set_public(h, source.data()));
    }
    using ISatState<TFloat, TState>::CopyFrom;
    void SetZero() override { HipSafe(asat_state_set_zero(h)); }
    asat_state_t* handle() const { return h; }

private:
    asat_state_t* h = nullptr;
    TState ptr = nullptr;
    TState& Refresh()
    {
        void* p = nullptr;
        HipSafe(asat_state_device_ptr(h, &p, nullptr));
        ptr = (TState)p;
        return ptr;
    }
};

template <typename TFloat>
class HipRandom : public ISatRandom<TFloat> {
public:
    HipRandom() : HipRandom((unsigned long long)std::chrono::system_clock::now().time_since_epoch().count()) {}
    explicit HipRandom(unsigned long long seed) { HipSafe(asat_rng_create(seed, &h)); }
    ~HipRandom() override { asat_rng_destroy(h); }
    // addr is a device pointer, as in the reference's CudaRandom
    void GenerateUniform(TFloat* addr, int length) override
    {
        HipSafe(asat_rng_uniform(h, addr, length, HipPrecision<TFloat>::value));
    }
    asat_rng_t* handle() const { return h; }

private:
    asat_rng_t* h = nullptr;
};

template <typename TFloat>
class HipSat : public ISat<TFloat, HipODEState<TFloat>> {
public:
    typedef HipODEState<TFloat> TState;
    typedef HipSatState<TFloat> State;
    HipSat() { HipSafe(asat_sys_create(HipPrecision<TFloat>::value, &h)); }
    ~HipSat() override { asat_sys_destroy(h); }

    void SetProblem(const SatProblem& problem) override
    {
        std::vector<int> lits, lens;
        problem.GetAllClauses(lits, lens);
        HipSafe(asat_sys_set_problem(h, lits.data(), lens.data(), (int)lens.size()));
    }
    int Get_N() const override { return nmk(0); }
    int Get_M() const override { return nmk(1); }
    int Get_K() const override { return nmk(2); }
    TFloat Get_B() const override
    {
        double b = 0;
        HipSafe(asat_sys_get_b(h, &b));
        return (TFloat)b;
    }
    void Set_B(TFloat b) override { HipSafe(asat_sys_set_b(h, (double)b)); }
    State* MakeState() const override
    {
        if (Get_K() == 0) throw std::runtime_error("SAT problem has not been set");
        return new State(h, Get_N(), Get_M());
    }
    void SetRandomState(ISatState<TFloat, TState>& state, ISatRandom<TFloat>& random) override
    {
        auto* st = dynamic_cast<State*>(&state);
        auto* rng = dynamic_cast<HipRandom<TFloat>*>(&random);
        if (st && rng) {
            HipSafe(asat_sys_set_random_state(h, st->handle(), rng->handle()));
            return;
        }
        // generic random source: uniforms on the host, then the reference's transform
        std::vector<TFloat> u((size_t)Get_N() + Get_M());
        random.GenerateUniform(u.data(), (int)u.size());
        for (int i = 0; i < Get_N(); i++) u[i] = u[i] * (TFloat)2 - (TFloat)1;
        for (size_t i = Get_N(); i < u.size(); i++) u[i] = u[i] + (TFloat)1;
        state.CopyFrom(u);
    }
    int GetClauseViolationCount(const ISatState<TFloat, TState>& state) const override
    {
        const State& st = dynamic_cast<const State&>(state);
        int c = 0;
        HipSafe(asat_sys_violations(h, st.handle(), &c));
        return c;
    }
    using ISat<TFloat, TState>::GetClauseViolationCount;
    void GetDerivatives(IODEState<TFloat, TState>& dxdt, const IODEState<TFloat, TState>& x, const TFloat t) override
    {
        State& d = dynamic_cast<State&>(dxdt);
        const State& s = dynamic_cast<const State&>(x);
        HipSafe(asat_sys_rhs(h, d.handle(), s.handle(), (double)t));
    }
    asat_sys_t* handle() const { return h; }

private:
    asat_sys_t* h = nullptr;
    int nmk(int which) const
    {
        int v[3] = {0, 0, 0};
        HipSafe(asat_sys_get_nmk(h, &v[0], &v[1], &v[2]));
        return v[which];
    }
};

// Tanh CTDS family, for CpuSatTanh<T> / CudaSatTanh1-3<T>
// (AnalogSAT/cpu_sat/CpuSatTanh.h, AnalogSAT/cuda_sat/CudaSatTanh3.cu)
template <typename TFloat>
class HipSatTanh : public HipSat<TFloat> {
public:
    HipSatTanh() { HipSafe(asat_sys_set_family(this->handle(), ASAT_FAMILY_TANH)); }
    TFloat Get_Q() const
    {
        double q = 0;
        HipSafe(asat_sys_get_q(this->handle(), &q));
        return (TFloat)q;
    }
    void Set_Q(TFloat q) { HipSafe(asat_sys_set_q(this->handle(), (double)q)); }
};

template <typename TFloat>
class HipRungeKutta : public IODEInt<TFloat, HipODEState<TFloat>> {
public:
    typedef HipODEState<TFloat> TState;
    HipRungeKutta() { HipSafe(asat_rk_create(&h)); }
    ~HipRungeKutta() override { asat_rk_destroy(h); }

    void Configure(std::shared_ptr<IODE<TFloat, TState>> ode) override
    {
        odePtr = ode;
        auto* sat = dynamic_cast<HipSat<TFloat>*>(ode.get());
        if (!sat) throw std::invalid_argument("HipRungeKutta integrates HipSat systems");
        HipSafe(asat_rk_configure(h, sat->handle()));
    }
    bool Step(IODEState<TFloat, TState>& oldState, IODEState<TFloat, TState>& newState, int count) override
    {
        auto& o = dynamic_cast<HipSatState<TFloat>&>(oldState);
        auto& n = dynamic_cast<HipSatState<TFloat>&>(newState);
        int rc = asat_rk_step(h, o.handle(), n.handle(), count);
        if (rc == ASAT_UNDERFLOW) return false;
        HipSafe(rc);
        return true;
    }
    TFloat GetLastStepSize() const override { return (TFloat)stats().hdid; }
    TFloat GetTime() const override { return (TFloat)stats().time; }
    int GetRHSCount() const override { return (int)stats().rhscount; }
    int GetStepCount() const override { return (int)stats().steps; }
    int GetAttemptCount() const override { return (int)stats().attempts; }
    void Reset() override { HipSafe(asat_rk_reset(h)); }
    bool SolveBatches(IODEState<TFloat, TState>& oldState, IODEState<TFloat, TState>& newState, int batch,
                      int maxSteps, TFloat maxTime, SatResult& result, int& steps, int& lastViolations) override
    {
        auto& o = dynamic_cast<HipSatState<TFloat>&>(oldState);
        auto& n = dynamic_cast<HipSatState<TFloat>&>(newState);
        asat_solve_result_t r;
        HipSafe(asat_rk_solve(h, o.handle(), n.handle(), batch, maxSteps, (double)maxTime,
                              sizeof(TFloat) == sizeof(float) ? 1 : 0, &r));
        result = (SatResult)r.result;
        steps = (int)r.steps;
        lastViolations = r.last_violations;
        return true;
    }
    TFloat GetEpsilon() const
    {
        double e = 0;
        HipSafe(asat_rk_get_eps(h, &e));
        return (TFloat)e;
    }
    void SetEpsilon(TFloat eps) { HipSafe(asat_rk_set_eps(h, (double)eps)); }
    // constant term of the error scale (reference: TINY = 1e-20, the default)
    void SetErrorFloor(double floor) { HipSafe(asat_rk_set_xscal_floor(h, floor)); }
    // allow (default) or forbid the one-workgroup persistent path for
    // problems whose integrator fits in one workgroup's LDS
    void SetOneWorkgroup(bool enable) { HipSafe(asat_rk_set_one_workgroup(h, enable ? 1 : 0)); }

private:
    asat_rk_t* h = nullptr;
    std::shared_ptr<IODE<TFloat, TState>> odePtr;
    asat_rk_stats_t stats() const
    {
        asat_rk_stats_t s;
        HipSafe(asat_rk_stats(h, &s));
        return s;
    }
};

template <typename TFloat>
using HipSatSolver = SatSolver<TFloat, HipODEState<TFloat>>;

}  // namespace analogsat

#endif
