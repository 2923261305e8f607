// Plugin interfaces of the AnalogSat solver, C++ side of the MI355X drop-in.
//
// Same class and method names as the reference's plugin API so that caller
// code compiles unchanged against either library:
//   IODEState  AnalogSAT/solver/IODEState.h:26-66
//   ISatState  AnalogSAT/solver/ISatState.h:31-161
//   IODE       AnalogSAT/solver/IODE.h:26-41
//   IODEInt    AnalogSAT/solver/IODEInt.h:28-60
//   ISatRandom AnalogSAT/solver/ISatRandom.h:25-33
//   ISat       AnalogSAT/solver/ISat.h:31-87
//   SatResult  AnalogSAT/solver/SatResult.h:24-32
// Header-only; nothing here touches the GPU.
#ifndef ASAT_INTERFACES_H
#define ASAT_INTERFACES_H

#include <memory>
#include <stdexcept>
#include <vector>

#include "sat_problem.h"

namespace analogsat {

enum SatResult {
    SAT_SOLUTION_FOUND,
    SAT_UNDERFLOW,
    SAT_ODE_INTERRUPTED,
    SAT_MAXTIME_REACHED,
    SAT_MAXITER_REACHED,
    SAT_UNKNOWN
};

inline const char* GetSatResultMessage(SatResult r)
{
    switch (r) {
    case SAT_SOLUTION_FOUND: return "solution found";
    case SAT_UNDERFLOW: return "underflow";
    case SAT_ODE_INTERRUPTED: return "user interrupt";
    case SAT_MAXTIME_REACHED: return "maximum time reached";
    case SAT_MAXITER_REACHED: return "maximum steps reached";
    default: return "unknown result";
    }
}

// storage type of a state: host vector or device pointer
template <typename TFloat>
using CpuODEState = std::vector<TFloat>;
template <typename TFloat>
using HipODEState = TFloat*;

template <typename TFloat, typename TState>
class IODEState {
public:
    virtual ~IODEState() {}
    virtual TState& GetState() = 0;
    virtual const TState& GetState() const = 0;
    virtual int GetAllocSize() const = 0;
    virtual void CopyTo(std::vector<TFloat>& target) const = 0;
    virtual void CopyFrom(const std::vector<TFloat>& source) = 0;
    virtual void CopyFrom(const IODEState<TFloat, TState>& other) = 0;
    virtual operator TState&() = 0;
    virtual operator const TState&() const = 0;
    virtual void SetZero() = 0;
};

// public view [s_1..s_N, a_1..a_M]; the internal layout stays hidden
template <typename TFloat, typename TState>
class ISatState : public IODEState<TFloat, TState> {
public:
    ~ISatState() override {}
    int Get_N() const { return N; }
    int Get_M() const { return M; }
    void CopyBooleanTo(std::vector<bool>& target) const
    {
        std::vector<TFloat> v;
        this->CopyTo(v);
        target.resize(N);
        for (int i = 0; i < N; i++) target[i] = v[i] > 0;
    }
    void CopyFrom(const IODEState<TFloat, TState>& other) override
    {
        std::vector<TFloat> tmp;
        other.CopyTo(tmp);
        this->CopyFrom(tmp);
    }
    using IODEState<TFloat, TState>::CopyFrom;

protected:
    ISatState(int n, int m) : N(n), M(m) {}
    int N, M;
};

template <typename TFloat, typename TState>
class IODE {
public:
    virtual ~IODE() {}
    virtual void GetDerivatives(IODEState<TFloat, TState>& dxdt, const IODEState<TFloat, TState>& state,
                                const TFloat time) = 0;
    virtual IODEState<TFloat, TState>* MakeState() const = 0;
};

template <typename TFloat, typename TState>
class IODEInt {
public:
    virtual ~IODEInt() {}
    virtual void Configure(std::shared_ptr<IODE<TFloat, TState>> odePtr) = 0;
    virtual bool Step(IODEState<TFloat, TState>& oldState, IODEState<TFloat, TState>& newState, int count) = 0;
    virtual TFloat GetLastStepSize() const = 0;
    virtual TFloat GetTime() const = 0;
    virtual int GetRHSCount() const = 0;
    virtual int GetStepCount() const = 0;
    virtual int GetAttemptCount() const = 0;
    virtual void Reset() = 0;
    // No reference counterpart: SatSolver's loop without a callback
    // (AnalogSAT/solver/SatSolver.h:313-331) run by the integrator itself.
    // Batches of `batch` accepted steps, the violated-clause count of
    // oldState after each; fills result, the solver's step count and the
    // last count, and returns true.  The default returns false, and
    // SatSolver then runs the loop on the host.
    virtual bool SolveBatches(IODEState<TFloat, TState>& oldState, IODEState<TFloat, TState>& newState, int batch,
                              int maxSteps, TFloat maxTime, SatResult& result, int& steps, int& lastViolations)
    {
        (void)oldState; (void)newState; (void)batch; (void)maxSteps; (void)maxTime; (void)result; (void)steps;
        (void)lastViolations;
        return false;
    }
};

template <class TFloat>
class ISatRandom {
public:
    virtual ~ISatRandom() {}
    virtual void GenerateUniform(TFloat* addr, int length) = 0;
};

template <typename TFloat, typename TState>
class ISat : public IODE<TFloat, TState> {
public:
    ~ISat() override {}
    virtual void SetProblem(const SatProblem& problem) = 0;
    virtual int Get_N() const = 0;
    virtual int Get_M() const = 0;
    virtual int Get_K() const = 0;
    virtual TFloat Get_B() const = 0;
    virtual void Set_B(TFloat b) = 0;
    ISatState<TFloat, TState>* MakeState() const override = 0;
    virtual void SetRandomState(ISatState<TFloat, TState>& state, ISatRandom<TFloat>& random) = 0;
    ISatState<TFloat, TState>* MakeRandomState(ISatRandom<TFloat>& random)
    {
        ISatState<TFloat, TState>* s = MakeState();
        SetRandomState(*s, random);
        return s;
    }
    virtual int GetClauseViolationCount(const ISatState<TFloat, TState>& state) const = 0;
    int GetClauseViolationCount(const std::vector<TFloat>& state) const
    {
        std::unique_ptr<ISatState<TFloat, TState>> tmp(MakeState());
        tmp->CopyFrom(state);
        return GetClauseViolationCount(*tmp);
    }
};

}  // namespace analogsat

#endif
