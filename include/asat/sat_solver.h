// SatSolver: the batch loop around an ISat and an IODEInt, C++ side of the
// drop-in.  Same public methods and the same semantics as the reference's
// header-only SatSolver (AnalogSAT/solver/SatSolver.h:41-356): Step in
// batches of batchSize accepted steps, take the violated-clause count after
// each batch, call the callback, stop on SAT / callback / underflow, and
// compare maxTime and maxSteps only between batches.
#ifndef ASAT_SAT_SOLVER_H
#define ASAT_SAT_SOLVER_H

#include <chrono>
#include <ctime>
#include <functional>
#include <memory>
#include <stdexcept>
#include <vector>

#include "interfaces.h"

namespace analogsat {

template <typename TFloat, typename TState>
class SatSolver {
public:
    typedef std::function<bool(SatSolver<TFloat, TState>&)> CallbackFunction;
    typedef SatSolver<TFloat, TState> Interface;

    SatSolver(std::unique_ptr<ISat<TFloat, TState>> sat, std::unique_ptr<IODEInt<TFloat, TState>> odeint)
        : satPtr(std::move(sat)), odeintPtr(std::move(odeint))
    {
        Reset();
    }
    SatSolver(ISat<TFloat, TState>* sat, IODEInt<TFloat, TState>* odeint)
        : SatSolver(std::unique_ptr<ISat<TFloat, TState>>(sat), std::unique_ptr<IODEInt<TFloat, TState>>(odeint))
    {
    }
    virtual ~SatSolver() {}

    int GetStepCount() const { return steps; }
    int GetRhsCount() const { return rhsCount; }
    TFloat GetElapsedTime() const { return time; }
    TFloat GetLastStepSize() const { return lastStepSize; }
    double GetElapsedWallTime() const
    {
        if (running) return std::chrono::duration<double>(std::chrono::steady_clock::now() - wallstart).count();
        return walltime;
    }
    double GetElapsedCpuTime() const
    {
        if (running) return double(std::clock() - cpustart) / CLOCKS_PER_SEC;
        return cputime;
    }
    int GetLastClauseViolationCount() const { return lastViolationCount; }

    void SetCallback(CallbackFunction cb) { callback = cb; useCallback = true; }
    void ClearCallback() { useCallback = false; }

    int GetBatchSize() const { return batchSize; }
    void SetBatchSize(int b)
    {
        if (b < 1) throw std::invalid_argument("Batch size must be positive");
        batchSize = b;
    }
    TFloat GetMaxTime() const { return maxTime; }
    void SetMaxTime(TFloat t)
    {
        if (t <= (TFloat)0) throw std::invalid_argument("Max time must be positive");
        maxTime = t;
    }
    int GetMaxSteps() const { return maxSteps; }
    void SetMaxSteps(int s)
    {
        if (s < 1) throw std::invalid_argument("Max iterations must be positive");
        maxSteps = s;
    }

    SatResult Solve()
    {
        running = true;
        ISat<TFloat, TState>& sat = *satPtr;
        IODEInt<TFloat, TState>& odeint = *odeintPtr;
        Reset();
        odeint.Configure(satPtr);
        wallstart = std::chrono::steady_clock::now();
        cpustart = std::clock();
        SatResult result = SAT_UNKNOWN;
        // without a callback the integrator may run the whole loop itself
        if (!(useCallback && callback) &&
            odeint.SolveBatches(*oldState, *newState, batchSize, maxSteps, maxTime, result, steps, lastViolationCount)) {
            walltime = std::chrono::duration<double>(std::chrono::steady_clock::now() - wallstart).count();
            cputime = double(std::clock() - cpustart) / CLOCKS_PER_SEC;
            time = odeint.GetTime();
            rhsCount = odeint.GetRHSCount();
            lastStepSize = odeint.GetLastStepSize();
            running = false;
            return result;
        }
        while (time < maxTime && steps < maxSteps) {
            int before = odeint.GetStepCount();
            if (!odeint.Step(*oldState, *newState, batchSize)) {
                result = SAT_UNDERFLOW;
                break;
            }
            steps += odeint.GetStepCount() - before;
            time = odeint.GetTime();
            lastViolationCount = sat.GetClauseViolationCount(*oldState);
            rhsCount = odeint.GetRHSCount();
            lastStepSize = odeint.GetLastStepSize();
            if (useCallback && callback && !callback(*this)) {
                result = SAT_ODE_INTERRUPTED;
                break;
            }
            if (lastViolationCount == 0) {
                result = SAT_SOLUTION_FOUND;
                break;
            }
        }
        if (result == SAT_UNKNOWN) {
            if (time >= maxTime) result = SAT_MAXTIME_REACHED;
            else if (steps >= maxSteps) result = SAT_MAXITER_REACHED;
        }
        walltime = std::chrono::duration<double>(std::chrono::steady_clock::now() - wallstart).count();
        cputime = double(std::clock() - cpustart) / CLOCKS_PER_SEC;
        time = odeint.GetTime();
        rhsCount = odeint.GetRHSCount();
        lastStepSize = odeint.GetLastStepSize();
        running = false;
        return result;
    }

    const ISatState<TFloat, TState>& GetSatState() const { return *oldState; }
    void SetSatState(const ISatState<TFloat, TState>& s)
    {
        oldState->CopyFrom(s);
        lastViolationCount = satPtr->GetClauseViolationCount(*oldState);
    }
    void SetSatState(const std::vector<TFloat>& s)
    {
        oldState->CopyFrom(s);
        lastViolationCount = satPtr->GetClauseViolationCount(*oldState);
    }
    void SetRandomInitialState(ISatRandom<TFloat>& random)
    {
        satPtr->SetRandomState(*oldState, random);
        lastViolationCount = satPtr->GetClauseViolationCount(*oldState);
    }
    void SetProblem(const SatProblem& problem)
    {
        satPtr->SetProblem(problem);
        Reset();
        oldState.reset(satPtr->MakeState());
        newState.reset(satPtr->MakeState());
        lastViolationCount = satPtr->GetClauseViolationCount(*oldState);
    }

protected:
    std::shared_ptr<ISat<TFloat, TState>> satPtr;
    std::shared_ptr<IODEInt<TFloat, TState>> odeintPtr;
    std::unique_ptr<ISatState<TFloat, TState>> oldState, newState;
    CallbackFunction callback;
    bool useCallback = false;
    TFloat time = 0;
    int steps = 0, rhsCount = 0, lastViolationCount = 0;
    double walltime = 0, cputime = 0;
    TFloat lastStepSize = 0;
    bool running = false;
    TFloat maxTime = (TFloat)100.0;   // reference defaults (SatSolver.h:60-62)
    int batchSize = 20;
    int maxSteps = 100000;
    std::chrono::steady_clock::time_point wallstart;
    std::clock_t cpustart = 0;

    void Reset()
    {
        rhsCount = 0;
        walltime = 0;
        cputime = 0;
        steps = 0;
        time = (TFloat)0;
        lastViolationCount = 0;
    }
};

template <typename TFloat>
using CpuSatSolver = SatSolver<TFloat, CpuODEState<TFloat>>;

}  // namespace analogsat

#endif
