// Drop-in check: the reference Frontend's GPU solve sequence
// (Frontend/src/cnf_runner.cpp:28-127, SolveOnGpu) written against the MI355X
// headers; only the class names differ (CudaSat1 -> HipSat, CudaRungeKutta ->
// HipRungeKutta, CudaSatSolver -> HipSatSolver, CudaRandom -> HipRandom).
// Usage: dropin_run <file.cnf> <eps> <batch> <stepmax> [public-start-file] [nocallback]
// Prints: result, steps, rhs count, violations, then the final spins' signs.
// With "nocallback" no callback is set (the reference's other branch of
// SatSolver::DoSolve); the minimum violation and batch counts print as -1.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "asat/hip_sat.h"

using namespace analogsat;

int main(int argc, char** argv)
{
    if (argc < 5) {
        std::fprintf(stderr, "usage: %s file.cnf eps batch stepmax [start]\n", argv[0]);
        return 2;
    }
    SatProblem problem;
    {
        CnfReader r(argv[1]);
        if (!r.Read(problem)) {
            std::fprintf(stderr, "%s\n", r.GetMessage());
            return 1;
        }
    }
    try {
        ISat<double, HipODEState<double>>* ctds = new HipSat<double>();
        ctds->Set_B(0.0);
        HipRungeKutta<double>* rk = new HipRungeKutta<double>();
        rk->SetEpsilon(std::atof(argv[2]));
        HipSatSolver<double> solver(ctds, rk);
        solver.SetProblem(problem);
        solver.SetMaxTime(1e8);
        solver.SetMaxSteps(std::atoi(argv[4]));
        solver.SetBatchSize(std::atoi(argv[3]));
        int n = problem.Get_N(), m = problem.Get_M();
        const bool callback = !(argc > 6 && std::string(argv[6]) == "nocallback");
        if (argc > 5) {
            std::vector<double> start(n + m);
            std::ifstream in(argv[5]);
            for (auto& v : start) in >> v;
            solver.SetSatState(start);
        } else {
            HipRandom<double> rand(1);
            solver.SetRandomInitialState(rand);
        }
        int minViolation = m, batches = 0;
        if (callback) {
            solver.SetCallback([&](HipSatSolver<double>::Interface& s) -> bool {
                int v = s.GetLastClauseViolationCount();
                if (v < minViolation) minViolation = v;
                batches++;
                return true;
            });
        } else {
            minViolation = batches = -1;
        }
        SatResult result = solver.Solve();
        std::vector<double> last;
        solver.GetSatState().CopyTo(last);
        std::printf("%s\n%d %d %d %d %d\n", GetSatResultMessage(result), solver.GetStepCount(), solver.GetRhsCount(),
                    solver.GetLastClauseViolationCount(), minViolation, batches);
        for (int i = 0; i < n; i++) std::printf("%d ", last[i] > 0 ? i + 1 : -(i + 1));
        std::printf("0\n");
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
