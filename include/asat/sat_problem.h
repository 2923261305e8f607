// Problem container, generator and DIMACS IO for the C++ side of the drop-in.
//
//   SatProblem    same methods as AnalogSAT/problem/SatProblem.h:32-73
//                 (zero literals and empty clauses dropped, N = max |lit|,
//                 K = longest clause)
//   FastSatMaker  same methods as AnalogSAT/problem/FastSatMaker.h:29-61;
//                 instances come from asat_problem_fastsat_stream, which
//                 follows the reference's generator stream seed for seed
//   CnfReader / CnfWriter  same methods as AnalogSAT/io/CnfReader.h,
//                 AnalogSAT/io/CnfWriter.h (DIMACS "p cnf" format)
#ifndef ASAT_SAT_PROBLEM_H
#define ASAT_SAT_PROBLEM_H

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include "../analogsat_hip.h"

namespace analogsat {

class SatProblem {
public:
    SatProblem() : n(0), m(0), k(0) {}

    void AddClause(const std::vector<int>& clause) { AddClauses(clause, std::vector<int>(1, (int)clause.size())); }

    void AddClauses(const std::vector<int>& literals, const std::vector<int>& lengths)
    {
        size_t off = 0;
        for (int len : lengths) {
            int start = (int)storage.size(), len1 = 0;
            for (int i = 0; i < len; i++) {
                int lit = literals[off + i];
                if (lit == 0) continue;
                storage.push_back(lit);
                n = std::max(n, std::abs(lit));
                len1++;
            }
            off += (size_t)len;
            if (len1 == 0) continue;
            starts.push_back(start);
            lens.push_back(len1);
            k = std::max(k, len1);
            m++;
        }
    }

    void AddClauses(const std::vector<std::vector<int>>& clauses)
    {
        for (const auto& c : clauses) AddClause(c);
    }

    void GetAllClauses(std::vector<int>& literals, std::vector<int>& lengths) const
    {
        literals = storage;
        lengths = lens;
    }

    void Clear()
    {
        storage.clear(); starts.clear(); lens.clear();
        n = m = k = 0;
    }

    const int* GetLiterals() const { return storage.data(); }
    int GetClauseStart(int index) const { return starts[index]; }
    int GetClauseLength(int index) const { return lens[index]; }
    int Get_K() const { return k; }
    int Get_N() const { return n; }
    int Get_M() const { return m; }

private:
    std::vector<int> storage, starts, lens;
    int n, m, k;
};

class FastSatMaker {
public:
    FastSatMaker() : FastSatMaker((unsigned long)std::chrono::system_clock::now().time_since_epoch().count()) {}
    explicit FastSatMaker(unsigned long seed) : k(3), n(100), m(325)
    {
        if (asat_problem_seed_engine(seed, &engine) != ASAT_OK) throw std::runtime_error(asat_last_error());
    }
    virtual ~FastSatMaker() {}
    void Set_K(int _k) { k = _k; }
    int Get_K() const { return k; }
    void Set_M(int _m) { m = _m; }
    int Get_M() const { return m; }
    void Set_N(int _n) { n = _n; }
    int Get_N() const { return n; }

    void MakeSatProblem(SatProblem& problem)
    {
        std::vector<int> lits((size_t)k * m), lens(m, k);
        if (asat_problem_fastsat_stream(n, m, k, &engine, lits.data()) != ASAT_OK)
            throw std::invalid_argument(asat_last_error());
        problem.Clear();
        problem.AddClauses(lits, lens);
    }

private:
    int k, n, m;
    unsigned long long engine;
};

class CnfReader {
public:
    explicit CnfReader(const char* filename) : f(std::fopen(filename, "r")) {}
    ~CnfReader()
    {
        if (f) std::fclose(f);
    }
    bool Read(SatProblem& problem)
    {
        if (!f) { message = "file open failed"; return false; }
        int n = -1, m = -1;
        char line[1024];
        while (!std::feof(f)) {
            int c = std::fgetc(f);
            if (c == 'p') {
                if (std::fscanf(f, " cnf %d %d", &n, &m) == 2) break;
                message = "expected: p cnf <variables> <clauses>";
                return false;
            } else if (c != EOF && c != '\n') {
                if (!std::fgets(line, sizeof(line), f)) break;
            }
        }
        if (n < 0 || m < 0) { message = "could not find a SAT declaration"; return false; }
        problem.Clear();
        if (n == 0 || m == 0) { message = "empty problem"; return true; }
        std::vector<int> lits, lens;
        int read = 0;
        while (read < m) {
            int len = 0, lit;
            for (;;) {
                if (std::fscanf(f, "%d", &lit) != 1) { message = "expected more literals in file"; return false; }
                if (lit == 0) break;
                lits.push_back(lit);
                len++;
            }
            lens.push_back(len);
            read++;
        }
        message = "read ok";
        problem.AddClauses(lits, lens);
        return true;
    }
    const char* GetMessage() const { return message.c_str(); }

private:
    std::FILE* f;
    std::string message;
};

class CnfWriter {
public:
    explicit CnfWriter(const char* filename) : f(std::fopen(filename, "w")) {}
    ~CnfWriter()
    {
        if (f) std::fclose(f);
    }
    bool Write(const SatProblem& problem)
    {
        if (!f) return false;
        std::fprintf(f, "p cnf %d %d\n", problem.Get_N(), problem.Get_M());
        for (int j = 0; j < problem.Get_M(); j++) {
            const int* l = problem.GetLiterals() + problem.GetClauseStart(j);
            for (int i = 0; i < problem.GetClauseLength(j); i++) std::fprintf(f, "%d ", l[i]);
            std::fprintf(f, "0\n");
        }
        return true;
    }

private:
    std::FILE* f;
};

}  // namespace analogsat

#endif
