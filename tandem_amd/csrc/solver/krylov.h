// GPU Krylov solvers over the matrix-free operator.
#pragma once

#include <cstdint>
#include <vector>

namespace tdg {

class Operator;

enum SolveMode : int { Solve_FP64_CG = 0, Solve_FP32_CG_FP64_IR = 1, Solve_FP32_CG_FP64_RELIABLE = 2 };
enum Precond : int { Precond_None = 0, Precond_BlockJacobi = 1, Precond_BlockJacobiF32 = 2 };

struct SolveOptions {
    double rtol = 1e-10;      // on the FP64 true residual ||b - A x|| / ||b||
    int max_iter = 100000;    // total (inner) CG iterations
    int mode = Solve_FP64_CG;
    double inner_rtol = 1e-4; // FP32 inner solve tolerance (restarted refinement mode)
    int max_outer = 50;
    int check_every = 10;     // host reads the residual every check_every iterations
    double reliable_delta = 0.1; // residual drop that triggers an FP64 replacement (mode 2)
    int precond = Precond_None;  // element-block Jacobi (built on first use, FP64 and FP32 inverses);
                                 // Precond_BlockJacobiF32: FP64 solves also apply the FP32-stored blocks
    int graphs = 1;              // replay the iterations between host checks as a captured HIP graph
                                 // (needs a non-default stream and an even check_every)
};

struct SolveResult {
    int iterations = 0;       // total CG iterations (inner for refinement)
    int outer_iterations = 0;
    int inner_iterations = 0;
    double rel_residual = 0.0;
    double seconds = 0.0;
    int converged = 0;
};

// b, x: device FP64 vectors of op.local_size(); x is overwritten (initial guess 0).
SolveResult solve(Operator& op, double const* b, double* x, SolveOptions const& opts);

struct MultiSolveResult {
    int total_iterations = 0;          // lockstep iterations (fused applies)
    int replacements = 0;              // FP64 residual replacements summed over columns
    std::vector<int> iterations;       // per column: iterations until it converged
    std::vector<double> rel_residual;  // per column: FP64 true relative residual
    double seconds = 0.0;
    int converged = 0;                 // all columns
};

// ncols right-hand sides B (column c at B + c ld) solved together, each iteration one batched
// apply; modes Solve_FP64_CG and Solve_FP32_CG_FP64_RELIABLE.  X overwritten (initial guess 0).
MultiSolveResult solve_multi(Operator& op, double const* B, double* X, int ncols, int64_t ld, SolveOptions const& o);

// Independent meshes held in one operator: segment c of b and x is [offsets[c], offsets[c+1]) (whole
// elements, no facet between segments); each segment is its own CG recursion (own scalars,
// convergence and FP64 replacements), all advanced in lockstep by one apply of the operator.
MultiSolveResult solve_segments(Operator& op, double const* b, double* x, int nseg, int64_t const* offsets,
                                SolveOptions const& o);

} // namespace tdg
