// p-multigrid preconditioner for the lockstep multi-column solver: levels are DG operators of
// decreasing polynomial degree on one mesh, joined by nodal interpolation.  tandem sets up the same
// hierarchy as a PETSc PCMG (app/common/PetscLinearSolver.cpp:53-114: levels from MGConfig, the
// interpolation NodalInterpolation::assemble, Galerkin coarse operators) with Chebyshev / Jacobi
// smoothing and a loosely solved CG coarse level (examples/options/mg_cheby.cfg).
#pragma once

#include "krylov.h"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <vector>

namespace tdg {

class Operator;

struct PMGOptions {
    int smooth = 6;             // Chebyshev steps per pre- and per post-smoothing
    double eig_lo = 0.05;       // Chebyshev interval [eig_lo, eig_hi] x estimated largest eigenvalue
    double eig_hi = 1.1;        //   of D^-1 A (D: element-block diagonal)
    int eig_iters = 12;         // preconditioned CG (Lanczos) steps of the estimate
    double coarse_rtol = 0.1;   // coarse CG (block Jacobi) tolerance per V-cycle and column
    int coarse_max_iter = 500;  // coarse CG iteration cap per V-cycle
    int galerkin_penalty = 1;   // coarse levels take the fine level's facet penalty
    int coarse_fp32 = 1;        // coarse CG in FP32 (FP64 right-hand side and solution)
    int vcycle_fp32 = 1;        // smoothing and transfers in FP32 (the outer flexible CG stays FP64)
    int coarse_assembled = 1;   // coarsest level on its assembled block rows where it has the register
                                //   kernel (P1 / P2)
    int coarse_amg = 1;         // coarse CG preconditioned by smoothed-aggregation AMG on those block
                                //   rows (elasticity), else block Jacobi
};

struct PMGLevel;
struct MultiWork;
class AMG;

class PMG {
public:
    // levels[0] is the fine operator, levels[1..] coarser ones on the same mesh (strictly
    // decreasing degree, same kind); all must have their material set.  The coarse operators run
    // on the fine operator's stream.
    PMG(std::vector<Operator*> levels, PMGOptions const& o);
    ~PMG();
    PMG(PMG const&) = delete;
    PMG& operator=(PMG const&) = delete;

    Operator& fine() { return *ops_[0]; }
    int num_levels() const { return int(ops_.size()); }
    // Block-Jacobi inverses of every level and the eigenvalue estimates (once; also done by the
    // first V-cycle).
    void setup();
    // z = M^-1 r: one V-cycle on ncols fine-level columns at stride ld.
    void vcycle(double const* r, double* z, int ncols, int64_t ld);
    std::vector<double> const& lambda_max() const { return lmax_; }
    // Coarse CG iterations summed over the V-cycles so far, and the V-cycle count.
    int64_t coarse_iterations() const { return coarse_its_; }
    int64_t cycles() const { return cycles_; }

private:
    template <class T> void smooth(int l, T const* b, T* x, int ncols, int64_t ld, bool zero_guess, bool residual);
    template <class T> void cycle(int l, T const* b, T* x, int ncols, int64_t ld);
    double estimate(int l);

    std::vector<Operator*> ops_;
    PMGOptions o_;
    std::vector<std::unique_ptr<PMGLevel>> lv_;
    std::vector<double> lmax_;
    std::unique_ptr<MultiWork> coarse_work_;
    std::unique_ptr<AMG> amg_;
    int64_t coarse_its_ = 0, cycles_ = 0;
    bool ready_ = false;
};

// ncols right-hand sides (column c at B + c ld) solved together by flexible CG (one truncation
// direction) preconditioned with one V-cycle per iteration, FP64; X overwritten (initial guess 0).
// o.rtol and o.max_iter apply per column.
MultiSolveResult solve_multi_pmg(PMG& M, double const* B, double* X, int ncols, int64_t ld, SolveOptions const& o);

} // namespace tdg
