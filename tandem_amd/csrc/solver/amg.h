// Smoothed-aggregation algebraic multigrid on the assembled block rows of a low-order elasticity
// operator: the preconditioner of the p-multigrid coarse CG.  tandem's p-multigrid configuration
// solves its coarsest level by CG preconditioned with PETSc's GAMG (examples/options/mg_cheby.cfg:
// -mg_coarse_ksp_type cg, -mg_coarse_pc_type gamg), a smoothed-aggregation AMG; this is that method
// built for the GPU:
//   set-up (host, FP64): level 0 = the operator's assembled block rows (one node per element);
//     aggregates of a node and its unaggregated neighbours (then leftovers join a neighbour's);
//     tentative prolongator from the rigid-body modes at the element's nodes, orthonormalised per
//     aggregate (QR; R is the next level's near-null space); prolongator smoothed once,
//     P = (I - 4/(3 rho) D^-1 A) P_tent with rho = lambda_max(D^-1 A) by power iteration (D the
//     block diagonal); coarse operator P^T A P; down to at most coarse_dofs unknowns, solved by a
//     dense inverse (an operator that small from the start is inverted directly);
//   V-cycle (device, FP32 or FP64, ncols columns at once): Chebyshev smoothing on D^-1 A over
//     [0.1, 1.1] x rho (GAMG's default interval), the same polynomial before and after the coarse
//     correction, restriction P^T: a fixed symmetric positive map, so plain CG may use it.
// Level 0 multiplies with the operator itself (its assembled apply) and its block-Jacobi inverses.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <vector>

namespace tdg {

class Operator;

struct AMGOptions {
    int smooth = 2;             // Chebyshev steps before and after the coarse correction
    double eig_lo = 0.1;        // Chebyshev interval [eig_lo, eig_hi] x lambda_max(D^-1 A)
    double eig_hi = 1.1;
    int64_t coarse_dofs = 800;  // the coarsest level's size at most (dense inverse)
    int max_levels = 10;
};

struct AMGLevelDev;

class AMG {
public:
    // op: elasticity with a nodal basis and assembled block rows (built here if needed); runs on
    // op's stream.
    AMG(Operator& op, AMGOptions const& o = AMGOptions{});
    ~AMG();
    AMG(AMG const&) = delete;
    AMG& operator=(AMG const&) = delete;

    // Z = M^-1 R for ncols columns of the operator's size at stride ld (one V-cycle, in the
    // precision of the vectors: FP32 operators and transfers, or FP64).
    void vcycle(float const* r, float* z, int ncols, int64_t ld);
    void vcycle(double const* r, double* z, int ncols, int64_t ld);
    int num_levels() const { return int(dofs_.size()); }
    std::vector<int64_t> const& level_dofs() const { return dofs_; }
    // sum over levels of the stored block entries / the level-0 entries
    double operator_complexity() const { return complexity_; }

private:
    template <class T> void run(T const* r, T* z, int ncols, int64_t ld);
    template <class T> void smooth(int l, T const* b, int64_t ldb, T* x, int64_t ldx, int ncols, bool zero_guess);
    template <class T>
    void residual(int l, T const* b, int64_t ldb, T const* x, int64_t ldx, T* r, int64_t ldr, int ncols);
    template <class T> void cycle(int l, T const* b, int64_t ldb, T* x, int64_t ldx, int ncols);

    Operator& op_;
    AMGOptions o_;
    std::vector<std::unique_ptr<AMGLevelDev>> lv_;
    std::vector<int64_t> dofs_;
    std::vector<double> rho_;
    double complexity_ = 1.0;
    int cap_cols_[2] = {0, 0}; // per precision (FP32, FP64)
    int64_t cap_ld_[2] = {0, 0};
};

} // namespace tdg
