// Element-block Jacobi preconditioner of the DG operator: the inverse of every element's diagonal
// block A_KK (bs x bs), in FP64 and FP32; blocks above 12 rows are stored as packed lower triangles
// (the inverses are exactly symmetric), half the bytes per application.
//
// The blocks are probed with the matrix-free apply on a colouring of the elements (no two elements
// sharing a facet have the same colour): with x = e_j on every element of one colour and zero
// elsewhere, y restricted to those elements is column j of their diagonal blocks.  This is the
// element-diagonal part of tandem's assembled matrix (assemble_volume / assemble_skeleton /
// assemble_boundary, Elasticity.cpp:317-519, Poisson.cpp:225-381; PetscDGMatrix.cpp:14-77) that a
// PCBJACOBI with one block per element uses.
#pragma once

#include "../kernels/kernels.h"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace tdg {

class Operator;

class BlockJacobi {
public:
    BlockJacobi() = default;
    BlockJacobi(BlockJacobi const&) = delete;
    BlockJacobi& operator=(BlockJacobi const&) = delete;
    ~BlockJacobi();

    // Probes, inverts and stores the blocks (FP64 and FP32 copies) on op's stream.
    void build(Operator& op);
    bool built() const { return bs_ > 0; }
    int block_size() const { return bs_; }
    int colours() const { return colours_; }

    // z = D^-1 r on device vectors (element-blocked), asynchronous on stream s; ncols columns with
    // leading dimension ld.
    void apply(double const* r, double* z, hipStream_t s, int ncols = 1, int64_t ld = 0) const;
    void apply(float const* r, float* z, hipStream_t s, int ncols = 1, int64_t ld = 0) const;
    // FP64 vectors and arithmetic with the FP32-stored blocks (half the bytes of the FP64 blocks;
    // the same fixed symmetric preconditioner for an FP64 solve as the FP32 one applies)
    void apply_fp32_blocks(double const* r, double* z, hipStream_t s, int ncols = 1, int64_t ld = 0) const;

    // Small blocks only (returns false otherwise, having done nothing): z = D^-1 r and the
    // per-column partial sums of r . z in partial[c kDotParts + i], i < kDotParts.
    bool apply_dot(float const* r, float* z, hipStream_t s, int ncols, int64_t ld, double* partial) const;
    bool apply_dot(double const* r, double* z, hipStream_t s, int ncols, int64_t ld, double* partial,
                   bool fp32_blocks) const;

    // FP64 inverse block of element e (bs x bs, column-major) copied to the host, for checks.
    void inverse_block(int64_t e, double* out) const;

private:
    void release();
    int bs_ = 0, colours_ = 0;
    int64_t E_ = 0, tp_ = 0;
    double* inv64_ = nullptr; // [E][bs][bs] column-major blocks of at most 12 rows
    float* inv32_ = nullptr;  //   and their FP32 copies
    double* p64_ = nullptr;   // [E][tp] packed lower triangles of larger blocks (exactly symmetric)
    float* p32_ = nullptr;    //   and their FP32 copies
};

// The preconditioner as a solver applies it: FP32 vectors always use the FP32 blocks; FP64 vectors
// use the FP64 blocks, or the FP32-stored blocks with FP64 arithmetic (fp32_blocks).
struct BlockJacobiApply {
    BlockJacobi const* bj = nullptr;
    bool fp32_blocks = false;
    void apply(double const* r, double* z, hipStream_t s, int ncols = 1, int64_t ld = 0) const {
        if (fp32_blocks) bj->apply_fp32_blocks(r, z, s, ncols, ld);
        else bj->apply(r, z, s, ncols, ld);
    }
    void apply(float const* r, float* z, hipStream_t s, int ncols = 1, int64_t ld = 0) const {
        bj->apply(r, z, s, ncols, ld);
    }
    // z = D^-1 r with the partial sums of r . z fused in (small blocks; false: nothing done)
    bool apply_dot(double const* r, double* z, hipStream_t s, int ncols, int64_t ld, double* partial) const {
        return bj->apply_dot(r, z, s, ncols, ld, partial, fp32_blocks);
    }
    bool apply_dot(float const* r, float* z, hipStream_t s, int ncols, int64_t ld, double* partial) const {
        return bj->apply_dot(r, z, s, ncols, ld, partial);
    }
};

// Greedy colouring of the elements over their facet neighbours (colour per element).
std::vector<int32_t> colour_elements(Operator const& op, int* ncolours);

} // namespace tdg
