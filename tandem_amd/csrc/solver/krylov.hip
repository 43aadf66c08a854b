// Krylov solvers on the GPU around the matrix-free operator: FP64 conjugate gradients, and FP32
// conjugate gradients inside an FP64 iterative-refinement loop (true residual and update in FP64).
//
// This is the role PETSc's KSPCG plays under tandem's PetscLinearSolver::solve
// (app/common/PetscLinearSolver.h:42, PetscLinearSolver.cpp:22-32) with the MatShell operator.
// Scalars stay on the device between kernels; the host reads the residual norm every
// `check_every` iterations only.  Reductions are two-stage with a fixed grid, so results are
// bitwise reproducible run to run.
#include "krylov.h"
#include "progress.h"
#include "replay.h"

#include "../host/operator.h"
#include "block_jacobi.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <stdexcept>
#include <string>

namespace tdg {

namespace {

constexpr int kBlock = 256;
constexpr int kGrid = 1024; // partial sums per reduction

#define KHIP(call)                                                                                 \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess) throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e_)); \
    } while (0)

__device__ inline double block_sum(double v) {
    __shared__ double s[kBlock / 64];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) s[w] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < kBlock / 64; ++i) r += s[i];
    return r;
}

// partial[b] = sum over the block's grid-stride range of a[i] * b[i] (FP64 accumulation)
template <class T>
__global__ __launch_bounds__(kBlock) void dot_partial(T const* __restrict__ a, T const* __restrict__ b, int64_t n,
                                                     double* __restrict__ partial) {
    double v = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += int64_t(gridDim.x) * kBlock)
        v += double(a[i]) * double(b[i]);
    v = block_sum(v);
    if (threadIdx.x == 0) partial[blockIdx.x] = v;
}

// out = sum(partial[0..kGrid))
__global__ __launch_bounds__(kBlock) void finalize_sum(double const* __restrict__ partial, double* __restrict__ out) {
    double v = 0.0;
    for (int i = threadIdx.x; i < kGrid; i += kBlock) v += partial[i];
    v = block_sum(v);
    if (threadIdx.x == 0) *out = v;
}

// Sum of partial[0 .. np) in a fixed order (the same in every block), returned to every thread.
__device__ inline double block_total(double const* __restrict__ partial, int64_t np) {
    __shared__ double tot;
    double v = 0.0;
    for (int64_t i = threadIdx.x; i < np; i += kBlock) v += partial[i];
    v = block_sum(v);
    if (threadIdx.x == 0) tot = v;
    __syncthreads();
    return tot;
}

// alpha = s[rho] / sum(pq[0 .. npq));  x += alpha p;  r -= alpha q;  partial sums of r.r.  Every block
// reduces the p.q partials itself, so no separate reduction launch sits between the apply and this.
template <class T, class TX>
__global__ __launch_bounds__(kBlock) void cg_xr_fused(TX* __restrict__ x, T* __restrict__ r, T const* __restrict__ p,
                                                     T const* __restrict__ q, int64_t n, double const* __restrict__ s,
                                                     int irho, double const* __restrict__ pq, int64_t npq,
                                                     double* __restrict__ partial) {
    const double alpha = s[irho] / block_total(pq, npq);
    const T a = T(alpha);
    double v = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += int64_t(gridDim.x) * kBlock) {
        x[i] += TX(a) * TX(p[i]);
        T ri = r[i] - a * q[i];
        r[i] = ri;
        v += double(ri) * double(ri);
    }
    v = block_sum(v);
    if (threadIdx.x == 0) partial[blockIdx.x] = v;
}

// rho_new = sum(partial[0 .. np)), written to s[inew] (block 0);  p = zr + (rho_new / s[iold]) p.
template <class T>
__global__ __launch_bounds__(kBlock) void cg_p_fused(T* __restrict__ p, T const* __restrict__ zr, int64_t n,
                                                    double* __restrict__ s, int inew, int iold,
                                                    double const* __restrict__ partial, int64_t np) {
    const double rho = block_total(partial, np);
    if (blockIdx.x == 0 && threadIdx.x == 0) s[inew] = rho;
    const T beta = T(rho / s[iold]);
    for (int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += int64_t(gridDim.x) * kBlock)
        p[i] = zr[i] + beta * p[i];
}

// alpha = s[rr] / s[pq];  x += alpha p;  r -= alpha q;  partial sums of r.r
template <class T, class TX>
__global__ __launch_bounds__(kBlock) void cg_xr(TX* __restrict__ x, T* __restrict__ r, T const* __restrict__ p,
                                               T const* __restrict__ q, int64_t n, double const* __restrict__ s,
                                               int irr, int ipq, double* __restrict__ partial) {
    const double alpha = s[irr] / s[ipq];
    const T a = T(alpha);
    double v = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += int64_t(gridDim.x) * kBlock) {
        x[i] += TX(a) * TX(p[i]);
        T ri = r[i] - a * q[i];
        r[i] = ri;
        v += double(ri) * double(ri);
    }
    v = block_sum(v);
    if (threadIdx.x == 0) partial[blockIdx.x] = v;
}

// p = r + (s[rr_new] / s[rr_old]) p
template <class T>
__global__ __launch_bounds__(kBlock) void cg_p(T* __restrict__ p, T const* __restrict__ r, int64_t n,
                                              double const* __restrict__ s, int inew, int iold) {
    const T beta = T(s[inew] / s[iold]);
    for (int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += int64_t(gridDim.x) * kBlock)
        p[i] = r[i] + beta * p[i];
}

// r = b - Ax (FP64), r32 = float(r), partial sums of r.r
__global__ __launch_bounds__(kBlock) void residual(double const* __restrict__ b, double const* __restrict__ Ax,
                                                   double* __restrict__ r, float* __restrict__ r32, int64_t n,
                                                   double* __restrict__ partial) {
    double v = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += int64_t(gridDim.x) * kBlock) {
        double ri = b[i] - Ax[i];
        r[i] = ri;
        if (r32) r32[i] = float(ri);
        v += ri * ri;
    }
    v = block_sum(v);
    if (threadIdx.x == 0) partial[blockIdx.x] = v;
}

template <class T> __global__ __launch_bounds__(kBlock) void copy_to(T* __restrict__ d, T const* __restrict__ s, int64_t n) {
    for (int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += int64_t(gridDim.x) * kBlock) d[i] = s[i];
}
template <class T> __global__ __launch_bounds__(kBlock) void fill(T* __restrict__ d, T v, int64_t n) {
    for (int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += int64_t(gridDim.x) * kBlock) d[i] = v;
}
// x += float-to-double(d)
__global__ __launch_bounds__(kBlock) void add_update(double* __restrict__ x, float const* __restrict__ d, int64_t n) {
    for (int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += int64_t(gridDim.x) * kBlock) x[i] += double(d[i]);
}

template <class T> struct Buf {
    T* p = nullptr;
    ~Buf() {
        if (p) (void)hipFree(p);
    }
    void alloc(std::size_t n) { KHIP(hipMalloc(reinterpret_cast<void**>(&p), n * sizeof(T))); }
};

double read_scalar(double const* d, hipStream_t s) {
    double h = 0.0;
    KHIP(hipMemcpyAsync(&h, d, sizeof(double), hipMemcpyDeviceToHost, s));
    KHIP(hipStreamSynchronize(s));
    return h;
}

// (Preconditioned) CG on A z = rhs in precision T from z = 0; returns iterations.  s: device
// scalars [6]: s[0], s[1] alternate rho = r.(M r), s[2] = p.Ap, s[4] = r.r when preconditioned.
// M = null is plain CG (rho = r.r).  w: work vector for M r.
template <class T, class TX>
int cg(Operator& op, BlockJacobiApply const* M, T const* rhs, TX* z, T* r, T* p, T* q, T* w, double* s, double* partial,
       double* pqp, int64_t n, double target_rr, int max_it, int check_every, bool graphs, hipStream_t st,
       double* rr_out) {
    const dim3 g(kGrid), b(kBlock);
    hipLaunchKernelGGL((fill<TX>), g, b, 0, st, z, TX(0), n);
    hipLaunchKernelGGL((copy_to<T>), g, b, 0, st, r, rhs, n);
    hipLaunchKernelGGL((dot_partial<T>), g, b, 0, st, r, r, n, partial);
    hipLaunchKernelGGL(finalize_sum, dim3(1), b, 0, st, partial, s + (M ? 4 : 0));
    if (M) {
        M->apply(r, w, st);
        hipLaunchKernelGGL((copy_to<T>), g, b, 0, st, p, w, n);
        hipLaunchKernelGGL((dot_partial<T>), g, b, 0, st, r, w, n, partial);
        hipLaunchKernelGGL(finalize_sum, dim3(1), b, 0, st, partial, s + 0);
    } else {
        hipLaunchKernelGGL((copy_to<T>), g, b, 0, st, p, rhs, n);
    }
    KHIP(hipGetLastError());
    double rr = read_scalar(s + (M ? 4 : 0), st);
    if (rr <= target_rr) {
        *rr_out = rr;
        return 0;
    }
    // one iteration: q = A p with the partials of p.q from the apply's store (or a dot pass for the
    // kernels without that epilogue); x, r update with the r.r partials; [z = M r, its r.z partials];
    // p update, which reduces the rho partials itself and records rho in s for the next step
    auto iteration = [&](int c) {
        int64_t npq = op.apply_dot(p, q, pqp);
        if (npq == 0) {
            hipLaunchKernelGGL((dot_partial<T>), g, b, 0, st, p, q, n, pqp);
            npq = kGrid;
        }
        hipLaunchKernelGGL((cg_xr_fused<T, TX>), g, b, 0, st, z, r, p, q, n, s, c, pqp, npq, partial);
        if (M) {
            hipLaunchKernelGGL(finalize_sum, dim3(1), b, 0, st, partial, s + 4);
            M->apply(r, w, st);
            hipLaunchKernelGGL((dot_partial<T>), g, b, 0, st, r, w, n, partial);
        }
        hipLaunchKernelGGL((cg_p_fused<T>), g, b, 0, st, p, M ? w : r, n, s, 1 - c, c, partial, int64_t(kGrid));
    };
    // check_every iterations between host checks; an even count keeps the rho slots in phase, so
    // the same captured sequence is replayed for every chunk
    Replay chunk(st, graphs && check_every % 2 == 0);
    Progress progress(sizeof(T) == 8 ? "FP64 CG" : "FP32 CG");
    int it = 0, cur = 0;
    while (it < max_it) {
        if (it + check_every <= max_it) {
            chunk.run([&] {
                for (int k = 0; k < check_every; ++k) iteration((cur + k) & 1);
            });
            it += check_every;
            cur = (cur + check_every) & 1;
        } else {
            iteration(cur);
            cur = 1 - cur;
            ++it;
        }
        if (it % check_every == 0 || it == max_it) {
            KHIP(hipGetLastError());
            rr = read_scalar(s + (M ? 4 : cur), st);
            if (!std::isfinite(rr)) throw std::runtime_error("CG breakdown (non-finite residual)");
            if (rr <= target_rr) break;
            progress.tick(it, 1, std::sqrt(rr));
        }
    }
    *rr_out = read_scalar(s + (M ? 4 : cur), st);
    return it;
}

} // namespace

SolveResult solve(Operator& op, double const* b, double* x, SolveOptions const& o) {
    if (!op.ready()) throw std::logic_error("set_material() must be called before solve");
    if (!(o.rtol > 0.0) || o.max_iter <= 0 || o.check_every <= 0) throw std::invalid_argument("bad solver options");
    op.upload_device();
    hipStream_t st = op.stream();
    const int64_t n = int64_t(op.local_size());
    const dim3 g(kGrid), bl(kBlock);
    Buf<double> s, partial;
    s.alloc(6);
    BlockJacobiApply bja;
    BlockJacobiApply const* M = nullptr;
    if (o.precond == Precond_BlockJacobi || o.precond == Precond_BlockJacobiF32) {
        bja.bj = &op.block_jacobi();
        bja.fp32_blocks = o.precond == Precond_BlockJacobiF32;
        M = &bja;
    } else if (o.precond != Precond_None) {
        throw std::invalid_argument("unknown preconditioner");
    }
    partial.alloc(kGrid);
    Buf<double> pqp; // partial sums of p.q from the apply's store
    pqp.alloc(std::size_t(std::max<int64_t>(kGrid, op.dot_partials_capacity())));
    SolveResult res{};
    KHIP(hipStreamSynchronize(st));
    auto t0 = std::chrono::steady_clock::now();
    // ||b||
    hipLaunchKernelGGL((dot_partial<double>), g, bl, 0, st, b, b, n, partial.p);
    hipLaunchKernelGGL(finalize_sum, dim3(1), bl, 0, st, partial.p, s.p + 3);
    KHIP(hipGetLastError());
    const double bb = read_scalar(s.p + 3, st);
    if (bb == 0.0) {
        hipLaunchKernelGGL((fill<double>), g, bl, 0, st, x, 0.0, n);
        res.converged = 1;
        return res;
    }
    const double target = o.rtol * o.rtol * bb;
    if (o.mode == Solve_FP64_CG) {
        Buf<double> r, p, q, w;
        r.alloc(n);
        p.alloc(n);
        q.alloc(n);
        if (M) w.alloc(n);
        double rr = 0.0;
        res.iterations = cg<double, double>(op, M, b, x, r.p, p.p, q.p, w.p, s.p, partial.p, pqp.p, n, target,
                                            o.max_iter, o.check_every, o.graphs != 0, st, &rr);
        // report the true residual ||b - A x|| / ||b||
        op.apply(x, q.p);
        hipLaunchKernelGGL(residual, g, bl, 0, st, b, q.p, r.p, (float*)nullptr, n, partial.p);
        hipLaunchKernelGGL(finalize_sum, dim3(1), bl, 0, st, partial.p, s.p + 3);
        KHIP(hipGetLastError());
        res.rel_residual = std::sqrt(read_scalar(s.p + 3, st) / bb);
        res.outer_iterations = 1;
        res.inner_iterations = res.iterations;
    } else if (o.mode == Solve_FP32_CG_FP64_IR) {
        Buf<double> r, Ax;
        Buf<float> r32, d32, p32, q32, w32, z32;
        r.alloc(n);
        Ax.alloc(n);
        r32.alloc(n);
        d32.alloc(n);
        p32.alloc(n);
        q32.alloc(n);
        w32.alloc(n);
        if (M) z32.alloc(n);
        // x = 0, r = b, r32 = float(b)
        hipLaunchKernelGGL((fill<double>), g, bl, 0, st, x, 0.0, n);
        hipLaunchKernelGGL((fill<double>), g, bl, 0, st, Ax.p, 0.0, n);
        hipLaunchKernelGGL(residual, g, bl, 0, st, b, Ax.p, r.p, r32.p, n, partial.p);
        KHIP(hipGetLastError());
        double rr = bb;
        int outer = 0, inner_total = 0;
        while (rr > target && outer < o.max_outer && inner_total < o.max_iter) {
            // FP32 CG on A d = r to relative tolerance inner_rtol, then the FP64 update x += d
            double inner_target = o.inner_rtol * o.inner_rtol * rr;
            double rr_in = 0.0;
            int its = cg<float, float>(op, M, r32.p, d32.p, w32.p, p32.p, q32.p, z32.p, s.p, partial.p, pqp.p, n,
                                       inner_target, o.max_iter - inner_total, o.check_every, o.graphs != 0, st,
                                       &rr_in);
            inner_total += its;
            hipLaunchKernelGGL(add_update, g, bl, 0, st, x, d32.p, n);
            ++outer;
            // FP64 true residual r = b - A x
            op.apply(x, Ax.p);
            hipLaunchKernelGGL(residual, g, bl, 0, st, b, Ax.p, r.p, r32.p, n, partial.p);
            hipLaunchKernelGGL(finalize_sum, dim3(1), bl, 0, st, partial.p, s.p + 3);
            KHIP(hipGetLastError());
            rr = read_scalar(s.p + 3, st);
            if (!std::isfinite(rr)) throw std::runtime_error("iterative refinement diverged");
            if (its == 0) break;
        }
        res.rel_residual = std::sqrt(rr / bb);
        res.outer_iterations = outer;
        res.inner_iterations = inner_total;
        res.iterations = inner_total;
    } else if (o.mode == Solve_FP32_CG_FP64_RELIABLE) {
        // FP32 CG whose residual is replaced by the FP64 true residual b - A x (and the FP32 update
        // folded into the FP64 x) whenever it has dropped by `reliable_delta` since the last
        // replacement (reliable updates; Sleijpen & van der Vorst 1996, as used for mixed-precision
        // CG by Clark et al. 2010).  The search direction is kept across replacements, so the
        // iteration count stays close to FP64 CG while all operator applies but the replacements
        // run in FP32.
        Buf<double> r, Ax;
        // With a preconditioner M the recursion is preconditioned CG: rho = r.(M r) in s[cur],
        // r.r in s[4] for the replacement and stopping tests.
        Buf<float> r32, xa32, p32, q32, z32;
        r.alloc(n);
        Ax.alloc(n);
        r32.alloc(n);
        xa32.alloc(n);
        p32.alloc(n);
        q32.alloc(n);
        if (M) z32.alloc(n);
        auto precondition = [&](int irho) {
            M->apply(r32.p, z32.p, st);
            hipLaunchKernelGGL((dot_partial<float>), g, bl, 0, st, r32.p, z32.p, n, partial.p);
            hipLaunchKernelGGL(finalize_sum, dim3(1), bl, 0, st, partial.p, s.p + irho);
        };
        hipLaunchKernelGGL((fill<double>), g, bl, 0, st, x, 0.0, n);
        hipLaunchKernelGGL((fill<double>), g, bl, 0, st, Ax.p, 0.0, n);
        hipLaunchKernelGGL((fill<float>), g, bl, 0, st, xa32.p, 0.0f, n);
        hipLaunchKernelGGL(residual, g, bl, 0, st, b, Ax.p, r.p, r32.p, n, partial.p);
        hipLaunchKernelGGL(finalize_sum, dim3(1), bl, 0, st, partial.p, s.p + (M ? 4 : 0));
        if (M) precondition(0);
        hipLaunchKernelGGL((copy_to<float>), g, bl, 0, st, p32.p, M ? z32.p : r32.p, n);
        KHIP(hipGetLastError());
        double rr_true = read_scalar(s.p + (M ? 4 : 0), st);
        double max_r = std::sqrt(rr_true);
        // iteration = part A (apply, step, r.r) + part B (preconditioner, direction); the host
        // check sits between A and B, so a chunk between checks is A B A B ... A
        auto partA = [&](int c) {
            int64_t npq = op.apply_dot(p32.p, q32.p, pqp.p);
            if (npq == 0) {
                hipLaunchKernelGGL((dot_partial<float>), g, bl, 0, st, p32.p, q32.p, n, pqp.p);
                npq = kGrid;
            }
            hipLaunchKernelGGL((cg_xr_fused<float, float>), g, bl, 0, st, xa32.p, r32.p, p32.p, q32.p, n, s.p, c,
                               pqp.p, npq, partial.p);
            hipLaunchKernelGGL(finalize_sum, dim3(1), bl, 0, st, partial.p, s.p + (M ? 4 : 1 - c));
        };
        auto partB = [&](int c) {
            if (M) precondition(1 - c);
            hipLaunchKernelGGL((cg_p<float>), g, bl, 0, st, p32.p, M ? z32.p : r32.p, n, s.p, 1 - c, c);
        };
        Replay chunk(st, o.graphs != 0 && o.check_every % 2 == 0);
        int it = 0, cur = 0, updates = 0;
        while (it < o.max_iter) {
            if (it % o.check_every == 0 && it + o.check_every <= o.max_iter) {
                // iterations it+1 .. it+check_every up to the last one's part A
                chunk.run([&] {
                    for (int k = 0; k < o.check_every; ++k) {
                        partA((cur + k) & 1);
                        if (k + 1 < o.check_every) partB((cur + k) & 1);
                    }
                });
                cur = (cur + o.check_every - 1) & 1;
                it += o.check_every;
            } else {
                partA(cur);
                ++it;
            }
            const int irr = M ? 4 : 1 - cur; // where r.r of the last part A landed
            bool replace = false;
            if (it % o.check_every == 0 || it == o.max_iter) {
                KHIP(hipGetLastError());
                double rr_it = read_scalar(s.p + irr, st);
                if (!std::isfinite(rr_it)) throw std::runtime_error("CG breakdown (non-finite residual)");
                replace = std::sqrt(rr_it) < o.reliable_delta * max_r || rr_it <= target || it == o.max_iter;
            }
            if (replace) {
                hipLaunchKernelGGL(add_update, g, bl, 0, st, x, xa32.p, n);
                hipLaunchKernelGGL((fill<float>), g, bl, 0, st, xa32.p, 0.0f, n);
                op.apply(x, Ax.p);
                hipLaunchKernelGGL(residual, g, bl, 0, st, b, Ax.p, r.p, r32.p, n, partial.p);
                hipLaunchKernelGGL(finalize_sum, dim3(1), bl, 0, st, partial.p, s.p + irr);
                KHIP(hipGetLastError());
                rr_true = read_scalar(s.p + irr, st);
                max_r = std::sqrt(rr_true);
                ++updates;
                if (rr_true <= target) break;
            }
            partB(cur);
            cur = 1 - cur;
        }
        res.rel_residual = std::sqrt(rr_true / bb);
        res.iterations = it;
        res.inner_iterations = it;
        res.outer_iterations = updates;
    } else {
        throw std::invalid_argument("unknown solver mode");
    }
    KHIP(hipStreamSynchronize(st));
    res.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    res.converged = res.rel_residual <= o.rtol ? 1 : 0;
    return res;
}

} // namespace tdg
