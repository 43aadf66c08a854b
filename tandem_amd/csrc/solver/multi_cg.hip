// Multi-right-hand-side conjugate gradients: ncols independent CG recursions advanced in lockstep,
// so every iteration costs one fused (batched) operator apply over all columns.  Used for the
// discrete Green's function, whose right-hand sides are the slip basis functions of the fault
// (tandem's SeasQDDiscreteGreenOperator solves them one at a time, app/common/
// SeasQDDiscreteGreenOperator.cpp:540-549).
//
// Per column: its own scalars (rho, p.Ap, r.r), its own convergence (a converged column is frozen:
// its step lengths are zero), and in the FP32 mode its own FP64 residual replacements (the
// reliable-update scheme of the single-vector solver, applied to the columns whose residual
// dropped by reliable_delta at a check).  Reductions are two-stage with a fixed grid (bitwise
// reproducible).  Vectors are column-major (column c at offset c ld, one batched apply per
// iteration) or, for independent meshes held in one operator, segments of one vector (segment c at
// off[c] with len[c] entries, one apply of the whole operator per iteration).
#include "krylov.h"

#include "../host/operator.h"
#include "block_jacobi.h"
#include "multi_kernels.h"
#include "progress.h"
#include "replay.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

namespace tdg {

// The lockstep recursions over ncols columns laid out by L (tot entries per vector).  Segments:
// one apply of the whole operator per iteration; columns: one batched apply.
MultiSolveResult lockstep_cg(Operator& op, double const* B, double* X, int ncols, Cols L, std::size_t tot,
                             bool segments, SolveOptions const& o, MultiWork* w, CallerPrecond const* pre) {
    if (o.mode != Solve_FP64_CG && o.mode != Solve_FP32_CG_FP64_RELIABLE)
        throw std::invalid_argument("multi-column solves support FP64 CG and FP32 CG with FP64 replacement");
    if (pre && (segments || o.precond != Precond_None ||
                (o.mode == Solve_FP64_CG ? !pre->f64 : !pre->f32)))
        throw std::invalid_argument("a caller's preconditioner takes no segments, no other preconditioner and "
                                    "a function for the mode's precision");
    op.upload_device();
    op.reserve_columns(segments ? 1 : ncols); // workspace of the batched apply, outside any capture
    BlockJacobiApply bja;
    BlockJacobiApply const* M = nullptr;
    if (o.precond == Precond_BlockJacobi || o.precond == Precond_BlockJacobiF32) {
        bja.bj = &op.block_jacobi();
        bja.fp32_blocks = o.precond == Precond_BlockJacobiF32;
        M = &bja;
    } else if (o.precond != Precond_None) {
        throw std::invalid_argument("unknown preconditioner");
    }
    hipStream_t st = op.stream();
    const int64_t ld = L.ld;
    const dim3 g(kG, ncols), bl(kB);
    auto applyA = [&](auto const* P, auto* Q) {
        if (segments) op.apply(P, Q);
        else op.apply_batched(P, Q, ncols, ld);
    };
    auto applyM = [&](auto const* R, auto* Z) {
        if (segments) M->apply(R, Z, st, 1, 0);
        else M->apply(R, Z, st, ncols, ld);
    };
    MultiWork local;
    MultiWork& W = w ? *w : local;
    DBuf<double>& S = W.S;
    DBuf<double>& partial = W.partial;
    static_assert(kDotParts == kG, "the fused kernels write the solver's partial-sum layout");
    // q = A p and the partial sums of p . q: one kernel on the assembled apply
    auto applyA_dot = [&](auto const* P, auto* Q) {
        if (segments || !op.apply_batched_dot(P, Q, ncols, ld, partial.p)) {
            applyA(P, Q);
            hipLaunchKernelGGL((mdot<std::remove_const_t<std::remove_pointer_t<decltype(P)>>,
                                     std::remove_pointer_t<decltype(Q)>>),
                               dim3(kG, ncols), dim3(kB), 0, st, P, Q, L, partial.p);
        }
    };
    // z = M^-1 r and the partial sums of r . z: one fused kernel for small blocks
    auto applyM_dot = [&](auto const* R, auto* Z) {
        if (segments || !M->apply_dot(R, Z, st, ncols, ld, partial.p)) {
            applyM(R, Z);
            hipLaunchKernelGGL((mdot<std::remove_const_t<std::remove_pointer_t<decltype(R)>>,
                                     std::remove_pointer_t<decltype(Z)>>),
                               dim3(kG, ncols), dim3(kB), 0, st, R, Z, L, partial.p);
        }
    };
    S.ensure(std::size_t(ncols) * kS);
    partial.ensure(std::size_t(ncols) * kG);
    std::vector<double> hs(std::size_t(ncols) * kS, 0.0);
    MultiSolveResult res;
    res.rel_residual.assign(ncols, 0.0);
    res.iterations.assign(ncols, 0);
    MHIP(hipStreamSynchronize(st));
    auto t0 = std::chrono::steady_clock::now();
    Progress progress(segments ? "segmented CG" : "multi-column CG");
    // largest relative recursive residual over the active systems
    auto worst = [&] {
        double w = 0.0;
        for (int c = 0; c < ncols; ++c)
            if (hs[c * kS + 5] != 0.0 && hs[c * kS + 3] > 0.0) w = std::max(w, std::sqrt(hs[c * kS + 4] / hs[c * kS + 3]));
        return w;
    };
    auto pull = [&] {
        MHIP(hipMemcpyAsync(hs.data(), S.p, hs.size() * sizeof(double), hipMemcpyDeviceToHost, st));
        MHIP(hipStreamSynchronize(st));
    };
    auto push = [&] {
        MHIP(hipMemcpyAsync(S.p, hs.data(), hs.size() * sizeof(double), hipMemcpyHostToDevice, st));
    };
    // b.b per column, all columns active
    for (int c = 0; c < ncols; ++c) hs[c * kS + 5] = 1.0;
    push();
    hipLaunchKernelGGL((mdot<double, double>), g, bl, 0, st, B, B, L, partial.p);
    hipLaunchKernelGGL(mfinal, dim3(ncols), bl, 0, st, partial.p, S.p, 3, -1);
    MHIP(hipGetLastError());
    pull();
    std::vector<double> target(ncols);
    for (int c = 0; c < ncols; ++c) {
        target[c] = o.rtol * o.rtol * hs[c * kS + 3];
        if (hs[c * kS + 3] == 0.0) hs[c * kS + 5] = 0.0;
    }
    hipLaunchKernelGGL((mfill<double>), g, bl, 0, st, X, 0.0, L);

    if (o.mode == Solve_FP64_CG) {
        DBuf<double>&r = W.r, &p = W.p, &q = W.q, &z = W.z;
        r.ensure(tot);
        p.ensure(tot);
        q.ensure(tot);
        if (M || pre) z.ensure(tot);
        push();
        hipLaunchKernelGGL((mcopy<double>), g, bl, 0, st, r.p, B, L);
        hipLaunchKernelGGL((mdot<double, double>), g, bl, 0, st, r.p, r.p, L, partial.p);
        hipLaunchKernelGGL(mfinal, dim3(ncols), bl, 0, st, partial.p, S.p, 4, -1);
        auto prec = [&](int slot) {
            if (pre) {
                pre->f64(r.p, z.p);
                hipLaunchKernelGGL((mdot<double, double>), g, bl, 0, st, r.p, z.p, L, partial.p);
            } else if (M) {
                applyM_dot(r.p, z.p);
            } else {
                hipLaunchKernelGGL((mdot<double, double>), g, bl, 0, st, r.p, r.p, L, partial.p);
            }
            hipLaunchKernelGGL(mfinal, dim3(ncols), bl, 0, st, partial.p, S.p, slot, -1);
        };
        prec(0);
        hipLaunchKernelGGL((mcopy<double>), g, bl, 0, st, p.p, (M || pre) ? z.p : r.p, L);
        MHIP(hipGetLastError());
        auto iteration = [&](int c) {
            applyA_dot(p.p, q.p);
            hipLaunchKernelGGL(mfinal, dim3(ncols), bl, 0, st, partial.p, S.p, 2, -1);
            hipLaunchKernelGGL((mxr<double, double>), g, bl, 0, st, X, r.p, p.p, q.p, L, S.p, c, partial.p);
            hipLaunchKernelGGL(mfinal, dim3(ncols), bl, 0, st, partial.p, S.p, 4, -1);
            prec(1 - c);
            hipLaunchKernelGGL((mp<double>), g, bl, 0, st, p.p, (M || pre) ? z.p : r.p, L, S.p, 1 - c, c);
        };
        // the check_every iterations between host checks replayed from one captured graph (an even
        // count keeps the rho slots in phase)
        Replay chunk(st, o.graphs != 0 && o.check_every % 2 == 0);
        int it = 0, cur = 0;
        while (it < o.max_iter) {
            if (it % o.check_every == 0 && it + o.check_every <= o.max_iter) {
                chunk.run([&] {
                    for (int k = 0; k < o.check_every; ++k) iteration((cur + k) & 1);
                });
                it += o.check_every;
                cur = (cur + o.check_every) & 1;
            } else {
                iteration(cur);
                cur = 1 - cur;
                ++it;
            }
            if (it % o.check_every == 0 || it == o.max_iter) {
                MHIP(hipGetLastError());
                pull();
                int active = 0;
                for (int c = 0; c < ncols; ++c) {
                    if (hs[c * kS + 5] == 0.0) continue;
                    const double rr = hs[c * kS + 4];
                    if (!std::isfinite(rr)) throw std::runtime_error("CG breakdown (non-finite residual)");
                    res.iterations[c] = it;
                    if (rr <= target[c]) hs[c * kS + 5] = 0.0;
                    else ++active;
                }
                push();
                if (!active) break;
                progress.tick(it, active, worst());
            }
        }
        res.total_iterations = it;
        // true residuals
        applyA(X, q.p);
        hipLaunchKernelGGL(mresidual, g, bl, 0, st, B, q.p, r.p, (float*)nullptr, L, S.p, 1, partial.p);
        hipLaunchKernelGGL(mfinal, dim3(ncols), bl, 0, st, partial.p, S.p, 4, -1);
        MHIP(hipGetLastError());
        pull();
    } else {
        DBuf<double>&r = W.r, &Ax = W.Ax;
        DBuf<float>&r32 = W.r32, &xa = W.xa, &p = W.pf, &q = W.qf, &z = W.zf;
        r.ensure(tot);
        Ax.ensure(tot);
        r32.ensure(tot);
        xa.ensure(tot);
        p.ensure(tot);
        q.ensure(tot);
        if (M || pre) z.ensure(tot);
        push();
        hipLaunchKernelGGL((mfill<float>), g, bl, 0, st, xa.p, 0.0f, L);
        hipLaunchKernelGGL((mfill<double>), g, bl, 0, st, Ax.p, 0.0, L);
        hipLaunchKernelGGL(mresidual, g, bl, 0, st, B, Ax.p, r.p, r32.p, L, S.p, 1, partial.p);
        hipLaunchKernelGGL(mfinal, dim3(ncols), bl, 0, st, partial.p, S.p, 4, -1);
        // rho into slot `slot` of the columns flagged in slot msk (all if msk < 0)
        auto prec = [&](int slot, int msk) {
            if (pre) {
                pre->f32(r32.p, z.p);
                hipLaunchKernelGGL((mdot<float, float>), g, bl, 0, st, r32.p, z.p, L, partial.p);
            } else if (M) {
                applyM_dot(r32.p, z.p);
            } else {
                hipLaunchKernelGGL((mdot<float, float>), g, bl, 0, st, r32.p, r32.p, L, partial.p);
            }
            hipLaunchKernelGGL(mfinal, dim3(ncols), bl, 0, st, partial.p, S.p, slot, msk);
        };
        prec(0, -1);
        hipLaunchKernelGGL((mcopy<float>), g, bl, 0, st, p.p, (M || pre) ? z.p : r32.p, L);
        MHIP(hipGetLastError());
        pull();
        std::vector<double> maxr(ncols);
        for (int c = 0; c < ncols; ++c) maxr[c] = std::sqrt(hs[c * kS + 4]);
        int it = 0, cur = 0;
        res.replacements = 0;
        // first half of an iteration: q = A p, step, residual update; second half: preconditioner
        // and new direction (after the host check of the iteration, which may replace residuals)
        auto half1 = [&](int c) {
            applyA_dot(p.p, q.p);
            hipLaunchKernelGGL(mfinal, dim3(ncols), bl, 0, st, partial.p, S.p, 2, -1);
            hipLaunchKernelGGL((mxr<float, float>), g, bl, 0, st, xa.p, r32.p, p.p, q.p, L, S.p, c, partial.p);
            hipLaunchKernelGGL(mfinal, dim3(ncols), bl, 0, st, partial.p, S.p, 4, -1);
        };
        auto half2 = [&](int c) {
            prec(1 - c, -1);
            hipLaunchKernelGGL((mp<float>), g, bl, 0, st, p.p, (M || pre) ? z.p : r32.p, L, S.p, 1 - c, c);
        };
        // between two checks: (half1 half2) x (check_every - 1), then half1, replayed from one
        // captured graph; the check's half2 follows it directly (an even count keeps the phase)
        Replay chunk(st, o.graphs != 0 && o.check_every % 2 == 0);
        while (it < o.max_iter) {
            if (it % o.check_every == 0 && it + o.check_every <= o.max_iter) {
                chunk.run([&] {
                    for (int k = 0; k < o.check_every; ++k) {
                        half1((cur + k) & 1);
                        if (k + 1 < o.check_every) half2((cur + k) & 1);
                    }
                });
                cur = (cur + o.check_every - 1) & 1;
                it += o.check_every;
            } else {
                half1(cur);
                ++it;
            }
            bool any_replace = false;
            int active = 0;
            if (it % o.check_every == 0 || it == o.max_iter) {
                MHIP(hipGetLastError());
                pull();
                for (int c = 0; c < ncols; ++c) {
                    hs[c * kS + 6] = 0.0;
                    if (hs[c * kS + 5] == 0.0) continue;
                    const double rr = hs[c * kS + 4];
                    if (!std::isfinite(rr)) throw std::runtime_error("CG breakdown (non-finite residual)");
                    res.iterations[c] = it;
                    if (std::sqrt(rr) < o.reliable_delta * maxr[c] || rr <= target[c] || it == o.max_iter) {
                        hs[c * kS + 6] = 1.0;
                        any_replace = true;
                    }
                }
                push();
            }
            if (any_replace) {
                // FP64 replacement of the flagged columns: x += xa, r = b - A x
                hipLaunchKernelGGL(mfold, g, bl, 0, st, X, xa.p, L, S.p);
                applyA(X, Ax.p);
                hipLaunchKernelGGL(mresidual, g, bl, 0, st, B, Ax.p, r.p, r32.p, L, S.p, 0, partial.p);
                hipLaunchKernelGGL(mfinal, dim3(ncols), bl, 0, st, partial.p, S.p, 4, 6);
                MHIP(hipGetLastError());
                pull();
                for (int c = 0; c < ncols; ++c) {
                    if (hs[c * kS + 6] == 0.0) continue;
                    ++res.replacements;
                    const double rr = hs[c * kS + 4];
                    maxr[c] = std::sqrt(rr);
                    if (rr <= target[c] || it == o.max_iter) hs[c * kS + 5] = 0.0;
                }
                push();
            }
            for (int c = 0; c < ncols; ++c)
                if (hs[c * kS + 5] != 0.0) ++active;
            if ((any_replace || it % o.check_every == 0) && !active) break;
            if (it % o.check_every == 0) progress.tick(it, active, worst());
            half2(cur);
            cur = 1 - cur;
        }
        res.total_iterations = it;
        // fold any remaining update and report the true residuals
        for (int c = 0; c < ncols; ++c) hs[c * kS + 6] = 1.0;
        push();
        hipLaunchKernelGGL(mfold, g, bl, 0, st, X, xa.p, L, S.p);
        applyA(X, Ax.p);
        hipLaunchKernelGGL(mresidual, g, bl, 0, st, B, Ax.p, r.p, (float*)nullptr, L, S.p, 1, partial.p);
        hipLaunchKernelGGL(mfinal, dim3(ncols), bl, 0, st, partial.p, S.p, 4, -1);
        MHIP(hipGetLastError());
        pull();
    }
    res.converged = 1;
    for (int c = 0; c < ncols; ++c) {
        const double bb = hs[c * kS + 3];
        res.rel_residual[c] = bb > 0.0 ? std::sqrt(hs[c * kS + 4] / bb) : 0.0;
        if (!(res.rel_residual[c] <= o.rtol)) res.converged = 0;
    }
    MHIP(hipStreamSynchronize(st));
    res.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return res;
}

MultiSolveResult solve_multi(Operator& op, double const* B, double* X, int ncols, int64_t ld, SolveOptions const& o) {
    if (!op.ready()) throw std::logic_error("set_material() must be called before solve");
    if (ncols < 1 || ld < int64_t(op.local_size())) throw std::invalid_argument("bad column layout");
    Cols L;
    L.ld = ld;
    L.n = int64_t(op.local_size());
    return lockstep_cg(op, B, X, ncols, L, std::size_t(ncols) * ld, false, o);
}

MultiSolveResult solve_segments(Operator& op, double const* b, double* x, int nseg, int64_t const* offsets,
                                SolveOptions const& o) {
    if (!op.ready()) throw std::logic_error("set_material() must be called before solve");
    const int64_t n = int64_t(op.local_size()), bs = int64_t(op.block_size());
    if (nseg < 1 || !offsets || offsets[0] != 0 || offsets[nseg] != n)
        throw std::invalid_argument("segment offsets must start at 0 and end at the vector size");
    for (int c = 0; c < nseg; ++c)
        if (offsets[c + 1] <= offsets[c] || offsets[c] % bs != 0)
            throw std::invalid_argument("segments must be non-empty ranges of whole elements");
    // the segments must be uncoupled: no interior facet joins elements of two segments
    std::vector<int32_t> seg(op.num_elements());
    for (int c = 0; c < nseg; ++c)
        for (int64_t e = offsets[c] / bs; e < offsets[c + 1] / bs; ++e) seg[std::size_t(e)] = c;
    for (auto const& f : op.topology().facets)
        if (seg[f.elem[0]] != seg[f.elem[1]])
            throw std::invalid_argument("a facet couples two segments: they are not independent meshes");
    std::vector<int64_t> off(offsets, offsets + nseg), len(nseg);
    for (int c = 0; c < nseg; ++c) len[c] = offsets[c + 1] - offsets[c];
    DBuf<int64_t> doff, dlen;
    doff.alloc(nseg);
    dlen.alloc(nseg);
    MHIP(hipMemcpy(doff.p, off.data(), nseg * sizeof(int64_t), hipMemcpyHostToDevice));
    MHIP(hipMemcpy(dlen.p, len.data(), nseg * sizeof(int64_t), hipMemcpyHostToDevice));
    Cols L;
    L.off = doff.p;
    L.len = dlen.p;
    return lockstep_cg(op, b, x, nseg, L, std::size_t(n), true, o);
}

} // namespace tdg
