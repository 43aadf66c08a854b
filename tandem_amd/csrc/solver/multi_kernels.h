// Kernels of the multi-column Krylov solvers: per-column dot products with a fixed two-stage
// reduction (bitwise reproducible), the CG vector updates with per-column scalars, copies and
// fills.  Vectors are column-major (column c at offset c ld) or segments of one vector (segment c at
// off[c] with len[c] entries).  Per-column scalars S[c][kS]: 0,1 rho; 2 p.Ap; 3 b.b; 4 r.r;
// 5 active; 6 replace; 7 z.Ap (flexible CG).
#pragma once

#include "krylov.h"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <stdexcept>
#include <string>

namespace tdg {

#define MHIP(call)                                                                                 \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess) throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e_)); \
    } while (0)

constexpr int kB = 256;
constexpr int kG = 256;  // partial sums per column
constexpr int kS = 8;    // scalars per column (see above)

// Where column c lives: c ld with n entries, or the device tables off/len (segments).
struct Cols {
    int64_t ld = 0, n = 0;
    int64_t const* off = nullptr;
    int64_t const* len = nullptr;
    __device__ int64_t base(int64_t c) const { return off ? off[c] : c * ld; }
    __device__ int64_t size(int64_t c) const { return off ? len[c] : n; }
};

template <class T> struct DBuf {
    T* p = nullptr;
    std::size_t n = 0;
    DBuf() = default;
    DBuf(DBuf const&) = delete;
    DBuf& operator=(DBuf const&) = delete;
    ~DBuf() {
        if (p) (void)hipFree(p);
    }
    void alloc(std::size_t m) {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        MHIP(hipMalloc(reinterpret_cast<void**>(&p), m * sizeof(T)));
        n = m;
    }
    // at least m entries (contents not kept when it grows)
    void ensure(std::size_t m) {
        if (n < m) alloc(m);
    }
};

// Reusable buffers of the lockstep CG (kept by callers that solve repeatedly).
struct MultiWork {
    DBuf<double> S, partial, r, p, q, z, Ax;
    DBuf<float> r32, xa, pf, qf, zf;
};

namespace {

__device__ inline double bsum(double v) {
    __shared__ double s[kB / 64];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) s[w] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < kB / 64; ++i) r += s[i];
    return r;
}

template <class A, class B>
__global__ __launch_bounds__(kB) void mdot(A const* __restrict__ a, B const* __restrict__ b, Cols L,
                                           double* __restrict__ partial) {
    const int64_t c = blockIdx.y, n = L.size(c);
    A const* ac = a + L.base(c);
    B const* bc = b + L.base(c);
    double v = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * kB + threadIdx.x; i < n; i += int64_t(gridDim.x) * kB)
        v += double(ac[i]) * double(bc[i]);
    v = bsum(v);
    if (threadIdx.x == 0) partial[c * kG + blockIdx.x] = v;
}

// S[c][slot] = sum of the column's partials (only where mask slot is set, if msk >= 0)
__global__ __launch_bounds__(kB) void mfinal(double const* __restrict__ partial, double* __restrict__ S, int slot,
                                             int msk) {
    const int c = blockIdx.x;
    double v = 0.0;
    for (int i = threadIdx.x; i < kG; i += kB) v += partial[c * kG + i];
    v = bsum(v);
    if (threadIdx.x == 0 && (msk < 0 || S[c * kS + msk] != 0.0)) S[c * kS + slot] = v;
}

// active columns: x += alpha p, r -= alpha q, alpha = rho / pAp; partial sums of r.r
template <class T, class TX>
__global__ __launch_bounds__(kB) void mxr(TX* __restrict__ x, T* __restrict__ r, T const* __restrict__ p,
                                          T const* __restrict__ q, Cols L, double const* __restrict__ S,
                                          int irho, double* __restrict__ partial) {
    const int64_t c = blockIdx.y, n = L.size(c), o = L.base(c);
    double const* s = S + c * kS;
    const bool act = s[5] != 0.0;
    const T a = act ? T(s[irho] / s[2]) : T(0);
    TX* xc = x + o;
    T* rc = r + o;
    T const* pc = p + o;
    T const* qc = q + o;
    double v = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * kB + threadIdx.x; i < n; i += int64_t(gridDim.x) * kB) {
        T ri = rc[i];
        if (act) {
            xc[i] += TX(a) * TX(pc[i]);
            ri = ri - a * qc[i];
            rc[i] = ri;
        }
        v += double(ri) * double(ri);
    }
    v = bsum(v);
    if (threadIdx.x == 0) partial[c * kG + blockIdx.x] = v;
}

// active columns: p = z + (rho_new / rho_old) p
template <class T>
__global__ __launch_bounds__(kB) void mp(T* __restrict__ p, T const* __restrict__ z, Cols L,
                                         double const* __restrict__ S, int inew, int iold) {
    const int64_t c = blockIdx.y, n = L.size(c);
    double const* s = S + c * kS;
    if (s[5] == 0.0) return;
    const T beta = T(s[inew] / s[iold]);
    T* pc = p + L.base(c);
    T const* zc = z + L.base(c);
    for (int64_t i = int64_t(blockIdx.x) * kB + threadIdx.x; i < n; i += int64_t(gridDim.x) * kB)
        pc[i] = zc[i] + beta * pc[i];
}

template <class T>
__global__ __launch_bounds__(kB) void mcopy(T* __restrict__ d, T const* __restrict__ s, Cols L) {
    const int64_t c = blockIdx.y, n = L.size(c), o = L.base(c);
    for (int64_t i = int64_t(blockIdx.x) * kB + threadIdx.x; i < n; i += int64_t(gridDim.x) * kB)
        d[o + i] = s[o + i];
}
template <class T>
__global__ __launch_bounds__(kB) void mfill(T* __restrict__ d, T v, Cols L) {
    const int64_t c = blockIdx.y, n = L.size(c), o = L.base(c);
    for (int64_t i = int64_t(blockIdx.x) * kB + threadIdx.x; i < n; i += int64_t(gridDim.x) * kB) d[o + i] = v;
}

// columns flagged for replacement (S[c][6]): x += xa (FP64), xa = 0
__global__ __launch_bounds__(kB) void mfold(double* __restrict__ x, float* __restrict__ xa, Cols L,
                                            double const* __restrict__ S) {
    const int64_t c = blockIdx.y, n = L.size(c), o = L.base(c);
    if (S[c * kS + 6] == 0.0) return;
    for (int64_t i = int64_t(blockIdx.x) * kB + threadIdx.x; i < n; i += int64_t(gridDim.x) * kB) {
        x[o + i] += double(xa[o + i]);
        xa[o + i] = 0.0f;
    }
}

// flagged columns (or all if all != 0): r = b - Ax in FP64, r32 = float(r); partial sums of r.r
__global__ __launch_bounds__(kB) void mresidual(double const* __restrict__ b, double const* __restrict__ Ax,
                                                double* __restrict__ r, float* __restrict__ r32, Cols L,
                                                double const* __restrict__ S, int all, double* __restrict__ partial) {
    const int64_t c = blockIdx.y, n = L.size(c), o = L.base(c);
    const bool on = all || S[c * kS + 6] != 0.0;
    double v = 0.0;
    if (on)
        for (int64_t i = int64_t(blockIdx.x) * kB + threadIdx.x; i < n; i += int64_t(gridDim.x) * kB) {
            const double ri = b[o + i] - Ax[o + i];
            r[o + i] = ri;
            if (r32) r32[o + i] = float(ri);
            v += ri * ri;
        }
    v = bsum(v);
    if (threadIdx.x == 0) partial[c * kG + blockIdx.x] = v;
}



// active columns: p = z + beta p with beta = -(z.Ap) / (p.Ap) (flexible CG, one truncation
// direction: p stays A-orthogonal to the previous direction for a variable preconditioner)
template <class T>
__global__ __launch_bounds__(kB) void mpflex(T* __restrict__ p, T const* __restrict__ z, Cols L,
                                             double const* __restrict__ S) {
    const int64_t c = blockIdx.y, n = L.size(c);
    double const* s = S + c * kS;
    if (s[5] == 0.0) return;
    const T beta = T(-s[7] / s[2]);
    T* pc = p + L.base(c);
    T const* zc = z + L.base(c);
    for (int64_t i = int64_t(blockIdx.x) * kB + threadIdx.x; i < n; i += int64_t(gridDim.x) * kB)
        pc[i] = zc[i] + beta * pc[i];
}

} // namespace

class Operator;
// A caller's preconditioner, z = M^-1 r on the solver's columns: f32 for the FP32 recursion of the
// FP32 mode, f64 for the FP64 mode.
struct CallerPrecond {
    std::function<void(float const* r, float* z)> f32;
    std::function<void(double const* r, double* z)> f64;
};
// ncols lockstep CG recursions over the columns laid out by L (tot entries per vector): one batched
// apply per iteration, or with segments one apply of the whole operator.  Buffers from w when given
// (grown as needed), else allocated for the call.  pre (no segments, in place of o.precond) is
// applied to the residuals of the mode's recursion.
MultiSolveResult lockstep_cg(Operator& op, double const* B, double* X, int ncols, Cols L, std::size_t tot,
                             bool segments, SolveOptions const& o, MultiWork* w = nullptr,
                             CallerPrecond const* pre = nullptr);

} // namespace tdg
