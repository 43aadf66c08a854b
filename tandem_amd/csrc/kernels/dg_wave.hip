// Elastic-wave right-hand side pieces and p-multigrid transfer on the GPU (FP64):
//   inverse mass (apply_inverse_mass, elasticity.py:236-238): y_e = Mhat^-1 E diag(w_e) E^T Mhat^-1 x_e
//     with w_e(q) = -W_q / |J_e(q)| (the weight-adjusted inverse of the curvilinear element mass
//     with density 1, Elasticity.cpp:227-253); one workgroup per element, stages in LDS;
//   nodal interpolation (NodalInterpolation.cpp:8-38): y_e = I x_e per quantity (prolongation
//     from degree `from` to degree `to`) or I^T (restriction);
//   axpy for assembling apply_<true> = A x - b_facet.
#include "kernels.h"

#include <algorithm>

namespace tdg {

namespace {

constexpr int kT = 128;

__global__ __launch_bounds__(kT) void inverse_mass_kernel(WaveArgs a, double const* __restrict__ x,
                                                          double* __restrict__ y) {
    extern __shared__ double sh[];
    const int Nbf = a.nbf, Nq = a.nq, Q = a.nquant, bs = Nbf * Q;
    double* t = sh;              // bs
    double* uq = sh + bs;        // Nq * Q
    const int64_t e = blockIdx.x;
    double const* xe = x + e * bs;
    for (int i = threadIdx.x; i < bs; i += kT) {
        const int k = i % Nbf, p = i / Nbf;
        double v = 0.0;
        for (int l = 0; l < Nbf; ++l) v += a.Minv[k + l * Nbf] * xe[l + p * Nbf];
        t[i] = v;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < Nq * Q; i += kT) {
        const int q = i % Nq, p = i / Nq;
        double v = 0.0;
        for (int k = 0; k < Nbf; ++k) v += a.E[k * Nq + q] * t[k + p * Nbf];
        uq[i] = v * a.w[e * Nq + q];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < bs; i += kT) {
        const int k = i % Nbf, p = i / Nbf;
        double v = 0.0;
        for (int q = 0; q < Nq; ++q) v += a.E[k * Nq + q] * uq[q + p * Nq];
        t[i] = v; // stage 3 reads only uq
    }
    __syncthreads();
    for (int i = threadIdx.x; i < bs; i += kT) {
        const int k = i % Nbf, p = i / Nbf;
        double v = 0.0;
        for (int l = 0; l < Nbf; ++l) v += a.Minv[k + l * Nbf] * t[l + p * Nbf];
        y[e * bs + i] = v;
    }
}

__global__ __launch_bounds__(256) void interpolate_kernel(double const* __restrict__ I, int n_to, int n_from, int nquant,
                                                          int64_t nelem, int transpose, double const* __restrict__ x,
                                                          double* __restrict__ y) {
    const int n_out = transpose ? n_from : n_to, n_in = transpose ? n_to : n_from;
    const int64_t total = nelem * n_out * nquant;
    for (int64_t t = int64_t(blockIdx.x) * 256 + threadIdx.x; t < total; t += int64_t(gridDim.x) * 256) {
        const int64_t e = t / (int64_t(n_out) * nquant);
        const int r = int(t % (int64_t(n_out) * nquant)), p = r / n_out, k = r % n_out;
        double const* xe = x + e * int64_t(n_in) * nquant + p * n_in;
        double v = 0.0;
        if (transpose)
            for (int l = 0; l < n_in; ++l) v += I[l * n_from + k] * xe[l];
        else
            for (int l = 0; l < n_in; ++l) v += I[k * n_from + l] * xe[l];
        y[e * int64_t(n_out) * nquant + p * n_out + k] = v;
    }
}

__global__ __launch_bounds__(256) void axpy_kernel(double* __restrict__ y, double alpha, double const* __restrict__ x,
                                                   int64_t n) {
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) y[i] += alpha * x[i];
}

unsigned grid256(int64_t n) { return unsigned(std::min<int64_t>((n + 255) / 256, 65536)); }

} // namespace

hipError_t launch_inverse_mass(WaveArgs const& a, double const* x, double* y, hipStream_t s) {
    const std::size_t shm = std::size_t(a.nbf * a.nquant + a.nq * a.nquant) * sizeof(double);
    if (shm > 64 * 1024) return hipErrorInvalidValue;
    hipLaunchKernelGGL(inverse_mass_kernel, dim3(unsigned(a.num_elements)), dim3(kT), shm, s, a, x, y);
    return hipGetLastError();
}

hipError_t launch_interpolate(double const* I, int n_to, int n_from, int nquant, int64_t nelem, int transpose,
                              double const* x, double* y, hipStream_t s) {
    const int64_t total = nelem * (transpose ? n_from : n_to) * nquant;
    hipLaunchKernelGGL(interpolate_kernel, dim3(grid256(total)), dim3(256), 0, s, I, n_to, n_from, nquant, nelem,
                       transpose, x, y);
    return hipGetLastError();
}

hipError_t launch_axpy(double* y, double alpha, double const* x, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(axpy_kernel, dim3(grid256(n)), dim3(256), 0, s, y, alpha, x, n);
    return hipGetLastError();
}

} // namespace tdg
