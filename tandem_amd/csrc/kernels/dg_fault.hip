// Fault adapter on the GPU (FP64): the projection of facet-point traction onto the fault basis
// (evaluate_traction, app/kernels/elasticity_adapter.py:25-27, called from
// app/localoperator/ElasticityAdapter.cpp:21-34) and the evaluation of a fault-basis slip at the
// facet points (evaluate_slip, elasticity_adapter.py:21-23; ElasticityAdapter.cpp:37-49).  Lane per
// output entry; the per-facet tables are a few hundred doubles and stay in L1/L2.
#include "kernels.h"

#include <algorithm>

namespace tdg {

namespace {

constexpr int kT = 256;

// T(k, p) = sum_l minv(l, k) sum_q e(l, q) w_q nl_q sum_o t(o, q) bas(o, p, q)
__global__ __launch_bounds__(kT) void fault_project_kernel(FaultArgs a, double const* __restrict__ tqs,
                                                           double* __restrict__ Ts, int64_t ld_q, int64_t ld_t) {
    const int D = a.dim, nbf = a.nbf, nq = a.nq;
    const int64_t per = int64_t(nbf) * D;
    const int64_t item = int64_t(blockIdx.x) * kT + threadIdx.x;
    if (item >= a.nfault * per) return;
    const int64_t m = item / per;
    const int r = int(item % per), k = r % nbf, p = r / nbf;
    double const* tq = tqs + int64_t(blockIdx.y) * ld_q + m * D * nq;
    double const* mi = a.minv + m * nbf * nbf;
    double acc = 0.0;
    for (int q = 0; q < nq; ++q) {
        double const* b = a.bas + ((m * nq + q) * D) * D;
        double tb = 0.0;
        for (int o = 0; o < D; ++o) tb += tq[o * nq + q] * b[o * D + p];
        double el = 0.0; // sum_l minv(l, k) e(l, q)
        for (int l = 0; l < nbf; ++l) el += mi[l + k * nbf] * a.e_q[q * nbf + l];
        acc += el * a.w[q] * a.nl[m * nq + q] * tb;
    }
    Ts[int64_t(blockIdx.y) * ld_t + m * per + k + p * nbf] = acc;
}

// F(f, p, q) = sum_l e(l, q) sum_n bas(p, n + 1, q) S(l, n) on fault facets, 0 elsewhere
__global__ __launch_bounds__(kT) void fault_slip_kernel(FaultArgs a, double const* __restrict__ Ss,
                                                        double* __restrict__ Fs, int64_t ld_s, int64_t ld_f) {
    const int D = a.dim, nbf = a.nbf, nq = a.nq;
    const int64_t per = int64_t(D) * nq;
    const int64_t item = int64_t(blockIdx.x) * kT + threadIdx.x;
    if (item >= a.num_facets * per) return;
    const int64_t f = item / per;
    const int r = int(item % per), p = r / nq, q = r % nq;
    const int64_t m = a.fmap[f];
    double v = 0.0;
    if (m >= 0) {
        double const* S = Ss + int64_t(blockIdx.y) * ld_s + m * nbf * (D - 1);
        double const* b = a.bas + ((m * nq + q) * D) * D;
        for (int n = 0; n + 1 < D; ++n) {
            double s = 0.0;
            for (int l = 0; l < nbf; ++l) s += a.e_q[q * nbf + l] * S[l + n * nbf];
            v += b[p * D + n + 1] * s;
        }
    }
    Fs[int64_t(blockIdx.y) * ld_f + item] = v;
}

unsigned blocks_for(int64_t n) { return unsigned((n + kT - 1) / kT); }

} // namespace

hipError_t launch_fault_project(FaultArgs const& a, double const* tq, double* T, int ncols, int64_t ld_q,
                                int64_t ld_t, hipStream_t s) {
    const int64_t items = a.nfault * a.nbf * a.dim;
    if (items == 0 || ncols < 1) return hipSuccess;
    hipLaunchKernelGGL(fault_project_kernel, dim3(blocks_for(items), unsigned(ncols)), dim3(kT), 0, s, a, tq, T, ld_q,
                       ld_t);
    return hipGetLastError();
}

hipError_t launch_fault_slip(FaultArgs const& a, double const* S, double* F, int ncols, int64_t ld_s, int64_t ld_f,
                             hipStream_t s) {
    const int64_t items = a.num_facets * a.dim * a.nq;
    if (items == 0 || ncols < 1) return hipSuccess;
    hipLaunchKernelGGL(fault_slip_kernel, dim3(blocks_for(items), unsigned(ncols)), dim3(kT), 0, s, a, S, F, ld_s,
                       ld_f);
    return hipGetLastError();
}

} // namespace tdg
