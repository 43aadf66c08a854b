// Right-hand side and facet traction of the elasticity operator on the GPU (FP64).
//
// rhs      : b = rhsVolume + rhsFacet over the Dirichlet and Fault facets, one lane per element
//            and basis function (Elasticity.cpp:521-696; elasticity.py:151-180).  Facet data are
//            the Dirichlet values on Dirichlet facets and the slip on Fault facets; a Fault on the
//            boundary takes half the slip (bc_boundary, Elasticity.cpp:559-573).  A skeleton facet
//            adds to both of its elements with c10 = eps/2 and c20 = +penalty (side 0) /
//            -penalty (side 1), both with the side-0 normal (rhs_skeleton, :575-643); a boundary
//            facet with c10 = eps, c20 = penalty (rhs_boundary, :645-696).
// traction : per facet and facet point, 1/2 (traction_0 + traction_1) - penalty (u_0 - u_1 - f) on
//            skeleton facets and traction_0 - penalty (u_0 - f) on boundary facets, with the unit
//            normal of side 0 (compute_traction / compute_traction_bnd, elasticity.py:242-250;
//            Elasticity.cpp:949-1018).
// Both use the apply kernels' facet data: the face-0 reference table with the node relabelling
// perm_f and the per-side facet geometry G' = J_f^-1 J^-1.
#include "kernels.h"

#include <type_traits>

namespace tdg {

namespace {

constexpr int binom_(int n, int k) { return k == 0 ? 1 : binom_(n - 1, k - 1) * n / k; }
constexpr int vol_points_(int D, int N) {
    return D == 2 ? (N == 1 ? 4 : N == 2 ? 7 : N == 3 ? 15 : N == 4 ? 19 : N == 5 ? 28 : 37)
                  : (N == 1 ? 6 : N == 2 ? 14 : N == 3 ? 31 : 57);
}
constexpr int fct_points_(int D, int N) { return D == 2 ? N + 1 : (N == 1 ? 4 : N == 2 ? 7 : N == 3 ? 15 : 19); }

constexpr int kBC_Fault = 3, kBC_Dirichlet = 5;
constexpr int kThreads = 256;

template <int D, int N>
__global__ __launch_bounds__(kThreads) void elasticity_rhs_kernel(RhsArgs a, double const* __restrict__ force,
                                                                  double const* __restrict__ fdata,
                                                                  double* __restrict__ b, int64_t ld_force,
                                                                  int64_t ld_fdata, int64_t ld_b) {
    constexpr int Nbf = binom_(N + D, D), Nq = vol_points_(D, N), nq = fct_points_(D, N);
    constexpr int Q = D, BS = Nbf * Q, DD = D * D, NF = D + 1, C = D + 1;
    const int64_t item = int64_t(blockIdx.x) * kThreads + threadIdx.x;
    if (item >= a.num_elements * Nbf) return;
    const int64_t e = item / Nbf;
    const int k = int(item % Nbf);
    double const* fo = force ? force + blockIdx.y * ld_force : nullptr;
    double const* fd = fdata ? fdata + blockIdx.y * ld_fdata : nullptr;
    double acc[Q];
#pragma unroll
    for (int p = 0; p < Q; ++p) acc[p] = 0.0;
    if (fo) {
        for (int q = 0; q < Nq; ++q) {
            const double t = a.jw[e * Nq + q] * a.tabE[k * Nq + q];
#pragma unroll
            for (int p = 0; p < Q; ++p) acc[p] += t * fo[(e * Q + p) * Nq + q];
        }
    }
    if (fd) {
#pragma unroll 1
        for (int f = 0; f < NF; ++f) {
            const int4 info = a.ef[e * NF + f];
            const int64_t fct = info.x;
            const int fb = a.fbc[fct];
            const int bc = fb & 0xff;
            const bool skel = (fb >> 8) & 1;
            if (bc != kBC_Dirichlet && bc != kBC_Fault) continue;
            const int side = (info.z >> 2) & 1;
            const double scale = (!skel && bc == kBC_Fault) ? 0.5 : 1.0;
            const double c10 = skel ? -0.5 : -1.0;
            const double c20 = side == 0 ? a.fpen[fct] : -a.fpen[fct];
            const int l = a.iperm[f * Nbf + k]; // face-0 label of basis function k on face f
            for (int q = 0; q < nq; ++q) {
                double G[DD];
#pragma unroll
                for (int i = 0; i < DD; ++i) G[i] = a.fG[((fct * 2 + side) * nq + q) * DD + i];
                const double lq = a.fc0[(fct * 2 + side) * nq + q], mq = a.fc1[(fct * 2 + side) * nq + q];
                double n[D], g[Q], dx[D];
                double nn = 0.0, fn = 0.0;
#pragma unroll
                for (int j = 0; j < D; ++j) {
                    n[j] = a.fn[(fct * nq + q) * D + j];
                    nn += n[j] * n[j];
                }
#pragma unroll
                for (int p = 0; p < Q; ++p) {
                    g[p] = scale * fd[(fct * Q + p) * nq + q];
                    fn += g[p] * n[p];
                }
                double const* t = &a.tabF[(l * C) * nq + q];
#pragma unroll
                for (int j = 0; j < D; ++j) {
                    double v = 0.0;
#pragma unroll
                    for (int ee = 0; ee < D; ++ee) v += G[ee + j * D] * t[(1 + ee) * nq];
                    dx[j] = v;
                }
                double nd = 0.0, dfd = 0.0;
#pragma unroll
                for (int j = 0; j < D; ++j) {
                    nd += dx[j] * n[j];
                    dfd += dx[j] * g[j];
                }
                const double w = a.w[q], nl = sqrt(nn), E = t[0];
#pragma unroll
                for (int p = 0; p < Q; ++p) {
                    const double tt = lq * dx[p] * fn + mq * (nd * g[p] + dfd * n[p]);
                    acc[p] += c10 * tt * w + c20 * w * E * g[p] * nl;
                }
            }
        }
    }
    double* bo = b + blockIdx.y * ld_b + e * BS;
#pragma unroll
    for (int p = 0; p < Q; ++p) bo[k + p * Nbf] = acc[p];
}

template <int D, int N>
__global__ __launch_bounds__(kThreads) void elasticity_traction_kernel(RhsArgs a, double const* __restrict__ xs,
                                                                       double const* __restrict__ fds,
                                                                       double* __restrict__ tqs,
                                                                       int32_t const* __restrict__ sub, int64_t nsub,
                                                                       int64_t ld_x, int64_t ld_f, int64_t ld_t) {
    constexpr int Nbf = binom_(N + D, D), nq = fct_points_(D, N);
    constexpr int Q = D, BS = Nbf * Q, DD = D * D, C = D + 1;
    const int64_t item = int64_t(blockIdx.x) * kThreads + threadIdx.x;
    if (item >= nsub * nq) return;
    // facet `fct` of the list (all facets, or the subset sub), output row `row`, column blockIdx.y
    const int64_t row = item / nq;
    const int64_t fct = sub ? int64_t(sub[row]) : row;
    const int q = int(item % nq);
    double const* __restrict__ x = xs + int64_t(blockIdx.y) * ld_x;
    double const* __restrict__ fdata = fds ? fds + int64_t(blockIdx.y) * ld_f : nullptr;
    double* __restrict__ tq = tqs + int64_t(blockIdx.y) * ld_t;
    const int4 fi = a.fct[fct]; // {element 0, element 1, local face 0 | local face 1 << 4, -}
    const int fb = a.fbc[fct];
    const int bc = fb & 0xff;
    const bool skel = (fb >> 8) & 1;
    const int ns = skel ? 2 : 1;
    double n[D], nn = 0.0;
#pragma unroll
    for (int j = 0; j < D; ++j) {
        n[j] = a.fn[(fct * nq + q) * D + j];
        nn += n[j] * n[j];
    }
    const double inl = 1.0 / sqrt(nn);
    double tr[Q], jump[Q];
#pragma unroll
    for (int p = 0; p < Q; ++p) tr[p] = jump[p] = 0.0;
    for (int s = 0; s < ns; ++s) {
        const int64_t el = s == 0 ? fi.x : fi.y;
        const int lf = s == 0 ? (fi.z & 15) : ((fi.z >> 4) & 15);
        double const* u = x + el * BS;
        int32_t const* pm = &a.perm[lf * Nbf];
        double uv[Q], gx[D][Q];
#pragma unroll
        for (int r = 0; r < Q; ++r) {
            uv[r] = 0.0;
#pragma unroll
            for (int ee = 0; ee < D; ++ee) gx[ee][r] = 0.0;
        }
        for (int l = 0; l < Nbf; ++l) {
            const int kk = pm[l];
            double const* t = &a.tabF[(l * C) * nq + q];
#pragma unroll
            for (int r = 0; r < Q; ++r) {
                const double ul = u[kk + r * Nbf];
                uv[r] += t[0] * ul;
#pragma unroll
                for (int ee = 0; ee < D; ++ee) gx[ee][r] += t[(1 + ee) * nq] * ul;
            }
        }
        double G[DD];
#pragma unroll
        for (int i = 0; i < DD; ++i) G[i] = a.fG[((fct * 2 + s) * nq + q) * DD + i];
        const double lq = a.fc0[(fct * 2 + s) * nq + q], mq = a.fc1[(fct * 2 + s) * nq + q];
        double J[Q][D]; // J[r][j] = du_r/dx_j
        double div = 0.0;
#pragma unroll
        for (int r = 0; r < Q; ++r)
#pragma unroll
            for (int j = 0; j < D; ++j) {
                double v = 0.0;
#pragma unroll
                for (int ee = 0; ee < D; ++ee) v += G[ee + j * D] * gx[ee][r];
                J[r][j] = v;
            }
#pragma unroll
        for (int r = 0; r < D; ++r) div += J[r][r];
        const double h = skel ? 0.5 : 1.0;
#pragma unroll
        for (int p = 0; p < Q; ++p) {
            double v = lq * div * n[p] * inl;
#pragma unroll
            for (int j = 0; j < D; ++j) v += mq * (J[p][j] + J[j][p]) * n[j] * inl;
            tr[p] += h * v;
            jump[p] += s == 0 ? uv[p] : -uv[p];
        }
    }
    const bool data = fdata && (bc == kBC_Dirichlet || bc == kBC_Fault);
    const double scale = (!skel && bc == kBC_Fault) ? 0.5 : 1.0;
    const double c00 = -a.fpen[fct];
#pragma unroll
    for (int p = 0; p < Q; ++p) {
        const double f = data ? scale * fdata[(fct * Q + p) * nq + q] : 0.0;
        tq[(row * Q + p) * nq + q] = tr[p] + c00 * (jump[p] - f);
    }
}

template <int D, int N>
hipError_t run_rhs(RhsArgs const& a, double const* force, double const* fdata, double* b, int ncols,
                   int64_t ld_force, int64_t ld_fdata, int64_t ld_b, hipStream_t s) {
    constexpr int Nbf = binom_(N + D, D);
    const int64_t items = a.num_elements * Nbf;
    dim3 grid(unsigned((items + kThreads - 1) / kThreads), unsigned(ncols));
    hipLaunchKernelGGL((elasticity_rhs_kernel<D, N>), grid, dim3(kThreads), 0, s, a, force, fdata, b, ld_force,
                       ld_fdata, ld_b);
    return hipGetLastError();
}

template <int D, int N>
hipError_t run_traction(RhsArgs const& a, double const* x, double const* fdata, double* t, int32_t const* sub,
                        int64_t nsub, int ncols, int64_t ld_x, int64_t ld_f, int64_t ld_t, hipStream_t s) {
    constexpr int nq = fct_points_(D, N);
    const int64_t items = nsub * nq;
    if (items == 0) return hipSuccess;
    dim3 grid(unsigned((items + kThreads - 1) / kThreads), unsigned(ncols));
    hipLaunchKernelGGL((elasticity_traction_kernel<D, N>), grid, dim3(kThreads), 0, s, a, x, fdata, t, sub, nsub, ld_x,
                       ld_f, ld_t);
    return hipGetLastError();
}

template <class F> hipError_t by_degree(int dim, int degree, F&& f) {
    if (dim == 3) {
        switch (degree) {
        case 1: return f(std::integral_constant<int, 3>{}, std::integral_constant<int, 1>{});
        case 2: return f(std::integral_constant<int, 3>{}, std::integral_constant<int, 2>{});
        case 3: return f(std::integral_constant<int, 3>{}, std::integral_constant<int, 3>{});
        case 4: return f(std::integral_constant<int, 3>{}, std::integral_constant<int, 4>{});
        }
    } else if (dim == 2) {
        switch (degree) {
        case 1: return f(std::integral_constant<int, 2>{}, std::integral_constant<int, 1>{});
        case 2: return f(std::integral_constant<int, 2>{}, std::integral_constant<int, 2>{});
        case 3: return f(std::integral_constant<int, 2>{}, std::integral_constant<int, 3>{});
        case 4: return f(std::integral_constant<int, 2>{}, std::integral_constant<int, 4>{});
        case 5: return f(std::integral_constant<int, 2>{}, std::integral_constant<int, 5>{});
        case 6: return f(std::integral_constant<int, 2>{}, std::integral_constant<int, 6>{});
        }
    }
    return hipErrorInvalidValue;
}

} // namespace

hipError_t launch_elasticity_rhs(int dim, int degree, RhsArgs const& a, double const* force, double const* fdata,
                                 double* b, int ncols, int64_t ld_force, int64_t ld_fdata, int64_t ld_b,
                                 hipStream_t s) {
    return by_degree(dim, degree, [&](auto Dc, auto Nc) {
        return run_rhs<decltype(Dc)::value, decltype(Nc)::value>(a, force, fdata, b, ncols, ld_force, ld_fdata, ld_b,
                                                                  s);
    });
}

hipError_t launch_elasticity_traction(int dim, int degree, RhsArgs const& a, double const* x, double const* fdata,
                                      double* t, hipStream_t s) {
    return by_degree(dim, degree, [&](auto Dc, auto Nc) {
        return run_traction<decltype(Dc)::value, decltype(Nc)::value>(a, x, fdata, t, nullptr, a.num_facets, 1, 0, 0,
                                                                       0, s);
    });
}

hipError_t launch_elasticity_traction_subset(int dim, int degree, RhsArgs const& a, double const* x,
                                             double const* fdata, double* t, int32_t const* sub, int64_t nsub,
                                             int ncols, int64_t ld_x, int64_t ld_f, int64_t ld_t, hipStream_t s) {
    return by_degree(dim, degree, [&](auto Dc, auto Nc) {
        return run_traction<decltype(Dc)::value, decltype(Nc)::value>(a, x, fdata, t, sub, nsub, ncols, ld_x, ld_f,
                                                                       ld_t, s);
    });
}

} // namespace tdg
