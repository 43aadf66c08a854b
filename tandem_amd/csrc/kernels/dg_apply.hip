// Element-centric matrix-free SIPG apply on CDNA4 (gfx950).
//
// One workgroup of 256 lanes owns EPB consecutive elements.  The element blocks of x are staged in
// LDS; the lanes then (1) evaluate the stress at the volume quadrature points (one lane per element
// and point) and the numerical fluxes at the facet points (one lane per element, face and point),
// writing the lifted coefficients to LDS, and (2) contract them back onto the basis (one lane per
// element and basis function).  Reference tables are streamed from L1/L2; per-element Jacobians and
// coefficients and per-facet data are streamed from HBM once per element / facet side.
//
// Elasticity uses one facet table: the warp-and-blend nodes are symmetric, so the trace on face f
// is the face-0 trace of the relabelled element vector U(pi_f(j)); the facet map's Jacobian is folded
// into the per-side facet geometry by the host precompute.
#include "kernels.h"

namespace tdg {

namespace {

constexpr int binom(int n, int k) { return k == 0 ? 1 : binom(n - 1, k - 1) * n / k; }
constexpr int pow2_at_least(int n) {
    int p = 1;
    while (p < n) p *= 2;
    return p;
}

// Rule sizes of tandem's autorule at order 2N+1 (AutoRule.h:41-51).
constexpr int vol_points(int D, int N) {
    return D == 2 ? (N == 1 ? 4 : N == 2 ? 7 : N == 3 ? 15 : N == 4 ? 19 : N == 5 ? 28 : 37)
                  : (N == 1 ? 6 : N == 2 ? 14 : N == 3 ? 31 : 57);
}
constexpr int fct_points(int D, int N) {
    return D == 2 ? N + 1 : (N == 1 ? 4 : N == 2 ? 7 : N == 3 ? 15 : 19);
}

template <int D, int N> struct Shape {
    static constexpr int Nbf = binom(N + D, D);
    static constexpr int Nq = vol_points(D, N);
    static constexpr int nq = fct_points(D, N);
    static constexpr int F = D + 1;
    static constexpr int QP = pow2_at_least(Nq);
    static constexpr int EPB = QP >= 256 ? 1 : 256 / QP;
};

constexpr int kThreads = 256;

template <class T> __device__ inline T rsqrt_t(T v);
template <> __device__ inline double rsqrt_t<double>(double v) { return 1.0 / sqrt(v); }
template <> __device__ inline float rsqrt_t<float>(float v) { return 1.0f / sqrtf(v); }

// ------------------------------------------------------------------------------ elasticity
template <int D, int N, class T>
__global__ __launch_bounds__(kThreads) void elasticity_apply_kernel(ApplyArgs<T> a, T const* __restrict__ xs,
                                                                    T* __restrict__ ys, int64_t ld) {
    using S = Shape<D, N>;
    constexpr int Nbf = S::Nbf, Nq = S::Nq, nq = S::nq, F = S::F, Q = D, BS = Nbf * Q, DD = D * D;
    constexpr int EPB = S::EPB, QP = S::QP, C = D + 1;
    __shared__ T sU[EPB * BS];
    __shared__ T sS[EPB * Nq * D * Q];
    __shared__ T sF[EPB * F * nq * C * Q];

    T const* __restrict__ x = xs + blockIdx.y * ld;
    T* __restrict__ y = ys + blockIdx.y * ld;
    const int tid = threadIdx.x;
    const int64_t e0 = int64_t(blockIdx.x) * EPB;
    const int ne = int(min<int64_t>(EPB, a.num_elements - e0));

    for (int i = tid; i < ne * BS; i += kThreads) sU[i] = x[e0 * BS + i];
    __syncthreads();

    // ---- volume points: S(e,u) = sum_j G(e,j) [lamWJ delta_uj div u + muWJ (du_u/dx_j + du_j/dx_u)]
    for (int item = tid; item < EPB * QP; item += kThreads) {
        const int el = item / QP, q = item % QP;
        if (el >= ne || q >= Nq) continue;
        const int64_t eg = e0 + el;
        T g[D][Q];
#pragma unroll
        for (int e = 0; e < D; ++e)
#pragma unroll
            for (int r = 0; r < Q; ++r) g[e][r] = T(0);
        T const* U = &sU[el * BS];
        for (int l = 0; l < Nbf; ++l) {
            T ul[Q];
#pragma unroll
            for (int r = 0; r < Q; ++r) ul[r] = U[l + r * Nbf];
#pragma unroll
            for (int e = 0; e < D; ++e) {
                T t = a.tabV[(l * D + e) * Nq + q];
#pragma unroll
                for (int r = 0; r < Q; ++r) g[e][r] += t * ul[r];
            }
        }
        T G[DD];
#pragma unroll
        for (int i = 0; i < DD; ++i) G[i] = a.G[(eg * Nq + q) * DD + i];
        const T lw = a.c0[eg * Nq + q], mw = a.c1[eg * Nq + q];
        T Ju[D][Q]; // Ju[s][r] = du_r/dx_s
#pragma unroll
        for (int s = 0; s < D; ++s)
#pragma unroll
            for (int r = 0; r < Q; ++r) {
                T v = T(0);
#pragma unroll
                for (int e = 0; e < D; ++e) v += G[e + s * D] * g[e][r];
                Ju[s][r] = v;
            }
        T tr = T(0);
#pragma unroll
        for (int r = 0; r < D; ++r) tr += Ju[r][r];
        T* out = &sS[(el * Nq + q) * D * Q];
#pragma unroll
        for (int e = 0; e < D; ++e)
#pragma unroll
            for (int u = 0; u < Q; ++u) {
                T v = T(0);
#pragma unroll
                for (int j = 0; j < D; ++j) {
                    T P = mw * (Ju[u][j] + Ju[j][u]);
                    if (u == j) P += lw * tr;
                    v += G[e + j * D] * P;
                }
                out[e * Q + u] = v;
            }
    }

    // ---- facet points: numerical fluxes (flux_u_*, flux_sigma_*) and lift coefficients
    for (int item = tid; item < ne * F * nq; item += kThreads) {
        const int el = item / (F * nq), rem = item % (F * nq), f = rem / nq, q = rem % nq;
        const int64_t eg = e0 + el;
        const int4 info = a.ef[eg * F + f];
        T* out = &sF[((el * F + f) * nq + q) * C * Q];
        const int kind = info.z & 3;
        if (kind == 0) {
#pragma unroll
            for (int i = 0; i < C * Q; ++i) out[i] = T(0);
            continue;
        }
        const int side = (info.z >> 2) & 1, nbf_ = (info.z >> 4) & 3;
        const int64_t fct = info.x;
        // own trace of the relabelled element vector with the face-0 table
        T u0[Q], g0[D][Q];
#pragma unroll
        for (int r = 0; r < Q; ++r) {
            u0[r] = T(0);
#pragma unroll
            for (int e = 0; e < D; ++e) g0[e][r] = T(0);
        }
        {
            T const* U = &sU[el * BS];
            int32_t const* pm = &a.perm[f * Nbf];
            for (int l = 0; l < Nbf; ++l) {
                const int j = pm[l];
                T ul[Q];
#pragma unroll
                for (int r = 0; r < Q; ++r) ul[r] = U[j + r * Nbf];
                T const* t = &a.tabF[(l * C) * nq + q];
#pragma unroll
                for (int r = 0; r < Q; ++r) u0[r] += t[0] * ul[r];
#pragma unroll
                for (int e = 0; e < D; ++e)
#pragma unroll
                    for (int r = 0; r < Q; ++r) g0[e][r] += t[(1 + e) * nq] * ul[r];
            }
        }
        T Gp[DD];
#pragma unroll
        for (int i = 0; i < DD; ++i) Gp[i] = a.fG[((fct * 2 + side) * nq + q) * DD + i];
        const T l0 = a.fc0[(fct * 2 + side) * nq + q], m0 = a.fc1[(fct * 2 + side) * nq + q];
        T J0[Q][D]; // J0[r][s] = du_r/dx_s
#pragma unroll
        for (int r = 0; r < Q; ++r)
#pragma unroll
            for (int s = 0; s < D; ++s) {
                T v = T(0);
#pragma unroll
                for (int e = 0; e < D; ++e) v += Gp[e + s * D] * g0[e][r];
                J0[r][s] = v;
            }
        T n[D];
        T nn = T(0);
#pragma unroll
        for (int j = 0; j < D; ++j) {
            n[j] = a.fn[(fct * nq + q) * D + j];
            if (side) n[j] = -n[j];
            nn += n[j] * n[j];
        }
        const T inl = rsqrt_t<T>(nn);
        const T pen = a.fpen[fct];
        T d[Q], sig[Q][D];
        {
            T tr0 = T(0);
#pragma unroll
            for (int r = 0; r < D; ++r) tr0 += J0[r][r];
#pragma unroll
            for (int i = 0; i < Q; ++i)
#pragma unroll
                for (int j = 0; j < D; ++j) {
                    sig[i][j] = m0 * (J0[i][j] + J0[j][i]);
                    if (i == j) sig[i][j] += l0 * tr0;
                }
        }
        if (kind == 1) {
            // neighbour trace (flux_u_skeleton / flux_sigma_skeleton)
            T u1[Q], g1[D][Q];
#pragma unroll
            for (int r = 0; r < Q; ++r) {
                u1[r] = T(0);
#pragma unroll
                for (int e = 0; e < D; ++e) g1[e][r] = T(0);
            }
            T const* Ux = &x[int64_t(info.y) * BS];
            int32_t const* pm = &a.perm[nbf_ * Nbf];
            for (int l = 0; l < Nbf; ++l) {
                const int j = pm[l];
                T ul[Q];
#pragma unroll
                for (int r = 0; r < Q; ++r) ul[r] = Ux[j + r * Nbf];
                T const* t = &a.tabF[(l * C) * nq + q];
#pragma unroll
                for (int r = 0; r < Q; ++r) u1[r] += t[0] * ul[r];
#pragma unroll
                for (int e = 0; e < D; ++e)
#pragma unroll
                    for (int r = 0; r < Q; ++r) g1[e][r] += t[(1 + e) * nq] * ul[r];
            }
            T G1[DD];
#pragma unroll
            for (int i = 0; i < DD; ++i) G1[i] = a.fG[((fct * 2 + (1 - side)) * nq + q) * DD + i];
            const T l1 = a.fc0[(fct * 2 + (1 - side)) * nq + q], m1 = a.fc1[(fct * 2 + (1 - side)) * nq + q];
            T J1[Q][D];
#pragma unroll
            for (int r = 0; r < Q; ++r)
#pragma unroll
                for (int s = 0; s < D; ++s) {
                    T v = T(0);
#pragma unroll
                    for (int e = 0; e < D; ++e) v += G1[e + s * D] * g1[e][r];
                    J1[r][s] = v;
                }
            T tr1 = T(0);
#pragma unroll
            for (int r = 0; r < D; ++r) tr1 += J1[r][r];
#pragma unroll
            for (int i = 0; i < Q; ++i) {
                d[i] = T(0.5) * (u1[i] - u0[i]);
#pragma unroll
                for (int j = 0; j < D; ++j) {
                    T s1 = m1 * (J1[i][j] + J1[j][i]);
                    if (i == j) s1 += l1 * tr1;
                    sig[i][j] = T(0.5) * (sig[i][j] + s1) - pen * (u0[i] - u1[i]) * (n[j] * inl);
                }
            }
        } else {
            // flux_u_boundary / flux_sigma_boundary
#pragma unroll
            for (int i = 0; i < Q; ++i) {
                d[i] = -u0[i];
#pragma unroll
                for (int j = 0; j < D; ++j) sig[i][j] = sig[i][j] - pen * u0[i] * (n[j] * inl);
            }
        }
        // apply_facet coefficients: -w (n . sigma_hat)_u on the value channel,
        // w G'(e,j) [lam delta_uj d.n + mu (d_u n_j + d_j n_u)] on the gradient channels
        const T wq = a.w[q];
        T dn = T(0);
#pragma unroll
        for (int r = 0; r < D; ++r) dn += d[r] * n[r];
#pragma unroll
        for (int u = 0; u < Q; ++u) {
            T v = T(0);
#pragma unroll
            for (int j = 0; j < D; ++j) v += n[j] * sig[u][j];
            out[u] = -wq * v;
        }
#pragma unroll
        for (int e = 0; e < D; ++e)
#pragma unroll
            for (int u = 0; u < Q; ++u) {
                T v = T(0);
#pragma unroll
                for (int j = 0; j < D; ++j) {
                    T Cuj = m0 * (d[u] * n[j] + d[j] * n[u]);
                    if (u == j) Cuj += l0 * dn;
                    v += Gp[e + j * D] * Cuj;
                }
                out[(1 + e) * Q + u] = wq * v;
            }
    }
    __syncthreads();

    // ---- contraction onto the basis: volume lift + facet lifts (face f relabelled back)
    for (int item = tid; item < ne * Nbf; item += kThreads) {
        const int el = item / Nbf, k = item % Nbf;
        T acc[Q];
#pragma unroll
        for (int u = 0; u < Q; ++u) acc[u] = T(0);
        T const* Sv = &sS[el * Nq * D * Q];
        for (int q = 0; q < Nq; ++q)
#pragma unroll
            for (int e = 0; e < D; ++e) {
                T t = a.tabVt[(e * Nq + q) * Nbf + k];
#pragma unroll
                for (int u = 0; u < Q; ++u) acc[u] += t * Sv[(q * D + e) * Q + u];
            }
#pragma unroll
        for (int f = 0; f < F; ++f) {
            const int j = a.iperm[f * Nbf + k];
            T const* Fl = &sF[(el * F + f) * nq * C * Q];
            for (int q = 0; q < nq; ++q)
#pragma unroll
                for (int c = 0; c < C; ++c) {
                    T t = a.tabFt[(c * nq + q) * Nbf + j];
#pragma unroll
                    for (int u = 0; u < Q; ++u) acc[u] += t * Fl[(q * C + c) * Q + u];
                }
        }
#pragma unroll
        for (int u = 0; u < Q; ++u) y[(e0 + el) * BS + k + u * Nbf] = acc[u];
    }
}

// ------------------------------------------------------------------------------ poisson
template <int D, int N, class T>
__global__ __launch_bounds__(kThreads) void poisson_apply_kernel(ApplyArgs<T> a, T const* __restrict__ xs,
                                                                 T* __restrict__ ys, int64_t ld) {
    using S = Shape<D, N>;
    constexpr int Nbf = S::Nbf, Nq = S::Nq, nq = S::nq, F = S::F, DD = D * D;
    constexpr int EPB = S::EPB, QP = S::QP, C = D + 1;
    __shared__ T sU[EPB * Nbf];
    __shared__ T sS[EPB * Nq * D];
    __shared__ T sF[EPB * F * nq * C];

    T const* __restrict__ x = xs + blockIdx.y * ld;
    T* __restrict__ y = ys + blockIdx.y * ld;
    const int tid = threadIdx.x;
    const int64_t e0 = int64_t(blockIdx.x) * EPB;
    const int ne = int(min<int64_t>(EPB, a.num_elements - e0));
    for (int i = tid; i < ne * Nbf; i += kThreads) sU[i] = x[e0 * Nbf + i];
    __syncthreads();

    // volume: S(e) = JWK sum_r G(e,r) du/dx_r  (apply_volume, poisson.py:139-142)
    for (int item = tid; item < EPB * QP; item += kThreads) {
        const int el = item / QP, q = item % QP;
        if (el >= ne || q >= Nq) continue;
        const int64_t eg = e0 + el;
        T g[D];
#pragma unroll
        for (int e = 0; e < D; ++e) g[e] = T(0);
        for (int l = 0; l < Nbf; ++l) {
            T ul = sU[el * Nbf + l];
#pragma unroll
            for (int e = 0; e < D; ++e) g[e] += a.tabV[(l * D + e) * Nq + q] * ul;
        }
        T G[DD];
#pragma unroll
        for (int i = 0; i < DD; ++i) G[i] = a.G[(eg * Nq + q) * DD + i];
        const T jwk = a.c0[eg * Nq + q];
        T du[D];
#pragma unroll
        for (int r = 0; r < D; ++r) {
            T v = T(0);
#pragma unroll
            for (int e = 0; e < D; ++e) v += G[e + r * D] * g[e];
            du[r] = v;
        }
#pragma unroll
        for (int e = 0; e < D; ++e) {
            T v = T(0);
#pragma unroll
            for (int r = 0; r < D; ++r) v += G[e + r * D] * du[r];
            sS[(el * Nq + q) * D + e] = jwk * v;
        }
    }

    // facets (flux_u_*, flux_sigma_*, apply_facet of poisson.py:123-144), per-face tables
    for (int item = tid; item < ne * F * nq; item += kThreads) {
        const int el = item / (F * nq), rem = item % (F * nq), f = rem / nq, q = rem % nq;
        const int64_t eg = e0 + el;
        const int4 info = a.ef[eg * F + f];
        T* out = &sF[((el * F + f) * nq + q) * C];
        const int kind = info.z & 3;
        if (kind == 0) {
#pragma unroll
            for (int i = 0; i < C; ++i) out[i] = T(0);
            continue;
        }
        const int side = (info.z >> 2) & 1, nbf_ = (info.z >> 4) & 3;
        const int64_t fct = info.x;
        T u0 = T(0), g0[D];
#pragma unroll
        for (int e = 0; e < D; ++e) g0[e] = T(0);
        for (int l = 0; l < Nbf; ++l) {
            T ul = sU[el * Nbf + l];
            T const* t = &a.tabF[((f * Nbf + l) * C) * nq + q];
            u0 += t[0] * ul;
#pragma unroll
            for (int e = 0; e < D; ++e) g0[e] += t[(1 + e) * nq] * ul;
        }
        T K0[DD];
#pragma unroll
        for (int i = 0; i < DD; ++i) K0[i] = a.fG[((fct * 2 + side) * nq + q) * DD + i];
        T n[D], nn = T(0);
#pragma unroll
        for (int j = 0; j < D; ++j) {
            n[j] = a.fn[(fct * nq + q) * D + j];
            if (side) n[j] = -n[j];
            nn += n[j] * n[j];
        }
        const T inl = rsqrt_t<T>(nn);
        const T pen = a.fpen[fct];
        T sig[D], uh;
#pragma unroll
        for (int p = 0; p < D; ++p) {
            T v = T(0);
#pragma unroll
            for (int e = 0; e < D; ++e) v += K0[e + p * D] * g0[e];
            sig[p] = v;
        }
        if (kind == 1) {
            T u1 = T(0), g1[D];
#pragma unroll
            for (int e = 0; e < D; ++e) g1[e] = T(0);
            T const* Ux = &x[int64_t(info.y) * Nbf];
            for (int l = 0; l < Nbf; ++l) {
                T ul = Ux[l];
                T const* t = &a.tabF[((nbf_ * Nbf + l) * C) * nq + q];
                u1 += t[0] * ul;
#pragma unroll
                for (int e = 0; e < D; ++e) g1[e] += t[(1 + e) * nq] * ul;
            }
            T K1[DD];
#pragma unroll
            for (int i = 0; i < DD; ++i) K1[i] = a.fG[((fct * 2 + (1 - side)) * nq + q) * DD + i];
            uh = T(0.5) * (u1 - u0);
#pragma unroll
            for (int p = 0; p < D; ++p) {
                T v = T(0);
#pragma unroll
                for (int e = 0; e < D; ++e) v += K1[e + p * D] * g1[e];
                sig[p] = T(0.5) * (sig[p] + v) - pen * (u0 - u1) * (n[p] * inl);
            }
        } else {
            uh = -u0;
#pragma unroll
            for (int p = 0; p < D; ++p) sig[p] = sig[p] - pen * u0 * (n[p] * inl);
        }
        const T wq = a.w[q];
        T V = T(0);
#pragma unroll
        for (int r = 0; r < D; ++r) V += n[r] * sig[r];
        out[0] = -wq * V;
#pragma unroll
        for (int e = 0; e < D; ++e) {
            T v = T(0);
#pragma unroll
            for (int r = 0; r < D; ++r) v += K0[e + r * D] * n[r];
            out[1 + e] = wq * uh * v;
        }
    }
    __syncthreads();

    for (int item = tid; item < ne * Nbf; item += kThreads) {
        const int el = item / Nbf, k = item % Nbf;
        T acc = T(0);
        for (int q = 0; q < Nq; ++q)
#pragma unroll
            for (int e = 0; e < D; ++e) acc += a.tabVt[(e * Nq + q) * Nbf + k] * sS[(el * Nq + q) * D + e];
#pragma unroll
        for (int f = 0; f < F; ++f) {
            T const* Fl = &sF[(el * F + f) * nq * C];
            for (int q = 0; q < nq; ++q)
#pragma unroll
                for (int c = 0; c < C; ++c) acc += a.tabFt[((f * C + c) * nq + q) * Nbf + k] * Fl[q * C + c];
        }
        y[(e0 + el) * Nbf + k] = acc;
    }
}

template <int D, int N, class T>
hipError_t run_elasticity(ApplyArgs<T> const& a, T const* x, T* y, int ncols, int64_t ld, hipStream_t s) {
    using Sh = Shape<D, N>;
    dim3 grid(unsigned((a.num_elements + Sh::EPB - 1) / Sh::EPB), unsigned(ncols));
    hipLaunchKernelGGL((elasticity_apply_kernel<D, N, T>), grid, dim3(kThreads), 0, s, a, x, y, ld);
    return hipGetLastError();
}
template <int D, int N, class T>
hipError_t run_poisson(ApplyArgs<T> const& a, T const* x, T* y, int ncols, int64_t ld, hipStream_t s) {
    using Sh = Shape<D, N>;
    dim3 grid(unsigned((a.num_elements + Sh::EPB - 1) / Sh::EPB), unsigned(ncols));
    hipLaunchKernelGGL((poisson_apply_kernel<D, N, T>), grid, dim3(kThreads), 0, s, a, x, y, ld);
    return hipGetLastError();
}

} // namespace

int kernel_rule_sizes(int dim, int degree, int* Nq, int* nq) {
    if ((dim != 2 && dim != 3) || degree < 1 || degree > (dim == 2 ? 6 : 4)) return -1;
    *Nq = vol_points(dim, degree);
    *nq = fct_points(dim, degree);
    return 0;
}

template <class T>
hipError_t launch_elasticity(int dim, int degree, ApplyArgs<T> const& a, T const* x, T* y, int ncols, int64_t ld,
                             hipStream_t s) {
    if (dim == 3) {
        switch (degree) {
        case 1: return run_elasticity<3, 1, T>(a, x, y, ncols, ld, s);
        case 2: return run_elasticity<3, 2, T>(a, x, y, ncols, ld, s);
        case 3: return run_elasticity<3, 3, T>(a, x, y, ncols, ld, s);
        case 4: return run_elasticity<3, 4, T>(a, x, y, ncols, ld, s);
        }
    } else if (dim == 2) {
        switch (degree) {
        case 1: return run_elasticity<2, 1, T>(a, x, y, ncols, ld, s);
        case 2: return run_elasticity<2, 2, T>(a, x, y, ncols, ld, s);
        case 3: return run_elasticity<2, 3, T>(a, x, y, ncols, ld, s);
        case 4: return run_elasticity<2, 4, T>(a, x, y, ncols, ld, s);
        case 5: return run_elasticity<2, 5, T>(a, x, y, ncols, ld, s);
        case 6: return run_elasticity<2, 6, T>(a, x, y, ncols, ld, s);
        }
    }
    return hipErrorInvalidValue;
}

template <class T>
hipError_t launch_poisson(int dim, int degree, ApplyArgs<T> const& a, T const* x, T* y, int ncols, int64_t ld,
                          hipStream_t s) {
    if (dim == 3) {
        switch (degree) {
        case 1: return run_poisson<3, 1, T>(a, x, y, ncols, ld, s);
        case 2: return run_poisson<3, 2, T>(a, x, y, ncols, ld, s);
        case 3: return run_poisson<3, 3, T>(a, x, y, ncols, ld, s);
        case 4: return run_poisson<3, 4, T>(a, x, y, ncols, ld, s);
        }
    } else if (dim == 2) {
        switch (degree) {
        case 1: return run_poisson<2, 1, T>(a, x, y, ncols, ld, s);
        case 2: return run_poisson<2, 2, T>(a, x, y, ncols, ld, s);
        case 3: return run_poisson<2, 3, T>(a, x, y, ncols, ld, s);
        case 4: return run_poisson<2, 4, T>(a, x, y, ncols, ld, s);
        case 5: return run_poisson<2, 5, T>(a, x, y, ncols, ld, s);
        case 6: return run_poisson<2, 6, T>(a, x, y, ncols, ld, s);
        }
    }
    return hipErrorInvalidValue;
}

template hipError_t launch_elasticity<double>(int, int, ApplyArgs<double> const&, double const*, double*, int, int64_t,
                                              hipStream_t);
template hipError_t launch_elasticity<float>(int, int, ApplyArgs<float> const&, float const*, float*, int, int64_t,
                                             hipStream_t);
template hipError_t launch_poisson<double>(int, int, ApplyArgs<double> const&, double const*, double*, int, int64_t,
                                           hipStream_t);
template hipError_t launch_poisson<float>(int, int, ApplyArgs<float> const&, float const*, float*, int, int64_t,
                                          hipStream_t);

} // namespace tdg
