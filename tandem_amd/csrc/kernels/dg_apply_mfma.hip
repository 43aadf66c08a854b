// Elasticity SIPG apply on the CDNA4 matrix cores (v_mfma_{f64,f32}_16x16x4).
//
// One wave owns 16 consecutive elements: the 16 columns of every MFMA tile are these elements.
//   volume   : C1(e,q ; el, r)  = sum_k TV(e,q ; k) U(k ; el, r)          [rows (e, 16-point block)]
//              per point in registers: S(e,u) = G (lamWJ tr(grad u) I + muWJ (grad u + grad u^T))
//              Y(k ; el, u)    += sum_(e,q) TV(e,q ; k) S(e,q ; el, u)     [K ordered so that each
//                                                                          lane's C1 rows are its
//                                                                          own B operands]
//   facets f : own trace rows (q, c), c = value / d/dxi_e, ordered so that lane group g receives
//              point 4t + g with channel c in register c; from the element vector U with
//              the face-f relabelling folded into the A operand; neighbour trace from the gathered,
//              relabelled neighbour vector with the face-0 table; per facet point in registers the
//              SIPG fluxes (flux_u_*, flux_sigma_*) and the apply_facet coefficients W(c, u); then
//              Y(k ; el, u) += sum_(q,c) T0(c, pi_f^-1 k, q) W(q,c ; el, u).
// MFMA operand layout (16x16x4): A[i][k] from lane i%16 + 16 k, B[k][j] from lane j + 16 k,
// D[i][j] in lane j + 16 g, register r with i = mfma_row<T>(g, r) (stacked for f32, interleaved
// for f64).  All tables are stored pre-swizzled per lane and per precision.
// Per-point coefficient data is stored element-innermost in groups of 16 so every load of a
// 16-lane row is one contiguous 16-element segment.
#include "kernels.h"
#include "mfma_layout.h"

namespace tdg {

namespace {

// D row held in register i of lane group g (D[row][j] lives in lane j + 16 g): the f32 form
// stacks the four registers of a lane on consecutive rows, the f64 form interleaves them.
template <class T> struct Mfma;
template <> struct Mfma<double> {
    using acc = double __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ acc mma(double a, double b, acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ constexpr int row(int g, int i) { return mfma_row<double>(g, i); }
};
template <> struct Mfma<float> {
    using acc = float __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ acc mma(float a, float b, acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ constexpr int row(int g, int i) { return mfma_row<float>(g, i); }
};

template <class T> __device__ __forceinline__ T inv_sqrt(T v);
template <> __device__ __forceinline__ double inv_sqrt<double>(double v) { return 1.0 / sqrt(v); }
template <> __device__ __forceinline__ float inv_sqrt<float>(float v) { return 1.0f / sqrtf(v); }

constexpr int kWaves = 4;

template <int D, int N, class T>
__global__ __launch_bounds__(64 * kWaves) void elasticity_mfma_kernel(MfmaArgs<T> a, T const* __restrict__ xs,
                                                                      T* __restrict__ ys, int64_t ld) {
    using L = MfmaLayout<D, N>;
    using M = Mfma<T>;
    using acc = typename M::acc;
    constexpr int Nbf = L::Nbf, Nq = L::Nq, nq = L::nq, Q = D, BS = Nbf * Q, DD = D * D;
    constexpr int KS = L::KS, QB = L::QB, MT = L::MT, FT = L::FT, NF = D + 1;
    constexpr int NCH = L::NCH, NFC = L::NFC;

    const int lane = threadIdx.x & 63, j = lane & 15, g = lane >> 4;
    const int64_t grp = int64_t(blockIdx.x) * kWaves + (threadIdx.x >> 6);
    if (grp >= a.num_groups) return; // uniform per wave
    T const* __restrict__ x = xs + blockIdx.y * ld;
    T* __restrict__ y = ys + blockIdx.y * ld;
    const int64_t el = grp * 16 + j;
    const bool valid = el < a.num_elements;

    // own element vector as B operands: Ub[s][r] = U(4s + g, r)
    T Ub[KS][Q];
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        const int k = 4 * s + g;
#pragma unroll
        for (int r = 0; r < Q; ++r) Ub[s][r] = (valid && k < Nbf) ? x[el * BS + k + r * Nbf] : T(0);
    }
    acc Y[MT][Q];
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int u = 0; u < Q; ++u) Y[m][u] = acc{0, 0, 0, 0};

    // ------------------------------------------------------------------ volume
#pragma unroll 1
    for (int b = 0; b < QB; ++b) {
        acc C1[D][Q];
#pragma unroll
        for (int e = 0; e < D; ++e)
#pragma unroll
            for (int r = 0; r < Q; ++r) C1[e][r] = acc{0, 0, 0, 0};
#pragma unroll
        for (int s = 0; s < KS; ++s)
#pragma unroll
            for (int e = 0; e < D; ++e) {
                const T av = a.A1[((e * QB + b) * KS + s) * 64 + lane];
#pragma unroll
                for (int r = 0; r < Q; ++r) C1[e][r] = M::mma(av, Ub[s][r], C1[e][r]);
            }
        // this lane holds points q = 16 b + 4 g + i, i = 0..3, of element el
        T S[D][Q][4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int q = 16 * b + M::row(g, i);
            const bool act = valid && q < Nq;
            T G[DD], lw = T(0), mw = T(0);
            T const* vd = a.VD + ((grp * Nq + (act ? q : 0)) * NCH) * 16 + j;
#pragma unroll
            for (int c = 0; c < DD; ++c) G[c] = act ? vd[c * 16] : T(0);
            if (act) {
                lw = vd[DD * 16];
                mw = vd[(DD + 1) * 16];
            }
            T Ju[D][Q]; // Ju[s][r] = du_r/dx_s
#pragma unroll
            for (int s2 = 0; s2 < D; ++s2)
#pragma unroll
                for (int r = 0; r < Q; ++r) {
                    T v = T(0);
#pragma unroll
                    for (int e = 0; e < D; ++e) v += G[e + s2 * D] * C1[e][r][i];
                    Ju[s2][r] = v;
                }
            T tr = T(0);
#pragma unroll
            for (int r = 0; r < D; ++r) tr += Ju[r][r];
#pragma unroll
            for (int e = 0; e < D; ++e)
#pragma unroll
                for (int u = 0; u < Q; ++u) {
                    T v = T(0);
#pragma unroll
                    for (int jj = 0; jj < D; ++jj) {
                        T P = mw * (Ju[u][jj] + Ju[jj][u]);
                        if (u == jj) P += lw * tr;
                        v += G[e + jj * D] * P;
                    }
                    S[e][u][i] = v;
                }
        }
#pragma unroll
        for (int e = 0; e < D; ++e)
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int m = 0; m < MT; ++m) {
                    const T av = a.A2[(m * (D * QB * 4) + (e * QB + b) * 4 + i) * 64 + lane];
#pragma unroll
                    for (int u = 0; u < Q; ++u) Y[m][u] = M::mma(av, S[e][u][i], Y[m][u]);
                }
    }

    // ------------------------------------------------------------------ facets
#pragma unroll 1
    for (int f = 0; f < NF; ++f) {
        const int4 info = a.FI[(grp * NF + f) * 16 + j];
        const int kind = valid ? (info.x & 3) : 0;
        const int nbface = (info.x >> 4) & 3;
        const int64_t nb = info.y;
        const T pen = a.FP[(grp * NF + f) * 16 + j];
        // neighbour element vector relabelled to its face-0 frame: Nb[s][r] = U_nb(pi_nbface(4s + g), r)
        T Nb[KS][Q];
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            const int k = 4 * s + g;
            const bool act = kind == 1 && k < Nbf;
            const int pk = act ? a.perm[nbface * Nbf + k] : 0;
#pragma unroll
            for (int r = 0; r < Q; ++r) Nb[s][r] = act ? x[nb * BS + pk + r * Nbf] : T(0);
        }
#pragma unroll 1
        for (int t = 0; t < FT; ++t) {
            acc Co[Q], Cn[Q];
#pragma unroll
            for (int r = 0; r < Q; ++r) Co[r] = Cn[r] = acc{0, 0, 0, 0};
#pragma unroll
            for (int s = 0; s < KS; ++s) {
                const T ao = a.AF[((f * FT + t) * KS + s) * 64 + lane];
                const T an = a.AF[(t * KS + s) * 64 + lane];
#pragma unroll
                for (int r = 0; r < Q; ++r) {
                    Co[r] = M::mma(ao, Ub[s][r], Co[r]);
                    Cn[r] = M::mma(an, Nb[s][r], Cn[r]);
                }
            }
            // this lane: facet point q = 4t + g; register c = value (0) / d/dxi_e (1 + e)
            const int q = 4 * t + g;
            const bool act = kind != 0 && q < nq;
            T W[4][Q];
#pragma unroll
            for (int c = 0; c < 4; ++c)
#pragma unroll
                for (int u = 0; u < Q; ++u) W[c][u] = T(0);
            if (act) {
                T const* fd = a.FD + (((grp * NF + f) * nq + q) * NFC) * 16 + j;
                T G0[DD], G1[DD], n[D];
#pragma unroll
                for (int c = 0; c < DD; ++c) {
                    G0[c] = fd[c * 16];
                    G1[c] = fd[(DD + 2 + c) * 16];
                }
                const T l0 = fd[DD * 16], m0 = fd[(DD + 1) * 16];
                const T l1 = fd[(2 * DD + 2) * 16], m1 = fd[(2 * DD + 3) * 16];
                T nn = T(0);
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    n[d] = fd[(2 * DD + 4 + d) * 16];
                    nn += n[d] * n[d];
                }
                const T inl = inv_sqrt<T>(nn);
                T u0[Q], u1[Q], J0[Q][D], J1[Q][D];
#pragma unroll
                for (int r = 0; r < Q; ++r) {
                    u0[r] = Co[r][0];
                    u1[r] = Cn[r][0];
#pragma unroll
                    for (int s2 = 0; s2 < D; ++s2) {
                        T v0 = T(0), v1 = T(0);
#pragma unroll
                        for (int e = 0; e < D; ++e) {
                            v0 += G0[e + s2 * D] * Co[r][1 + e];
                            v1 += G1[e + s2 * D] * Cn[r][1 + e];
                        }
                        J0[r][s2] = v0;
                        J1[r][s2] = v1;
                    }
                }
                T tr0 = T(0), tr1 = T(0);
#pragma unroll
                for (int r = 0; r < D; ++r) {
                    tr0 += J0[r][r];
                    tr1 += J1[r][r];
                }
                T d[Q], sig[Q][D];
                const bool skel = kind == 1;
#pragma unroll
                for (int i = 0; i < Q; ++i) {
                    d[i] = skel ? T(0.5) * (u1[i] - u0[i]) : -u0[i];
                    const T jump = skel ? (u0[i] - u1[i]) : u0[i];
#pragma unroll
                    for (int jj = 0; jj < D; ++jj) {
                        T s0 = m0 * (J0[i][jj] + J0[jj][i]);
                        T s1 = m1 * (J1[i][jj] + J1[jj][i]);
                        if (i == jj) {
                            s0 += l0 * tr0;
                            s1 += l1 * tr1;
                        }
                        sig[i][jj] = (skel ? T(0.5) * (s0 + s1) : s0) - pen * jump * (n[jj] * inl);
                    }
                }
                const T wq = a.w[q];
                T dn = T(0);
#pragma unroll
                for (int r = 0; r < D; ++r) dn += d[r] * n[r];
#pragma unroll
                for (int u = 0; u < Q; ++u) {
                    T v = T(0);
#pragma unroll
                    for (int jj = 0; jj < D; ++jj) v += n[jj] * sig[u][jj];
                    W[0][u] = -wq * v;
                }
#pragma unroll
                for (int e = 0; e < D; ++e)
#pragma unroll
                    for (int u = 0; u < Q; ++u) {
                        T v = T(0);
#pragma unroll
                        for (int jj = 0; jj < D; ++jj) {
                            T Cuj = m0 * (d[u] * n[jj] + d[jj] * n[u]);
                            if (u == jj) Cuj += l0 * dn;
                            v += G0[e + jj * D] * Cuj;
                        }
                        W[1 + e][u] = wq * v;
                    }
            }
#pragma unroll
            for (int c = 0; c < 4; ++c)
#pragma unroll
                for (int m = 0; m < MT; ++m) {
                    const T av = a.AL[((f * MT + m) * (FT * 4) + t * 4 + c) * 64 + lane];
#pragma unroll
                    for (int u = 0; u < Q; ++u) Y[m][u] = M::mma(av, W[c][u], Y[m][u]);
                }
        }
    }

    // ------------------------------------------------------------------ store: Y(16 m + 4 g + i, u)
    if (valid) {
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int k = 16 * m + M::row(g, i);
                if (k < Nbf) {
#pragma unroll
                    for (int u = 0; u < Q; ++u) y[el * BS + k + u * Nbf] = Y[m][u][i];
                }
            }
    }
}

template <int D, int N, class T>
hipError_t run(MfmaArgs<T> const& a, T const* x, T* y, int ncols, int64_t ld, hipStream_t s) {
    dim3 grid(unsigned((a.num_groups + kWaves - 1) / kWaves), unsigned(ncols));
    hipLaunchKernelGGL((elasticity_mfma_kernel<D, N, T>), grid, dim3(64 * kWaves), 0, s, a, x, y, ld);
    return hipGetLastError();
}

} // namespace

template <class T>
hipError_t launch_elasticity_mfma(int dim, int degree, MfmaArgs<T> const& a, T const* x, T* y, int ncols, int64_t ld,
                                  hipStream_t s) {
    if (dim == 3) {
        switch (degree) {
        case 2: return run<3, 2, T>(a, x, y, ncols, ld, s);
        case 3: return run<3, 3, T>(a, x, y, ncols, ld, s);
        case 4: return run<3, 4, T>(a, x, y, ncols, ld, s);
        }
    } else if (dim == 2) {
        switch (degree) {
        case 2: return run<2, 2, T>(a, x, y, ncols, ld, s);
        case 4: return run<2, 4, T>(a, x, y, ncols, ld, s);
        case 6: return run<2, 6, T>(a, x, y, ncols, ld, s);
        }
    }
    return hipErrorInvalidValue;
}

bool mfma_supported(int dim, int degree) {
    return (dim == 3 && degree >= 2 && degree <= 4) || (dim == 2 && (degree == 2 || degree == 4 || degree == 6));
}

template hipError_t launch_elasticity_mfma<double>(int, int, MfmaArgs<double> const&, double const*, double*, int,
                                                   int64_t, hipStream_t);
template hipError_t launch_elasticity_mfma<float>(int, int, MfmaArgs<float> const&, float const*, float*, int,
                                                  int64_t, hipStream_t);

} // namespace tdg
