// Elasticity SIPG apply in two matrix-core kernels: a facet kernel that evaluates every facet once
// (both sides' traces, the flux and both lifts) and an element kernel that applies the volume term
// and adds the four lifted facet contributions of each element.
//
// Facet kernel.  One wave owns 8 facets; the 16 columns of every 16x16x4 MFMA tile are their two
// sides (side 0 of facet c in column c, side 1 in column 15 - c), so each lane works on one side
// and the two sides of a point trade their traces across a DPP row mirror.  Two-sided facets come
// first in the facet order, one-sided (Dirichlet or fault boundary) after them; the side-1 columns
// of a one-sided facet carry nothing.  Everything is computed in the face-0 frame of each side:
// the element vector of each column is staged in LDS relabelled to face 0
// (U'(k') = x(perm_f(k'))), so one table T0 of face-0 values and reference gradients is the A
// operand of every trace and every lift whatever the local face numbers of the 16 facets, and the
// lifted rows are relabelled back (row k' -> element row perm_f(k')) when they are written:
//   traces   : C_s(c,q ; j, r) = sum_k' T0(c, q ; k') U'_s(k' ; j, r)        rows = 4 points x 4 channels
//   per point: geometry from the stored facet Jacobian M = J J_f (tangents shared by both sides,
//              the third column per side): G'_s = M_s^-1, the area-scaled normal n = -+ t1 x t2
//              (outward from side 0); u_s, grad u_s = G'_s^T grad_xi u_s, sigma_s n, and from the
//              mirrored lane the other side's u and sigma n; the flux
//              tau = 1/2 (sigma_0 + sigma_1) n - delta |n| (u_0 - u_1), d = 1/2 (u_1 - u_0)
//              (one-sided: tau = sigma_0 n - delta |n| u_0, d = -u_0) and the lift weights
//              W_0(0) = -w tau, W_1(0) = +w tau, W_s(1+e) = w sum_j G'_s(e, j) C_s(u, j),
//              C_s = mu_s (d n^T + n d^T) + lam_s (d . n) I
//   lifts    : Y_s(k' ; j, u) += sum_(q,c) T0(c, q ; k') W_s(c, q ; j, u)
// and the result of side s goes to the slot L(element_s, face_s) (one 105-vector at P4 3D).
//
// Element kernel.  One wave owns 16 consecutive elements; the volume term is the element-centric
// kernel's (reference gradients on the matrix cores, the stress per point in registers, the
// divergence on the matrix cores), with the accumulators started from the sum of the element's
// facet slots.  Slots of faces that carry no facet term (natural boundaries) are zero: the
// operator clears the slot buffer when it allocates it and the facet kernel never writes them.
#pragma once

#include "kernels.h"
#include "mfma_layout.h"
#include "mfma_ops.h"

#include <algorithm>
#include <type_traits>

namespace tdg {

namespace split {

// Reciprocal from the hardware estimate refined by one Newton step (relative error near the rounding
// unit, far below what the result needs); the argument is a nonzero Jacobian determinant.
template <class T> __device__ __forceinline__ T rcp_(T v);
template <> __device__ __forceinline__ double rcp_<double>(double v) {
    const double r = __builtin_amdgcn_rcp(v);
    return fma(r, fma(-v, r, 1.0), r);
}
template <> __device__ __forceinline__ float rcp_<float>(float v) {
    const float r = __builtin_amdgcn_rcpf(v);
    return fmaf(r, fmaf(-v, r, 1.0f), r);
}
template <class T> __device__ __forceinline__ T sqrt_(T v);
template <> __device__ __forceinline__ double sqrt_<double>(double v) { return sqrt(v); }
template <> __device__ __forceinline__ float sqrt_<float>(float v) { return sqrtf(v); }

template <class T> struct Strides {
    // FP32 pads table rows and vectors to twice an odd number of words (distinct banks for the 32
    // lanes of a half wave reading 16 rows at 2 consecutive offsets); FP64 keeps them unpadded.
    static constexpr int pad(int n) { return sizeof(T) == 4 ? (n + 1) / 4 * 4 + 2 : n; }
};

// Lane index within the wave.  In a persistent kernel it is hidden from the optimiser, so that the
// lane-dependent addresses of the block loop are computed where they are used instead of being
// hoisted out of the loop and held in registers across it.
template <bool PERSIST> __device__ __forceinline__ int lane_of(int tid) {
    int v = tid & 63;
    if constexpr (PERSIST) __asm__ volatile("" : "+v"(v));
    return v & 63;
}

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Workgroup w of the grid -> (column, workgroup index within the column).  Workgroups are dealt
// round-robin to the 8 XCDs (w to XCD w % 8); with several columns, w = 8 (ncols r + c) + x takes
// column c of block 8 r + x, so all columns of a block run on one XCD one after another and the
// block's data are read from HBM once into that XCD's L2.
// One column with xcd_remap: each XCD works through one contiguous range of blocks (nvb, the
// number of blocks, is a multiple of 8).  Blocks are virtual: a persistent launch (xcd_remap 2) runs
// fewer workgroups, each taking blocks w, w + gridDim.x, ... (gridDim.x a multiple of 8, so every
// block stays on the XCD it would have had).
__device__ __forceinline__ void block_of(unsigned w, unsigned nvb, int ncols, int xcd_remap, unsigned& col,
                                         unsigned& lb) {
    const unsigned xcd = w % 8u, r = w / 8u;
    col = ncols > 1 ? r % unsigned(ncols) : 0u;
    lb = ncols > 1 ? (r / unsigned(ncols)) * 8u + xcd : (xcd_remap ? xcd * (nvb / 8u) + r : w);
}

// =====================================================================================  facets
// Column c of a facet wave holds side 0 of facet c (c < 8) or side 1 of facet 15 - c, so that the
// two sides of a facet sit in lanes j and 15 - j of the same 16-lane row and trade their traces
// with one DPP row mirror.
__device__ __forceinline__ int column_facet(int c) { return c < 8 ? c : 15 - c; }

__device__ __forceinline__ float mirror16(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x140, 0xf, 0xf, false));
}
__device__ __forceinline__ double mirror16(double v) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const int lo = __builtin_amdgcn_update_dpp(0, int(u & 0xffffffffu), 0x140, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, int(u >> 32), 0x140, 0xf, 0xf, false);
    return __builtin_bit_cast(double, (static_cast<unsigned long long>(unsigned(hi)) << 32) | unsigned(lo));
}

template <int D, int N, class T, int W, int OCC, bool VTAIL_ON, bool PERSIST>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(OCC)))
void elasticity_facet_kernel(SplitArgs<T> a, T const* __restrict__ xs, T* __restrict__ Ls, int64_t ld,
                             int64_t ldL, int ncols, unsigned nvb) {
    using L = MfmaLayout<D, N>;
    using acc = typename Mma<T>::acc;
    constexpr int Nbf = L::Nbf, nq = L::nq, Q = D, BS = Nbf * Q, NF = D + 1, DD = D * D;
    constexpr int KS = L::KS, MT = L::MT, FT = L::FT;
    constexpr int NFR = DD + D + 4; // tangents (D-1 columns), c_0, c_1, lam_0, mu_0, lam_1, mu_1
    constexpr int NbfP = Strides<T>::pad(Nbf), BSP = Strides<T>::pad(BS);
    constexpr int nT = NF * nq * Nbf;
    constexpr int TAIL = Nbf - 16 * (MT - 1);
    constexpr bool VTAIL = VTAIL_ON && D == 3 && MT > 1 && TAIL <= 4;
    constexpr int MM = VTAIL ? MT - 1 : MT;
    constexpr int NT = VTAIL ? TAIL : 1;
    constexpr int NI = (BS + 3) / 4; // staging steps of 4 entries per column

    __shared__ T sT[NF * nq * NbfP];
    __shared__ int sP[2 * NF * Nbf];
    __shared__ T sW[nq];
    __shared__ T sX[W][16 * BSP];

    // canonical tables, once per workgroup
    {
        const int tid = threadIdx.x;
#pragma unroll
        for (int i0 = 0; i0 < 2 * NF * Nbf; i0 += 64 * W) {
            const int i = i0 + tid;
            if (i < 2 * NF * Nbf) sP[i] = a.perm[i];
        }
#pragma unroll
        for (int i0 = 0; i0 < nT; i0 += 64 * W) {
            const int i = i0 + tid;
            if (i < nT) sT[(i / Nbf) * NbfP + i % Nbf] = a.T0[i];
        }
        if (tid < nq) sW[tid] = a.w[tid];
    }
    __syncthreads();

    auto block = [&](unsigned vb) {
        const int tid = int(threadIdx.x), wv = tid >> 6, lane = lane_of<PERSIST>(tid), j = lane & 15, g = lane >> 4;
        // this lane's column: facet fl, side sd
        const int fl = column_facet(j), sd = j < 8 ? 0 : 1;
        T* const X = sX[wv];
        unsigned col, lb;
        block_of(vb, nvb, ncols, a.xcd_remap, col, lb);
        const int64_t grp = int64_t(lb) * W + wv;
        const bool live = grp < a.num_fgroups; // uniform per wave
        const int64_t gj = live ? grp : 0;
        T const* __restrict__ x = xs + int64_t(col) * ld;
        T* __restrict__ Lc = Ls + int64_t(col) * ldL;

        const int4 info = a.FInfo[gj * 8 + fl];
        const T pen = a.FPen[gj * 8 + fl];
        const int kind = live ? (info.z >> 8) & 3 : 0;
        const bool two = kind == 1;
        const bool ok = sd == 0 ? kind != 0 : two; // this column carries a facet side
        const int el = ok ? (sd == 0 ? info.x : info.y) : 0;
        const int lf = ok ? (sd == 0 ? (info.z & 15) : ((info.z >> 4) & 15)) : 0;
        // record channels of this side: c_s and (lam_s, mu_s)
        const int chc = sd == 0 ? (D - 1) * D : DD, chl = DD + D + 2 * sd;
        T const* const recb = a.FR + gj * nq * NFR * 8 + fl;
        auto load_rec = [&](int t, T (&rc)[DD + 2]) {
            const int qc = min(4 * t + g, nq - 1);
            T const* r = recb + qc * NFR * 8;
#pragma unroll
            for (int c = 0; c < (D - 1) * D; ++c) rc[c] = r[c * 8];
#pragma unroll
            for (int c = 0; c < D; ++c) rc[(D - 1) * D + c] = r[(chc + c) * 8];
            rc[DD] = r[chl * 8];
            rc[DD + 1] = r[(chl + 1) * 8];
        };
        T rc[DD + 2];
        load_rec(0, rc);

        // element vector of this column in LDS, relabelled to face 0: U'(k', r) = x(perm_f(k'), r);
        // lane (j, g) moves entries 4 i + g
        {
            T v[NI];
            T const* xe = x + int64_t(el) * BS + g;
#pragma unroll
            for (int i = 0; i < NI; ++i) v[i] = (4 * i + 3 < BS || 4 * i + g < BS) ? xe[4 * i] : T(0);
            int const* const ip = sP + NF * Nbf + lf * Nbf;
#pragma unroll
            for (int i = 0; i < NI; ++i) {
                const int rem = 4 * i + g;
                if (4 * i + 3 < BS || rem < BS) {
                    const int r = rem / Nbf, k = rem - r * Nbf;
                    X[j * BSP + ip[k] + r * Nbf] = ok ? v[i] : T(0);
                }
            }
        }
        wave_sync(); // staged vectors ready (the region is this wave's)

        acc Y[MM][Q];
        T Yt[NT][Q];
#pragma unroll
        for (int m = 0; m < MM; ++m)
#pragma unroll
            for (int u = 0; u < Q; ++u) Y[m][u] = acc{0, 0, 0, 0};
#pragma unroll
        for (int kt = 0; kt < NT; ++kt)
#pragma unroll
            for (int u = 0; u < Q; ++u) Yt[kt][u] = T(0);

        if (live) {
            const int tq = mfma_group<T>(j);
            const int tc = min(mfma_reg<T>(j), D);
            T const* const U = X + j * BSP;

#pragma unroll 1
            for (int t = 0; t < FT; ++t) {
                const int q = 4 * t + g;
                const int qc = min(q, nq - 1);
                const bool act = ok && q < nq;
                T rn[DD + 2];
                load_rec(min(t + 1, FT - 1), rn); // next tile's record, one tile ahead
                // quadrature weight, zero on lanes without a facet point (every lane computes finite
                // values: padding facets carry the identity Jacobian, empty columns a zero vector)
                const T wq = act ? sW[qc] : T(0);
                const T ws = sd == 0 ? -wq : wq; // the traction term lifts with -w on side 0, +w on side 1
                acc C[Q];
#pragma unroll
                for (int r = 0; r < Q; ++r) C[r] = acc{0, 0, 0, 0};
                const int ab = (tc * nq + min(4 * t + tq, nq - 1)) * NbfP;
                // matrix-core phases run at raised wave priority, so that the SIMD's other wave's
                // point arithmetic fills the matrix core's busy cycles rather than delaying the MFMAs
                __builtin_amdgcn_s_setprio(1);
                {
                    T av = sT[ab + min(g, Nbf - 1)], b[Q];
#pragma unroll
                    for (int r = 0; r < Q; ++r) b[r] = U[min(g, Nbf - 1) + r * Nbf];
#pragma unroll
                    for (int s = 0; s < KS; ++s) {
                        const int k = 4 * s + g;
                        const bool kin = 4 * s + 3 < Nbf || k < Nbf;
                        T avn = av, bn[Q];
#pragma unroll
                        for (int r = 0; r < Q; ++r) bn[r] = b[r];
                        if (s + 1 < KS) {
                            const int kn = min(k + 4, Nbf - 1);
                            avn = sT[ab + kn];
#pragma unroll
                            for (int r = 0; r < Q; ++r) bn[r] = U[kn + r * Nbf];
                        }
#pragma unroll
                        for (int r = 0; r < Q; ++r) C[r] = Mma<T>::mma(av, kin ? b[r] : T(0), C[r]);
                        av = avn;
#pragma unroll
                        for (int r = 0; r < Q; ++r) b[r] = bn[r];
                    }
                }
                __builtin_amdgcn_s_setprio(0);
                // ---- lane (j, g): side sd of facet fl at point q; register c = value (0) / d/dxi_e (1 + e)
                // rows of G'_s = M_s^-1 with M_s = [t_1 .. t_(D-1), c_s], raw normal t_1 x t_2 (3D)
                T R[D][D], nr[D], dt;
                T const* cv = rc + (D - 1) * D;
                if constexpr (D == 3) {
                    T const* ta = rc;
                    T const* tb = rc + 3;
                    nr[0] = ta[1] * tb[2] - ta[2] * tb[1];
                    nr[1] = ta[2] * tb[0] - ta[0] * tb[2];
                    nr[2] = ta[0] * tb[1] - ta[1] * tb[0];
                    R[0][0] = tb[1] * cv[2] - tb[2] * cv[1];
                    R[0][1] = tb[2] * cv[0] - tb[0] * cv[2];
                    R[0][2] = tb[0] * cv[1] - tb[1] * cv[0];
                    R[1][0] = cv[1] * ta[2] - cv[2] * ta[1];
                    R[1][1] = cv[2] * ta[0] - cv[0] * ta[2];
                    R[1][2] = cv[0] * ta[1] - cv[1] * ta[0];
                    dt = cv[0] * nr[0] + cv[1] * nr[1] + cv[2] * nr[2];
                    const T id = rcp_(dt);
#pragma unroll
                    for (int i = 0; i < 3; ++i) {
                        R[0][i] *= id;
                        R[1][i] *= id;
                        R[2][i] = nr[i] * id;
                    }
                } else {
                    T const* ta = rc;
                    nr[0] = ta[1];
                    nr[1] = -ta[0];
                    dt = ta[0] * cv[1] - ta[1] * cv[0];
                    const T id = rcp_(dt);
                    R[0][0] = cv[1] * id;
                    R[0][1] = -cv[0] * id;
                    R[1][0] = -ta[1] * id;
                    R[1][1] = ta[0] * id;
                }
                // n outward from side 0: c_0 points into element 0, c_1 into element 1
                // (3D: nr . c_s = det_s; 2D: nr . c_s = -det_s)
                const bool flip = (D == 3 ? dt > T(0) : dt < T(0)) != (sd == 1);
                T n[D];
#pragma unroll
                for (int d = 0; d < D; ++d) n[d] = flip ? -nr[d] : nr[d];
                T nn = T(0);
#pragma unroll
                for (int d = 0; d < D; ++d) nn += n[d] * n[d];
                const T nl = sqrt_(nn);
                const T lam = rc[DD], mu = rc[DD + 1];
                // own side: u, physical gradient, sigma n
                T us[Q], sn[Q];
                {
                    T Jg[Q][D];
#pragma unroll
                    for (int r = 0; r < Q; ++r) {
                        us[r] = C[r][0];
#pragma unroll
                        for (int i = 0; i < D; ++i) {
                            T v = T(0);
#pragma unroll
                            for (int e = 0; e < D; ++e) v += R[e][i] * C[r][1 + e];
                            Jg[r][i] = v;
                        }
                    }
                    T tr = T(0);
#pragma unroll
                    for (int r = 0; r < D; ++r) tr += Jg[r][r];
#pragma unroll
                    for (int i = 0; i < Q; ++i) {
                        T v = T(0);
#pragma unroll
                        for (int jj = 0; jj < D; ++jj) v += (Jg[i][jj] + Jg[jj][i]) * n[jj];
                        sn[i] = mu * v + lam * tr * n[i];
                    }
                }
                // the other side's u and sigma n from the mirrored lane
                T tau[Q], dl[Q];
#pragma unroll
                for (int i = 0; i < Q; ++i) {
                    const T uo = mirror16(us[i]), so = mirror16(sn[i]);
                    const T jump = sd == 0 ? us[i] - uo : uo - us[i]; // u_0 - u_1
                    if (two) {
                        tau[i] = T(0.5) * (sn[i] + so) - pen * nl * jump;
                        dl[i] = T(-0.5) * jump;
                    } else {
                        tau[i] = sn[i] - pen * nl * us[i];
                        dl[i] = -us[i];
                    }
                }
                T dn = T(0);
#pragma unroll
                for (int d = 0; d < D; ++d) dn += dl[d] * n[d];
                T Wt[D + 1][Q];
#pragma unroll
                for (int u = 0; u < Q; ++u) Wt[0][u] = ws * tau[u];
#pragma unroll
                for (int e = 0; e < D; ++e)
#pragma unroll
                    for (int u = 0; u < Q; ++u) {
                        T v = T(0);
#pragma unroll
                        for (int jj = 0; jj < D; ++jj) {
                            T Cuj = mu * (dl[u] * n[jj] + dl[jj] * n[u]);
                            if (u == jj) Cuj += lam * dn;
                            v += R[e][jj] * Cuj;
                        }
                        Wt[1 + e][u] = wq * v;
                    }
                __builtin_amdgcn_s_setprio(1);
                // ---- lift: Y(k') += sum_(q,c) T0(c, q ; k') W(c, q)
                {
                    T av = sT[qc * NbfP + min(j, Nbf - 1)];
#pragma unroll
                    for (int c = 0; c <= D; ++c) {
#pragma unroll
                        for (int m = 0; m < MM; ++m) {
                            const int nx = c * MM + m + 1;
                            const T avn = nx < (D + 1) * MM
                                              ? sT[((nx / MM) * nq + qc) * NbfP + min(16 * (nx % MM) + j, Nbf - 1)]
                                              : av;
#pragma unroll
                            for (int u = 0; u < Q; ++u) Y[m][u] = Mma<T>::mma(av, Wt[c][u], Y[m][u]);
                            av = avn;
                        }
                        if constexpr (VTAIL) {
#pragma unroll
                            for (int kt = 0; kt < NT; ++kt) {
                                const T tv = sT[(c * nq + qc) * NbfP + 16 * MM + kt];
#pragma unroll
                                for (int u = 0; u < Q; ++u) Yt[kt][u] += tv * Wt[c][u];
                            }
                        }
                    }
                }
#pragma unroll
                for (int c = 0; c < DD + 2; ++c) rc[c] = rn[c];
            }
            __builtin_amdgcn_s_setprio(0);
        }

        // ---- lifted rows back to element labels through LDS, then one contiguous slot per column
        if constexpr (VTAIL) {
#pragma unroll
            for (int kt = 0; kt < NT; ++kt)
#pragma unroll
                for (int u = 0; u < Q; ++u) {
                    T v = Yt[kt][u];
                    v += __shfl_xor(v, 16, 64);
                    v += __shfl_xor(v, 32, 64);
                    Yt[kt][u] = v;
                }
        }
        wave_sync(); // every read of the staged vectors is complete
        {
            int const* const pf = sP + lf * Nbf;
#pragma unroll
            for (int m = 0; m < MM; ++m)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int k = 16 * m + mfma_row<T>(g, i);
                    if (k < Nbf) {
                        const int p = pf[k];
#pragma unroll
                        for (int u = 0; u < Q; ++u) X[j * BSP + p + u * Nbf] = Y[m][u][i];
                    }
                }
            if constexpr (VTAIL) {
                if (g == 0) {
#pragma unroll
                    for (int kt = 0; kt < NT; ++kt) {
                        const int p = pf[16 * MM + kt];
#pragma unroll
                        for (int u = 0; u < Q; ++u) X[j * BSP + p + u * Nbf] = Yt[kt][u];
                    }
                }
            }
        }
        wave_sync();
        if (ok) {
            T* const dst = Lc + (int64_t(el) * NF + lf) * BS + g;
#pragma unroll
            for (int i = 0; i < NI; ++i)
                if (4 * i + 3 < BS || 4 * i + g < BS) dst[4 * i] = X[j * BSP + 4 * i + g];
        }
        wave_sync(); // every read of the staged region is done before the next block restages it
    };
    if constexpr (PERSIST) {
        for (unsigned vb = blockIdx.x; vb < nvb; vb += gridDim.x) block(vb);
    } else {
        block(blockIdx.x);
    }
}

// =====================================================================================  elements
// DEEP (one wave per SIMD, up to 512 registers): the point data of block b + 1 are loaded during
// block b and the facet slots during the last block, so only the first loads of a wave wait on
// memory.  Otherwise (two or more waves per SIMD) each block loads its own point data up front and
// the slots are read at the end; the other waves of the SIMD cover the waits.
template <int D, int N, class T, int W, int OCC, bool DEEP, bool VTAIL_ON, bool PERSIST>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(OCC)))
void elasticity_element_kernel(SplitArgs<T> a, T const* __restrict__ xs, T const* __restrict__ Ls,
                               T* __restrict__ ys, int64_t ld, int64_t ldL, int ncols, unsigned nvb) {
    using L = MfmaLayout<D, N>;
    using acc = typename Mma<T>::acc;
    constexpr int Nbf = L::Nbf, Nq = L::Nq, Q = D, BS = Nbf * Q, DD = D * D, NF = D + 1;
    constexpr int KS = L::KS, QB = L::QB, MT = L::MT, NCH = L::NCH;
    constexpr int NbfP = Strides<T>::pad(Nbf), BSP = Strides<T>::pad(BS);
    constexpr int nV = D * Nq * Nbf;
    constexpr int TAIL = Nbf - 16 * (MT - 1);
    constexpr bool VTAIL = VTAIL_ON && D == 3 && MT > 1 && TAIL <= 4;
    constexpr int MM = VTAIL ? MT - 1 : MT;
    constexpr int NT = VTAIL ? TAIL : 1;
    constexpr int NO = (16 * BS + 63) / 64; // output entries per lane

    __shared__ T sV[D * Nq * NbfP];
    __shared__ T sU[W][16 * BSP];

    // the volume table, once per workgroup
#pragma unroll
    for (int i0 = 0; i0 < nV; i0 += 64 * W) {
        const int i = i0 + int(threadIdx.x);
        if (i < nV) sV[(i / Nbf) * NbfP + i % Nbf] = a.V[i];
    }
    __syncthreads();

    auto block = [&](unsigned vb) {
        const int tid = int(threadIdx.x), wv = tid >> 6, lane = lane_of<PERSIST>(tid), j = lane & 15, g = lane >> 4;
        T* const U = sU[wv];
        unsigned col, lb;
        block_of(vb, nvb, ncols, a.xcd_remap, col, lb);
        const int64_t grp = int64_t(lb) * W + wv;
        const bool live = grp < a.num_groups; // uniform per wave
        T const* __restrict__ x = xs + int64_t(col) * ld;
        T const* __restrict__ Lc = Ls + int64_t(col) * ldL;
        T* __restrict__ y = ys + int64_t(col) * ld;
        const int64_t e0 = grp * 16;
        const int nel = live ? int(min(int64_t(16), a.num_elements - e0)) : 0;
        const bool valid = j < nel;
        const int64_t gj = live ? grp : 0;
        T const* const xw = x + (live ? e0 : 0) * BS; // this wave's 16 element vectors (contiguous)
        const int nout = nel * BS;

        // point data of point q = 16 b + mfma_row(g, i) (clamped), for registers i0 <= i < i1
        auto load_pd = [&](int b, T (&pd)[4][NCH], int i0 = 0, int i1 = 4) {
#pragma unroll
            for (int i = i0; i < i1; ++i) {
                const int q = min(16 * b + mfma_row<T>(g, i), Nq - 1);
                T const* vd = a.VD + ((gj * Nq + q) * NCH) * 16 + j;
#pragma unroll
                for (int c = 0; c < NCH; ++c) pd[i][c] = vd[c * 16];
            }
        };
        // the element's facet slots for output entry i0 + lane (16 consecutive elements: contiguous)
        T const* const lsl = Lc + (live ? e0 : 0) * NF * BS;
        auto load_slots = [&](T (&ls)[NO]) {
#pragma unroll
            for (int o = 0; o < NO; ++o) {
                const int i = 64 * o + lane;
                const int ic = min(i, max(nout - 1, 0));
                const int jj = ic / BS, rem = ic - jj * BS;
                T v = T(0);
#pragma unroll
                for (int f = 0; f < NF; ++f) v += lsl[(jj * NF + f) * BS + rem];
                ls[o] = v;
            }
        };

        T pd[4][NCH];
        if constexpr (DEEP) load_pd(0, pd);
        if (live) {
            // every load first (clamped addresses), then the LDS stores: no load waits on another
            T v[NO];
#pragma unroll
            for (int o = 0; o < NO; ++o) v[o] = xw[min(64 * o + lane, nout - 1)];
#pragma unroll
            for (int o = 0; o < NO; ++o) {
                const int i = 64 * o + lane;
                if (i < 16 * BS) U[(i / BS) * BSP + i % BS] = i < nout ? v[o] : T(0);
            }
        }
        wave_sync(); // staged vectors ready (the region is this wave's)

        T const* const Uj = U + j * BSP;
        acc Y[MM][Q];
        T Yt[NT][Q];
#pragma unroll
        for (int m = 0; m < MM; ++m)
#pragma unroll
            for (int u = 0; u < Q; ++u) Y[m][u] = acc{0, 0, 0, 0};
#pragma unroll
        for (int kt = 0; kt < NT; ++kt)
#pragma unroll
            for (int u = 0; u < Q; ++u) Yt[kt][u] = T(0);
        T ls[NO];

        // late: the point data of registers 2 and 3 are loaded here, once the block's reference
        // gradients are on the matrix cores
        auto volume_block = [&](int b, T (&P)[4][NCH], bool late) {
            acc C[D][Q];
#pragma unroll
            for (int e = 0; e < D; ++e)
#pragma unroll
                for (int r = 0; r < Q; ++r) C[e][r] = acc{0, 0, 0, 0};
            const int qa = min(16 * b + j, Nq - 1) * NbfP;
#pragma unroll
            for (int s = 0; s < KS; ++s) {
                const int k = 4 * s + g;
                const int kc = min(k, Nbf - 1);
                T bu[Q];
#pragma unroll
                for (int r = 0; r < Q; ++r) {
                    const T v = Uj[kc + r * Nbf];
                    bu[r] = (4 * s + 3 < Nbf || k < Nbf) ? v : T(0);
                }
#pragma unroll
                for (int e = 0; e < D; ++e) {
                    const T av = sV[e * Nq * NbfP + qa + kc];
#pragma unroll
                    for (int r = 0; r < Q; ++r) C[e][r] = Mma<T>::mma(av, bu[r], C[e][r]);
                }
            }
            if (late) load_pd(b, P, 2, 4);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                if (16 * b + mfma_row<T>(0, i) >= Nq) continue; // all four lane groups past the last point
                const bool act = valid && 16 * b + mfma_row<T>(g, i) < Nq;
                T la[D][MM];
                {
                    const int qa2 = min(16 * b + mfma_row<T>(g, i), Nq - 1) * NbfP;
#pragma unroll
                    for (int e = 0; e < D; ++e)
#pragma unroll
                        for (int m = 0; m < MM; ++m) la[e][m] = sV[e * Nq * NbfP + qa2 + min(16 * m + j, Nbf - 1)];
                }
                T const* G = P[i];
                const T lw = P[i][DD], mw = P[i][DD + 1];
                T Ju[D][Q];
#pragma unroll
                for (int s2 = 0; s2 < D; ++s2)
#pragma unroll
                    for (int r = 0; r < Q; ++r) {
                        T v = T(0);
#pragma unroll
                        for (int e = 0; e < D; ++e) v += G[e + s2 * D] * C[e][r][i];
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
                            T Pv = mw * (Ju[u][jj] + Ju[jj][u]);
                            if (u == jj) Pv += lw * tr;
                            v += G[e + jj * D] * Pv;
                        }
                        C[e][u][i] = act ? v : T(0);
                    }
                const int qr = min(16 * b + mfma_row<T>(g, i), Nq - 1) * NbfP;
#pragma unroll
                for (int e = 0; e < D; ++e) {
#pragma unroll
                    for (int m = 0; m < MM; ++m) {
                        const T av = la[e][m];
#pragma unroll
                        for (int u = 0; u < Q; ++u) Y[m][u] = Mma<T>::mma(av, C[e][u][i], Y[m][u]);
                    }
                    if constexpr (VTAIL) {
#pragma unroll
                        for (int kt = 0; kt < NT; ++kt) {
                            const T tv = sV[e * Nq * NbfP + qr + 16 * MM + kt];
#pragma unroll
                            for (int u = 0; u < Q; ++u) Yt[kt][u] += tv * C[e][u][i];
                        }
                    }
                }
            }
        };

        if (live) {
            if constexpr (DEEP) {
#pragma unroll
                for (int b = 0; b < QB; ++b) {
                    T pn[4][NCH];
                    if (b + 1 < QB) load_pd(b + 1, pn);
                    else load_slots(ls);
                    volume_block(b, pd, false);
                    if (b + 1 < QB) {
#pragma unroll
                        for (int i = 0; i < 4; ++i)
#pragma unroll
                            for (int c = 0; c < NCH; ++c) pd[i][c] = pn[i][c];
                    }
                }
            } else {
#pragma unroll 1
                for (int b = 0; b < QB; ++b) {
                    // the point data are loaded before the matrix-core work that does not need them:
                    // FP32 all four registers; FP64 (fewer registers to spare) two, and the other two
                    // once the block's reference-gradient MFMAs are issued
                    constexpr bool half = sizeof(T) == 8;
                    load_pd(b, pd, 0, half ? 2 : 4);
                    __builtin_amdgcn_sched_barrier(0);
                    volume_block(b, pd, half);
                }
            }
        }

        if constexpr (VTAIL) {
#pragma unroll
            for (int kt = 0; kt < NT; ++kt)
#pragma unroll
                for (int u = 0; u < Q; ++u) {
                    T v = Yt[kt][u];
                    v += __shfl_xor(v, 16, 64);
                    v += __shfl_xor(v, 32, 64);
                    Yt[kt][u] = v;
                }
        }
        wave_sync(); // every read of this wave's U is complete (sU is per wave)
        if (live) {
#pragma unroll
            for (int m = 0; m < MM; ++m)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int k = 16 * m + mfma_row<T>(g, i);
                    if (k < Nbf) {
#pragma unroll
                        for (int u = 0; u < Q; ++u) U[j * BSP + k + u * Nbf] = Y[m][u][i];
                    }
                }
            if constexpr (VTAIL) {
                if (g == 0) {
#pragma unroll
                    for (int kt = 0; kt < NT; ++kt)
#pragma unroll
                        for (int u = 0; u < Q; ++u) U[j * BSP + 16 * MM + kt + u * Nbf] = Yt[kt][u];
                }
            }
            if constexpr (!DEEP) load_slots(ls);
        }
        // x again for x . y (one column only: Krylov solvers), all loads issued before the first use
        T xv[NO];
        if (a.dot && live) {
#pragma unroll
            for (int o = 0; o < NO; ++o) xv[o] = xw[min(64 * o + lane, nout - 1)];
        }
        wave_sync();
        double dot = 0.0;
        if (live) {
            // y = volume + the element's facet slots
            T* dst = y + e0 * BS;
#pragma unroll
            for (int o = 0; o < NO; ++o) {
                const int i = 64 * o + lane;
                if (i < nout) {
                    const int jj = i / BS, rem = i - jj * BS;
                    const T v = U[jj * BSP + rem] + ls[o];
                    dst[i] = v;
                    if (a.dot) dot += double(xv[o]) * double(v);
                }
            }
        }
        if (a.dot) {
            // one partial per block: the wave sums, then the waves' sums in a fixed order
            __shared__ double red[W];
            for (int o = 32; o > 0; o >>= 1) dot += __shfl_down(dot, o, 64);
            if (lane == 0) red[wv] = dot;
            __syncthreads();
            if (tid == 0) {
                double t = 0.0;
#pragma unroll
                for (int w = 0; w < W; ++w) t += red[w];
                a.dot[vb] = t;
            }
            __syncthreads(); // red is rewritten by the next block
        }
        wave_sync(); // every read of the staged region is done before the next block restages it
    };
    if constexpr (PERSIST) {
        for (unsigned vb = blockIdx.x; vb < nvb; vb += gridDim.x) block(vb);
    } else {
        block(blockIdx.x);
    }
}

// Launch shapes of the two kernels, searched by the autotuner: waves per workgroup and the
// occupancy (waves per SIMD) the register allocation targets, for each kernel; DEEP element kernel;
// the last basis tile of the lifts on the vector ALU (VTAIL) or on the matrix cores.
template <int FW, int FOCC, int EW, int EOCC, bool DEEP, bool VT> struct Shape {
    static constexpr int fw = FW, focc = FOCC, ew = EW, eocc = EOCC;
    static constexpr bool deep = DEEP, vtail = VT;
};

template <class S> int64_t element_workgroups(int64_t num_groups) {
    const int64_t eb = (num_groups + S::ew - 1) / S::ew;
    return (eb + 7) / 8 * 8;
}

// Workgroups a persistent launch of kernel k runs: as many as are resident on the device at once
// (a multiple of 8), at most nvb.
inline int64_t resident_workgroups(void const* k, int threads, int64_t nvb) {
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, threads, 0) != hipSuccess || cus <= 0 || per_cu <= 0)
        return nvb;
    const int64_t cap = int64_t(cus) * per_cu / 8 * 8;
    return cap >= 8 ? std::min(nvb, cap) : nvb;
}

template <int D, int N, class T, class S>
hipError_t run_shape(SplitArgs<T> const& a, T const* x, T* y, T* Lbuf, int ncols, int64_t ld, int64_t ldL,
                     hipStream_t s) {
    const int64_t fb = (a.num_fgroups + S::fw - 1) / S::fw;
    const int64_t ft = (fb + 7) / 8 * 8 * int64_t(ncols);
    const int64_t et = element_workgroups<S>(a.num_groups) * int64_t(ncols);
    if (ft > int64_t(0x7fffffff) || et > int64_t(0x7fffffff)) return hipErrorInvalidValue;
    if (a.xcd_remap == 2) {
        auto fk = elasticity_facet_kernel<D, N, T, S::fw, S::focc, S::vtail, true>;
        auto ek = elasticity_element_kernel<D, N, T, S::ew, S::eocc, S::deep, S::vtail, true>;
        if (ft > 0) {
            const int64_t fg = resident_workgroups(reinterpret_cast<void const*>(fk), 64 * S::fw, ft);
            hipLaunchKernelGGL(fk, dim3(unsigned(fg)), dim3(64 * S::fw), 0, s, a, x, Lbuf, ld, ldL, ncols,
                               unsigned(ft));
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) return e;
        }
        const int64_t eg = resident_workgroups(reinterpret_cast<void const*>(ek), 64 * S::ew, et);
        hipLaunchKernelGGL(ek, dim3(unsigned(eg)), dim3(64 * S::ew), 0, s, a, x, Lbuf, y, ld, ldL, ncols,
                           unsigned(et));
        return hipGetLastError();
    }
    if (ft > 0) {
        hipLaunchKernelGGL((elasticity_facet_kernel<D, N, T, S::fw, S::focc, S::vtail, false>), dim3(unsigned(ft)),
                           dim3(64 * S::fw), 0, s, a, x, Lbuf, ld, ldL, ncols, unsigned(ft));
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((elasticity_element_kernel<D, N, T, S::ew, S::eocc, S::deep, S::vtail, false>),
                       dim3(unsigned(et)), dim3(64 * S::ew), 0, s, a, x, Lbuf, y, ld, ldL, ncols, unsigned(et));
    return hipGetLastError();
}

// The shape table (index < kSplitShapes): FP64 kernels hold up to two waves per SIMD, FP32 more.
template <int D, int N, class T>
hipError_t run_split(int shape, SplitArgs<T> const& a, T const* x, T* y, T* Lbuf, int ncols, int64_t ld, int64_t ldL,
                     hipStream_t s) {
    if constexpr (sizeof(T) == 8) {
        switch (shape) {
        case 0: return run_shape<D, N, T, Shape<4, 2, 8, 2, false, true>>(a, x, y, Lbuf, ncols, ld, ldL, s);
        case 1: return run_shape<D, N, T, Shape<8, 2, 8, 2, false, true>>(a, x, y, Lbuf, ncols, ld, ldL, s);
        case 2: return run_shape<D, N, T, Shape<4, 2, 4, 1, true, true>>(a, x, y, Lbuf, ncols, ld, ldL, s);
        case 3: return run_shape<D, N, T, Shape<4, 2, 8, 2, false, false>>(a, x, y, Lbuf, ncols, ld, ldL, s);
        }
    } else {
        switch (shape) {
        case 0: return run_shape<D, N, T, Shape<4, 3, 4, 3, false, false>>(a, x, y, Lbuf, ncols, ld, ldL, s);
        case 1: return run_shape<D, N, T, Shape<4, 3, 4, 3, false, true>>(a, x, y, Lbuf, ncols, ld, ldL, s);
        case 2: return run_shape<D, N, T, Shape<4, 4, 4, 3, false, false>>(a, x, y, Lbuf, ncols, ld, ldL, s);
        case 3: return run_shape<D, N, T, Shape<8, 3, 8, 2, false, false>>(a, x, y, Lbuf, ncols, ld, ldL, s);
        }
    }
    return hipErrorInvalidValue;
}

template <class T> int64_t element_workgroups_of(int shape, int64_t num_groups) {
    if constexpr (sizeof(T) == 8) {
        switch (shape) {
        case 0: return element_workgroups<Shape<4, 2, 8, 2, false, true>>(num_groups);
        case 1: return element_workgroups<Shape<8, 2, 8, 2, false, true>>(num_groups);
        case 2: return element_workgroups<Shape<4, 2, 4, 1, true, true>>(num_groups);
        case 3: return element_workgroups<Shape<4, 2, 8, 2, false, false>>(num_groups);
        }
    } else {
        switch (shape) {
        case 0: return element_workgroups<Shape<4, 3, 4, 3, false, false>>(num_groups);
        case 1: return element_workgroups<Shape<4, 3, 4, 3, false, true>>(num_groups);
        case 2: return element_workgroups<Shape<4, 4, 4, 3, false, false>>(num_groups);
        case 3: return element_workgroups<Shape<8, 3, 8, 2, false, false>>(num_groups);
        }
    }
    return -1;
}

} // namespace split
} // namespace tdg
