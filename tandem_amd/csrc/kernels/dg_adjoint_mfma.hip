// Adjoint material gradient on the CDNA4 matrix cores: the same sensitivities as the
// thread-per-point kernel (dg_adjoint.hip),
//   g_lam(e, q) = d/d lam_Q(e, q) psi^T A(lam, mu) u,   g_mu likewise,
// with every contraction over the basis or over points laid out as 16x16x4 MFMA tiles whose 16
// columns are the 16 elements of one wave, as in the apply (dg_apply_mfma.hip):
//
//   volume     Cu(e,q ; el, r) = sum_k V(e,q ; k) U(k ; el, r), Cp likewise for psi   rows = points
//              per point: s_lam = W|J| div u div psi, s_mu = W|J| (grad u + grad u^T) : grad psi
//              Y(k ; el, ch) += sum_q E(k, q) s_ch(q ; el)                             K = 4 points
//   facet f    own value/gradient traces of u and psi (face-0 table, own vector in the face-f frame),
//              neighbour value traces (rows ordered so that each lane gets the point of its own
//              trace tile), the point sensitivities s0, s1 and the penalty sum, and
//              Y(k ; el, ch) += sum_q T0(0, iperm_f(k), q) s_ch(q ; el)
//   penalty    the derivative of delta through the extremal nodal coefficients, added to Y at those
//              nodes
//   projection y = M_e^-1 Y per element (the element's own dense matrix: a VALU matvec, lanes over
//              (element, row), coalesced rows of M^-1), then g(q ; el, ch) = W|J| sum_k E(k, q) y(k)
//
// LDS: the volume gradient table, replaced by the face-0 trace table between the phases, the
// transposed material basis, the facet relabellings, and per wave the 16 u and psi vectors (later the
// nodal sensitivities, their projections and the output staging).  The number of waves per
// workgroup is the largest (up to 8) whose LDS fits one CU.
#include "kernels.h"
#include "mfma_layout.h"
#include "mfma_ops.h"

#include <climits>

namespace tdg {

namespace {

template <int D, int N, class T> struct GradShape {
    using L = MfmaLayout<D, N>;
    static constexpr int Nbf = L::Nbf, Nq = L::Nq, nq = L::nq, BS = Nbf * D, NF = D + 1;
    static constexpr int nV = D * Nq * Nbf, nT = NF * nq * Nbf, nE = Nq * Nbf;
    static constexpr int nTab = nV > nT ? nV : nT;
    static constexpr int fixed = int(sizeof(T)) * (nTab + nE) + int(sizeof(int)) * 2 * NF * Nbf;
    static constexpr int per_wave = int(sizeof(T)) * 2 * 16 * BS;
    static constexpr int fit = (160 * 1024 - fixed) / per_wave;
    // FP64 keeps one wave per SIMD (512 registers); FP32 up to two
    static constexpr int cap = sizeof(T) == 8 ? 4 : 8;
    static constexpr int waves = fit > cap ? cap : fit;
    static_assert(waves >= 1, "LDS of one wave exceeds the CU");
    static_assert(BS >= Nq, "output staging needs Nq <= Nbf D");
};

// register i of an accumulator tile, i not known at compile time (selects, no indexing)
template <class A> __device__ __forceinline__ auto reg_of(A const& v, int i) {
    return i == 0 ? v[0] : i == 1 ? v[1] : i == 2 ? v[2] : v[3];
}

template <int D, int N, class T>
__global__ __launch_bounds__((64 * GradShape<D, N, T>::waves)) void material_gradient_mfma_kernel(
    MfmaArgs<T> a, GradArgs<T> g, T const* __restrict__ u, T const* __restrict__ psi, T* __restrict__ g_lam,
    T* __restrict__ g_mu) {
    using L = MfmaLayout<D, N>;
    using S = GradShape<D, N, T>;
    using acc = typename Mma<T>::acc;
    constexpr int Nbf = L::Nbf, Nq = L::Nq, nq = L::nq, Q = D, BS = Nbf * Q, DD = D * D, NF = D + 1;
    constexpr int KS = L::KS, QB = L::QB, MT = L::MT, FT = L::FT, NCH = L::NCH, NFC = L::NFC;
    constexpr int W = S::waves, NV = (FT + 3) / 4; // neighbour value tiles of 16 facet points
    constexpr int nV = S::nV, nT = S::nT, nE = S::nE;

    __shared__ T sTab[S::nTab];        // V[e][q][k], then T0[c][q][k]
    __shared__ T sE[nE];               // E[q][k] material basis at the volume points
    __shared__ int sP[2 * NF * Nbf];   // perm_f, then iperm_f
    __shared__ T sW[W][2][16 * BS];    // per wave: u and psi of its 16 elements

    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63, j = lane & 15, gq = lane >> 4;
    // workgroup w runs on XCD w % 8; with the XCD variant each XCD takes one contiguous range
    const unsigned w = blockIdx.x, xcd = w % 8u, r8 = w / 8u;
    const unsigned lb = a.xcd_remap ? xcd * (gridDim.x / 8u) + r8 : w;
    const int64_t grp = int64_t(lb) * W + wv;
    const bool live = grp < a.num_groups; // uniform per wave
    const int64_t e0 = grp * 16;
    const int nel = live ? int(min(int64_t(16), a.num_elements - e0)) : 0;
    const bool valid = j < nel;
    const int64_t gj = live ? grp : 0;        // grouped point data of this lane (in bounds)
    const int64_t ej = valid ? e0 + j : 0;    // element-major data of this lane's element (in bounds)

#pragma unroll
    for (int i0 = 0; i0 < nV; i0 += 64 * W) {
        const int i = i0 + tid;
        if (i < nV) sTab[i] = a.V[i];
    }
#pragma unroll
    for (int i0 = 0; i0 < nE; i0 += 64 * W) {
        const int i = i0 + tid;
        if (i < nE) sE[i] = g.tabEt[i];
    }
#pragma unroll
    for (int i0 = 0; i0 < 2 * NF * Nbf; i0 += 64 * W) {
        const int i = i0 + tid;
        if (i < 2 * NF * Nbf) sP[i] = a.perm[i];
    }
    T* const Uw = sW[wv][0];
    T* const Pw = sW[wv][1];
    {
        const int n = nel * BS;
#pragma unroll
        for (int i0 = 0; i0 < 16 * BS; i0 += 64) {
            const int i = i0 + lane;
            if (i < 16 * BS) {
                Uw[i] = i < n ? u[e0 * BS + i] : T(0);
                Pw[i] = i < n ? psi[e0 * BS + i] : T(0);
            }
        }
    }
    __syncthreads();

    T const* const Uj = Uw + j * BS; // element j: U(k, r) at Uj[k + r Nbf]
    T const* const Pj = Pw + j * BS;
    acc Y[MT][2]; // nodal sensitivities, rows k = 16 m + mfma_row, columns = elements; lam, mu
#pragma unroll
    for (int m = 0; m < MT; ++m) Y[m][0] = Y[m][1] = acc{0, 0, 0, 0};

    // ------------------------------------------------------------------ volume
    if (live) {
#pragma unroll 1
        for (int b = 0; b < QB; ++b) {
            T wj[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) wj[i] = g.jw[ej * Nq + min(16 * b + mfma_row<T>(gq, i), Nq - 1)];
            acc Cu[D][Q], Cp[D][Q];
#pragma unroll
            for (int e = 0; e < D; ++e)
#pragma unroll
                for (int r = 0; r < Q; ++r) Cu[e][r] = Cp[e][r] = acc{0, 0, 0, 0};
            const int qa = min(16 * b + j, Nq - 1) * Nbf; // A row: point 16 b + j
#pragma unroll
            for (int s = 0; s < KS; ++s) {
                const int k = 4 * s + gq, kc = min(k, Nbf - 1);
                const bool kin = 4 * s + 3 < Nbf || k < Nbf;
                T bu[Q], bp[Q];
#pragma unroll
                for (int r = 0; r < Q; ++r) {
                    bu[r] = kin ? Uj[kc + r * Nbf] : T(0);
                    bp[r] = kin ? Pj[kc + r * Nbf] : T(0);
                }
#pragma unroll
                for (int e = 0; e < D; ++e) {
                    const T av = sTab[e * Nq * Nbf + qa + kc];
#pragma unroll
                    for (int r = 0; r < Q; ++r) {
                        Cu[e][r] = Mma<T>::mma(av, bu[r], Cu[e][r]);
                        Cp[e][r] = Mma<T>::mma(av, bp[r], Cp[e][r]);
                    }
                }
            }
            // register i: point 16 b + mfma_row(g, i); K step i of the nodal reduction
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const bool act = valid && 16 * b + mfma_row<T>(gq, i) < Nq;
                T G[DD];
                {
                    const int q = min(16 * b + mfma_row<T>(gq, i), Nq - 1);
                    T const* vd = a.VD + ((gj * Nq + q) * NCH) * 16 + j;
#pragma unroll
                    for (int c = 0; c < DD; ++c) G[c] = vd[c * 16];
                }
                T Ju[Q][D], Jp[Q][D]; // d v_r / d x_s
#pragma unroll
                for (int r = 0; r < Q; ++r)
#pragma unroll
                    for (int s = 0; s < D; ++s) {
                        T vu = T(0), vp = T(0);
#pragma unroll
                        for (int e = 0; e < D; ++e) {
                            vu += G[e + s * D] * Cu[e][r][i];
                            vp += G[e + s * D] * Cp[e][r][i];
                        }
                        Ju[r][s] = vu;
                        Jp[r][s] = vp;
                    }
                T du = T(0), dp = T(0), ss = T(0);
#pragma unroll
                for (int r = 0; r < D; ++r) {
                    du += Ju[r][r];
                    dp += Jp[r][r];
                }
#pragma unroll
                for (int r = 0; r < D; ++r)
#pragma unroll
                    for (int s = 0; s < D; ++s) ss += (Ju[r][s] + Ju[s][r]) * Jp[r][s];
                const T sl = act ? wj[i] * du * dp : T(0);
                const T sm = act ? wj[i] * ss : T(0);
                const int qr = min(16 * b + mfma_row<T>(gq, i), Nq - 1) * Nbf;
#pragma unroll
                for (int m = 0; m < MT; ++m) {
                    const T av = sE[qr + min(16 * m + j, Nbf - 1)];
                    Y[m][0] = Mma<T>::mma(av, sl, Y[m][0]);
                    Y[m][1] = Mma<T>::mma(av, sm, Y[m][1]);
                }
            }
        }
    }
    __syncthreads(); // every read of the volume table is complete
#pragma unroll
    for (int i0 = 0; i0 < nT; i0 += 64 * W) {
        const int i = i0 + tid;
        if (i < nT) sTab[i] = a.T0[i];
    }
    __syncthreads();

    // ------------------------------------------------------------------ facets
    // penalty derivatives of this lane's element: mu at argmin 2 mu_i, lam and mu at argmax D lam + 2 mu
    T pmx = T(0), ply = T(0), pmy = T(0);
    const int4 pinf = g.pinfo[ej];
    if (live) {
        const T c0 = g.pc[ej * 2 + 0], c1 = g.pc[ej * 2 + 1];
        // own trace tiles: A row j = facet point 4 t + mfma_group(j), channel mfma_reg(j)
        const int tq = mfma_group<T>(j), tc = min(mfma_reg<T>(j), D);
        // neighbour value tiles v: A row j = facet point 16 v + vq, so that register i of lane group
        // g holds point 4 (4 v + i) + g, the point of own trace tile 4 v + i in that lane
        const int vq = 4 * mfma_reg<T>(j) + mfma_group<T>(j);
#pragma unroll 1
        for (int f = 0; f < NF; ++f) {
            const int4 info = a.FI[(gj * NF + f) * 16 + j];
            const int kind = valid ? (info.x & 3) : 0;
            const int nbface = (info.x >> 4) & 3;
            const bool skel = kind == 1;
            const int64_t nb = skel ? int64_t(info.y) : 0;
            int po[KS], pn[KS];
#pragma unroll
            for (int s = 0; s < KS; ++s) {
                const int kc = min(4 * s + gq, Nbf - 1);
                po[s] = sP[f * Nbf + kc];
                pn[s] = sP[nbface * Nbf + kc];
            }
            // neighbour value traces of u and psi (B operands gathered into registers)
            acc Vn[2][NV][Q];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                T const* X = (h == 0 ? u : psi) + nb * BS;
                T Nb[KS][Q];
#pragma unroll
                for (int s = 0; s < KS; ++s) {
                    const int k = 4 * s + gq;
                    const bool kin = 4 * s + 3 < Nbf || k < Nbf;
#pragma unroll
                    for (int r = 0; r < Q; ++r) {
                        const T v = X[pn[s] + r * Nbf];
                        Nb[s][r] = kin ? v : T(0);
                    }
                }
#pragma unroll
                for (int v = 0; v < NV; ++v) {
#pragma unroll
                    for (int r = 0; r < Q; ++r) Vn[h][v][r] = acc{0, 0, 0, 0};
                    const int ab = min(16 * v + vq, nq - 1) * Nbf; // value channel
#pragma unroll
                    for (int s = 0; s < KS; ++s) {
                        const T av = sTab[ab + min(4 * s + gq, Nbf - 1)];
#pragma unroll
                        for (int r = 0; r < Q; ++r) Vn[h][v][r] = Mma<T>::mma(av, Nb[s][r], Vn[h][v][r]);
                    }
                }
            }
            int il[MT]; // lift A operand columns: face-0 label of row k = 16 m + j
#pragma unroll
            for (int m = 0; m < MT; ++m) il[m] = sP[NF * Nbf + f * Nbf + min(16 * m + j, Nbf - 1)];
            T const* const fdf = a.FD + ((gj * NF + f) * nq) * NFC * 16 + j;
            T gd = T(0);
#pragma unroll 1
            for (int t = 0; t < FT; ++t) {
                const int q = 4 * t + gq, qc = min(q, nq - 1);
                const bool act = kind != 0 && q < nq;
                T G0[DD], n[D];
#pragma unroll
                for (int c = 0; c < DD; ++c) G0[c] = fdf[(qc * NFC + c) * 16];
#pragma unroll
                for (int d = 0; d < D; ++d) n[d] = fdf[(qc * NFC + 2 * DD + 4 + d) * 16];
                const T wq = a.w[qc];
                acc Cu[Q], Cp[Q];
#pragma unroll
                for (int r = 0; r < Q; ++r) Cu[r] = Cp[r] = acc{0, 0, 0, 0};
                const int ab = (tc * nq + min(4 * t + tq, nq - 1)) * Nbf;
#pragma unroll
                for (int s = 0; s < KS; ++s) {
                    const int k = 4 * s + gq;
                    const bool kin = 4 * s + 3 < Nbf || k < Nbf;
                    const T av = sTab[ab + min(k, Nbf - 1)];
#pragma unroll
                    for (int r = 0; r < Q; ++r) {
                        const T bu = Uj[po[s] + r * Nbf], bp = Pj[po[s] + r * Nbf];
                        Cu[r] = Mma<T>::mma(av, kin ? bu : T(0), Cu[r]);
                        Cp[r] = Mma<T>::mma(av, kin ? bp : T(0), Cp[r]);
                    }
                }
                // this lane: facet point q; register c = value (0) / d/dxi_e (1 + e)
                T ju[Q], jp[Q], Ju[Q][D], Jp[Q][D];
#pragma unroll
                for (int r = 0; r < Q; ++r) {
                    T nu = T(0), np = T(0); // neighbour values: register t % 4 of value tile t / 4
#pragma unroll
                    for (int v = 0; v < NV; ++v) {
                        if (v == t / 4) {
                            nu = reg_of(Vn[0][v][r], t % 4);
                            np = reg_of(Vn[1][v][r], t % 4);
                        }
                    }
                    ju[r] = Cu[r][0] - (skel ? nu : T(0));
                    jp[r] = Cp[r][0] - (skel ? np : T(0));
#pragma unroll
                    for (int s = 0; s < D; ++s) {
                        T vu = T(0), vp = T(0);
#pragma unroll
                        for (int e = 0; e < D; ++e) {
                            vu += G0[e + s * D] * Cu[r][1 + e];
                            vp += G0[e + s * D] * Cp[r][1 + e];
                        }
                        Ju[r][s] = vu;
                        Jp[r][s] = vp;
                    }
                }
                T nn = T(0), du = T(0), dp = T(0), jun = T(0), jpn = T(0), jj = T(0), su = T(0), sp = T(0);
#pragma unroll
                for (int r = 0; r < D; ++r) {
                    nn += n[r] * n[r];
                    du += Ju[r][r];
                    dp += Jp[r][r];
                    jun += ju[r] * n[r];
                    jpn += jp[r] * n[r];
                    jj += ju[r] * jp[r];
                }
#pragma unroll
                for (int r = 0; r < D; ++r)
#pragma unroll
                    for (int s = 0; s < D; ++s) {
                        su += jp[r] * (Ju[r][s] + Ju[s][r]) * n[s];
                        sp += ju[r] * (Jp[r][s] + Jp[s][r]) * n[s];
                    }
                const T wh = skel ? T(0.5) * wq : wq;
                const T s0 = act ? wh * (-dp * jun - du * jpn) : T(0);
                const T s1 = act ? wh * (-sp - su) : T(0);
                gd += act ? wq * sqrt(nn) * jj : T(0);
                // Y(k) += T0(0, iperm_f(k), q) s(q): K step = the tile's 4 points
#pragma unroll
                for (int m = 0; m < MT; ++m) {
                    const T av = sTab[qc * Nbf + il[m]];
                    Y[m][0] = Mma<T>::mma(av, s0, Y[m][0]);
                    Y[m][1] = Mma<T>::mma(av, s1, Y[m][1]);
                }
            }
            gd += __shfl_xor(gd, 16, 64);
            gd += __shfl_xor(gd, 32, 64);
            if (kind != 0) {
                const T gpen = skel ? T(0.25) * gd : gd;
                const T Cf = g.pcoef[ej * NF + f];
                pmx += gpen * (-Cf * c1 * c1 / (c0 * c0)) * T(2);
                ply += gpen * (T(2) * Cf * c1 / c0) * T(D);
                pmy += gpen * (T(2) * Cf * c1 / c0) * T(2);
            }
        }
    }
    __syncthreads(); // every read of u and psi is complete

    // ------------------------------------------------------------------ nodal sensitivities
    T* const gN = Uw; // [el][ch][k]
    T* const yN = Pw; // [el][ch][k]
    if (live) {
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int k = 16 * m + mfma_row<T>(gq, i);
                if (k < Nbf) {
                    gN[(j * 2 + 0) * Nbf + k] = Y[m][0][i];
                    gN[(j * 2 + 1) * Nbf + k] = Y[m][1][i];
                }
            }
    }
    __syncthreads();
    if (live && valid && gq == 0) {
        gN[(j * 2 + 1) * Nbf + pinf.x] += pmx;
        gN[(j * 2 + 0) * Nbf + pinf.y] += ply;
        gN[(j * 2 + 1) * Nbf + pinf.y] += pmy;
    }
    __syncthreads();
    // y = M_e^-1 gN: lane per (element, row k); row k of M^-1 read along k (M^-1 symmetric)
    if (live) {
        for (int it = lane; it < 16 * Nbf; it += 64) {
            const int el = it / Nbf, k = it % Nbf;
            T yl = T(0), ym = T(0);
            if (el < nel) {
                T const* Mi = g.Minv + (e0 + el) * Nbf * Nbf + k;
                T const* gl = gN + (el * 2) * Nbf;
#pragma unroll
                for (int i = 0; i < Nbf; ++i) {
                    const T m = Mi[i * Nbf];
                    yl += m * gl[i];
                    ym += m * gl[Nbf + i];
                }
            }
            yN[(el * 2 + 0) * Nbf + k] = yl;
            yN[(el * 2 + 1) * Nbf + k] = ym;
        }
    }
    __syncthreads();

    // ------------------------------------------------------------------ back to the points
    // g(q ; el, ch) = W|J| sum_k E(k, q) y(k ; el, ch), staged in LDS for a contiguous store
    T* const out = Uw; // [el][q]
#pragma unroll 1
    for (int ch = 0; ch < 2; ++ch) {
        if (live) {
#pragma unroll 1
            for (int b = 0; b < QB; ++b) {
                acc C = acc{0, 0, 0, 0};
                const int qa = min(16 * b + j, Nq - 1) * Nbf;
#pragma unroll
                for (int s = 0; s < KS; ++s) {
                    const int k = 4 * s + gq, kc = min(k, Nbf - 1);
                    const T bv = yN[(j * 2 + ch) * Nbf + kc];
                    C = Mma<T>::mma(sE[qa + kc], k < Nbf ? bv : T(0), C);
                }
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int q = 16 * b + mfma_row<T>(gq, i);
                    if (q < Nq) out[j * Nq + q] = C[i];
                }
            }
        }
        __syncthreads();
        if (live) {
            T* dst = (ch == 0 ? g_lam : g_mu) + e0 * Nq;
            T const* jw = g.jw + e0 * Nq;
            const int n = nel * Nq;
            for (int i = lane; i < n; i += 64) dst[i] = jw[i] * out[i];
        }
        __syncthreads();
    }
}

template <int D, int N, class T>
hipError_t run(MfmaArgs<T> const& a, GradArgs<T> const& g, T const* u, T const* psi, T* gl, T* gm, hipStream_t s) {
    constexpr int W = GradShape<D, N, T>::waves;
    const int64_t blocks = (a.num_groups + W - 1) / W;
    const int64_t total = (blocks + 7) / 8 * 8;
    if (total > int64_t(INT_MAX)) return hipErrorInvalidValue;
    hipLaunchKernelGGL((material_gradient_mfma_kernel<D, N, T>), dim3(unsigned(total)), dim3(64 * W), 0, s, a, g, u,
                       psi, gl, gm);
    return hipGetLastError();
}

} // namespace

template <class T>
hipError_t launch_material_gradient_mfma(int dim, int degree, MfmaArgs<T> const& a, GradArgs<T> const& g,
                                         T const* u, T const* psi, T* g_lam, T* g_mu, hipStream_t s) {
    if (a.num_elements == 0) return hipSuccess;
    if (dim == 3) {
        switch (degree) {
        case 2: return run<3, 2, T>(a, g, u, psi, g_lam, g_mu, s);
        case 3: return run<3, 3, T>(a, g, u, psi, g_lam, g_mu, s);
        case 4: return run<3, 4, T>(a, g, u, psi, g_lam, g_mu, s);
        }
    } else if (dim == 2) {
        switch (degree) {
        case 2: return run<2, 2, T>(a, g, u, psi, g_lam, g_mu, s);
        case 3: return run<2, 3, T>(a, g, u, psi, g_lam, g_mu, s);
        case 4: return run<2, 4, T>(a, g, u, psi, g_lam, g_mu, s);
        case 6: return run<2, 6, T>(a, g, u, psi, g_lam, g_mu, s);
        }
    }
    return hipErrorInvalidValue;
}

template hipError_t launch_material_gradient_mfma<double>(int, int, MfmaArgs<double> const&, GradArgs<double> const&,
                                                          double const*, double const*, double*, double*,
                                                          hipStream_t);
template hipError_t launch_material_gradient_mfma<float>(int, int, MfmaArgs<float> const&, GradArgs<float> const&,
                                                         float const*, float const*, float*, float*, hipStream_t);

} // namespace tdg
