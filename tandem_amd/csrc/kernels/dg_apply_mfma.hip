// Elasticity SIPG apply on the CDNA4 matrix cores with the reference tables and the element
// vectors staged in LDS.
//
// One wave owns 16 consecutive elements: the 16 columns of every 16x16x4 MFMA tile are these
// elements.  A block of kWaves waves first copies into LDS
//   sV[e][q][k]   = d phi_k / d xi_e at volume point q          (gradient and divergence operand)
//   sT[c][q][k]   = face-0 value (c = 0) / d/dxi_(c-1) at facet point q (trace and lift operand)
//   sP            = the facet relabellings perm[f][k] and their inverses,
//   sU[wave]      = the 16 element vectors of each wave (one contiguous slice of x; FP32 pads
//                   the table rows and the vectors to bank-friendly strides NbfP, BSP),
// so each MFMA A operand is one LDS read whatever the tile orientation, and B operands of the own
// element are LDS reads in any relabelled order.
//
//   volume   : C(e,q ; el, r) = sum_k V(e,q ; k) U(k ; el, r)                  rows = 16 points
//              per point in registers, in place: S(e,u) = G (lamWJ tr(Ju) I + muWJ (Ju + Ju^T))
//              Y(k ; el, u) += sum_(e,q) V(e,q ; k) S(e,q ; el, u)             K = 4 points
//   facet f  : own and neighbour traces share the face-0 operand T0; the own vector is read in
//              the face-f frame, U(perm_f(k')), the neighbour's in its face frame; the point
//              fluxes give W(c ; el, u) and Y(k ; el, u) += sum_(q,c) T0(c, iperm_f(k), q) W(q,c).
// Output rows of the trace tiles are ordered so that lane group g holds facet point 4t + g with
// channel c in register c (see mfma_group / mfma_reg).  Padded rows and K entries read a clamped,
// finite table entry; their B operand (K) is zero or their output row is dropped.  Loads use
// clamped, in-bounds addresses and the per-point arithmetic is branch-free (results of inactive
// points are discarded by selects), so the unrolled facet-point loop is one basic block in which
// the compiler can overlap the point arithmetic of one tile with the MFMAs of the next.
#include "kernels.h"
#include "mfma_layout.h"
#include "mfma_ops.h"

#include <type_traits>

namespace tdg {

namespace {

template <class T> __device__ __forceinline__ T rsqrt_(T v);
template <> __device__ __forceinline__ double rsqrt_<double>(double v) { return 1.0 / sqrt(v); }
template <> __device__ __forceinline__ float rsqrt_<float>(float v) { return 1.0f / sqrtf(v); }

constexpr int kWaves = 4;

// FP32: two waves per SIMD (the LDS footprint allows two blocks per CU); FP64: one.
template <int D, int N, class T>
__global__ __launch_bounds__(64 * kWaves) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 2 : 1)))
void elasticity_mfma_kernel(MfmaArgs<T> a, T const* __restrict__ xs, T* __restrict__ ys, int64_t ld, int ncols) {
    using L = MfmaLayout<D, N>;
    using acc = typename Mma<T>::acc;
    constexpr int Nbf = L::Nbf, Nq = L::Nq, nq = L::nq, Q = D, BS = Nbf * Q, DD = D * D;
    constexpr int KS = L::KS, QB = L::QB, MT = L::MT, FT = L::FT, NF = D + 1;
    constexpr int NCH = L::NCH, NFC = L::NFC;
    constexpr int nV = D * Nq * Nbf, nT = NF * nq * Nbf;
    // FP32 LDS strides of the table rows and of the element vectors: twice an odd number, so the
    // 32 lanes of a half wave that read 16 rows (or elements) at 2 consecutive offsets hit distinct
    // banks (measured: 30% fewer bank-conflict cycles than the unpadded 35 / 105, while multiples of
    // 4 such as 36 / 108 add conflicts).  FP64 keeps the unpadded layout.
    constexpr int NbfP = sizeof(T) == 4 ? (Nbf + 1) / 4 * 4 + 2 : Nbf;
    constexpr int BSP = sizeof(T) == 4 ? (BS + 1) / 4 * 4 + 2 : BS;
    constexpr int nVP = D * Nq * NbfP, nTP = NF * nq * NbfP;

    // FP32 stages the 16 neighbour vectors of each face in LDS, in the volume table's space once the
    // volume phase is done (contiguous loads; the FP32 kernel has no registers to hold them); FP64
    // keeps them in registers, gathered per lane
    constexpr bool STAGE = sizeof(T) == 4;
    constexpr int nVN = STAGE && kWaves * 16 * BSP > nVP ? kWaves * 16 * BSP : nVP;
    __shared__ T sVN[nVN];
    T* const sV = sVN;
    __shared__ T sT[nTP];
    __shared__ int sP[2 * NF * Nbf];
    __shared__ T sU[kWaves][16 * BSP];

    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63, j = lane & 15, g = lane >> 4;
    // Workgroups are dealt round-robin to the 8 XCDs (workgroup w to XCD w % 8).  Several columns:
    // workgroup w = 8 (ncols r + c) + x takes column c of block 8 r + x, so all columns of a block
    // run on one XCD one after another and its point and facet data are read from HBM once into
    // that XCD's L2 instead of once per column.  One column with the XCD variant: each XCD works
    // through one contiguous range of blocks.  The grid is a multiple of 8 workgroups.
    const unsigned w = blockIdx.x, xcd = w % 8u, r = w / 8u;
    const unsigned col = ncols > 1 ? r % unsigned(ncols) : 0u;
    const unsigned lb = ncols > 1 ? (r / unsigned(ncols)) * 8u + xcd
                                  : (a.xcd_remap ? xcd * (gridDim.x / 8u) + r : w);
    const int64_t grp = int64_t(lb) * kWaves + wv;
    const bool live = grp < a.num_groups; // uniform per wave
    T const* __restrict__ x = xs + int64_t(col) * ld;
    T* __restrict__ y = ys + int64_t(col) * ld;
    const int64_t e0 = grp * 16;
    const int nel = live ? int(min(int64_t(16), a.num_elements - e0)) : 0;

    // unrolled copies: every load of the preamble is in flight at once (a rolled loop waits for each)
#pragma unroll
    for (int i0 = 0; i0 < nV; i0 += 64 * kWaves) {
        const int i = i0 + tid;
        if (i < nV) sV[(i / Nbf) * NbfP + i % Nbf] = a.V[i];
    }
#pragma unroll
    for (int i0 = 0; i0 < nT; i0 += 64 * kWaves) {
        const int i = i0 + tid;
        if (i < nT) sT[(i / Nbf) * NbfP + i % Nbf] = a.T0[i];
    }
#pragma unroll
    for (int i0 = 0; i0 < 2 * NF * Nbf; i0 += 64 * kWaves) {
        const int i = i0 + tid;
        if (i < 2 * NF * Nbf) sP[i] = a.perm[i];
    }
    T* const U = sU[wv];
    {
        T const* src = x + e0 * BS;
        const int n = nel * BS;
#pragma unroll
        for (int i0 = 0; i0 < 16 * BS; i0 += 64) {
            const int i = i0 + lane;
            if (i < 16 * BS) U[(i / BS) * BSP + i % BS] = i < n ? src[i] : T(0);
        }
    }
    __syncthreads();

    const bool valid = j < nel;
    const int64_t gj = live ? grp : 0; // point data row of this lane (in bounds for every lane)
    T const* const Uj = U + j * BSP;    // element j: U(k, r) at Uj[k + r Nbf]
    // Output rows k of the lift: the first MM tiles of 16 on the matrix cores; when the last tile
    // holds at most 4 real rows (3D P3: 4 of 16, P4: 3 of 16) those rows are summed on the vector
    // ALU instead (each lane over its own K entries, then across the four lane groups), which
    // removes a third of the lift MFMAs, which would mostly multiply padding.
    constexpr int TAIL = Nbf - 16 * (MT - 1);
    constexpr bool VTAIL = D == 3 && MT > 1 && TAIL <= 4;
    constexpr int MM = VTAIL ? MT - 1 : MT;
    constexpr int NT = VTAIL ? TAIL : 1;
    acc Y[MT][Q];
    T Yt[NT][Q];
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int u = 0; u < Q; ++u) Y[m][u] = acc{0, 0, 0, 0};
#pragma unroll
    for (int kt = 0; kt < NT; ++kt)
#pragma unroll
        for (int u = 0; u < Q; ++u) Yt[kt][u] = T(0);

    // ------------------------------------------------------------------ volume
    // one 16-point block; LAST marks the final block, whose fully padded K steps are skipped
    auto volume_block = [&](int b, auto last) {
        constexpr bool LAST = decltype(last)::value;
        // point data of this lane's points q = 16 b + mfma_row(g, i), issued before the MFMAs
        T pd[4][NCH];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int q = min(16 * b + mfma_row<T>(g, i), Nq - 1);
            T const* vd = a.VD + ((gj * Nq + q) * NCH) * 16 + j;
#pragma unroll
            for (int c = 0; c < NCH; ++c) pd[i][c] = vd[c * 16];
        }
        acc C[D][Q];
#pragma unroll
        for (int e = 0; e < D; ++e)
#pragma unroll
            for (int r = 0; r < Q; ++r) C[e][r] = acc{0, 0, 0, 0};
        const int qa = min(16 * b + j, Nq - 1) * NbfP; // A row: point 16 b + j
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
        // stress at each point, in place (C[e][u][i] <- S(e, u) of point i), then that point's
        // share of the divergence: K step (e, i) covers points 16 b + mfma_row(kk, i), kk = 0..3
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            if (LAST && 16 * (QB - 1) + mfma_row<T>(0, i) >= Nq) continue; // compile-time: all padded
            const bool act = valid && 16 * b + mfma_row<T>(g, i) < Nq;
            // this K step's A operands, read before the point arithmetic they do not depend on
            T la[D][MM];
            {
                const int qa2 = min(16 * b + mfma_row<T>(g, i), Nq - 1) * NbfP;
#pragma unroll
                for (int e = 0; e < D; ++e)
#pragma unroll
                    for (int m = 0; m < MM; ++m) la[e][m] = sV[e * Nq * NbfP + qa2 + min(16 * m + j, Nbf - 1)];
            }
            T const* G = pd[i];
            const T lw = pd[i][DD], mw = pd[i][DD + 1];
            T Ju[D][Q]; // Ju[s][r] = du_r/dx_s
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
                        T P = mw * (Ju[u][jj] + Ju[jj][u]);
                        if (u == jj) P += lw * tr;
                        v += G[e + jj * D] * P;
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
                    // this lane's point is K entry mfma_row(g, i) of the step; its share of the
                    // tail rows (C = 0 on padded points)
#pragma unroll
                    for (int kt = 0; kt < NT; ++kt) {
                        const T a = sV[e * Nq * NbfP + qr + 16 * MM + kt];
#pragma unroll
                        for (int u = 0; u < Q; ++u) Yt[kt][u] += a * C[e][u][i];
                    }
                }
            }
        }
    };

    if (live) {
#pragma unroll 1
        for (int b = 0; b < QB - 1; ++b) volume_block(b, std::false_type{});
        volume_block(QB - 1, std::true_type{});
    }
    if constexpr (STAGE) __syncthreads(); // every read of the volume table is complete
    T* const NbU = sVN + (STAGE ? wv * 16 * BSP : 0); // this wave's 16 neighbour vectors (FP32)
    if (live) {
        // -------------------------------------------------------------- facets
        // trace A operand row rho = j: facet point 4 t + mfma_group(rho), channel mfma_reg(rho)
        const int tq = mfma_group<T>(j);
        const int tc = min(mfma_reg<T>(j), D);
#pragma unroll 1
        for (int f = 0; f < NF; ++f) {
            const int4 info = a.FI[(gj * NF + f) * 16 + j];
            const int kind = valid ? (info.x & 3) : 0;
            const int nbface = (info.x >> 4) & 3;
            const int64_t nb = kind == 1 ? int64_t(info.y) : 0;
            const T pen = a.FP[(gj * NF + f) * 16 + j];
            // own and neighbour B operand offsets in the face frames (FP32: packed in one register,
            // neighbour vectors staged in LDS; FP64: neighbour values gathered into registers)
            int pp[KS];
            T Nb[STAGE ? 1 : KS][Q];
            if constexpr (STAGE) {
                // element jj's neighbour index comes from lane jj; lanes without one read element 0
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier(); // reads of the previous face's neighbours are done
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                const int nb32 = int(nb);
#pragma unroll
                for (int d0 = 0; d0 < 16 * BS; d0 += 64) {
                    const int d = d0 + lane;
                    const int jj = min(d / BS, 15);
                    const uint32_t off = uint32_t(__shfl(nb32, jj, 64)) * uint32_t(BS) + uint32_t(d - jj * BS);
                    if (d < 16 * BS) NbU[jj * BSP + (d - jj * BS)] = x[off];
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
#pragma unroll
            for (int s = 0; s < KS; ++s) {
                const int k = 4 * s + g;
                const int kc = min(k, Nbf - 1);
                const int pk = sP[nbface * Nbf + kc];
                if constexpr (STAGE) {
                    pp[s] = sP[f * Nbf + kc] | (pk << 16);
                } else {
                    pp[s] = sP[f * Nbf + kc];
#pragma unroll
                    for (int r = 0; r < Q; ++r) {
                        const T v = x[nb * BS + pk + r * Nbf];
                        Nb[s][r] = (4 * s + 3 < Nbf || k < Nbf) ? v : T(0);
                    }
                }
            }
            T const* const Nj = NbU + j * BSP;
            // lift A operand columns: face-0 label of row k = 16 m + j
            int il[MT];
#pragma unroll
            for (int m = 0; m < MT; ++m) il[m] = sP[NF * Nbf + f * Nbf + min(16 * m + j, Nbf - 1)];
            int const* const iperm_f = sP + NF * Nbf + f * Nbf; // face-0 labels (tail rows read at use)
            T const* const fdf = a.FD + ((gj * NF + f) * nq) * NFC * 16 + j;
            // lift of tile t's point weights: Y(k) += sum_(q,c) T0(c, iperm_f(k), q) W(q, c)
            auto lift = [&](int t, T const (&W)[D + 1][Q]) {
                const int qc = min(4 * t + g, nq - 1);
                // A operands one (c, m) step ahead of their MFMAs
                T av = sT[qc * NbfP + il[0]];
#pragma unroll
                for (int c = 0; c <= D; ++c) {
#pragma unroll
                    for (int m = 0; m < MM; ++m) {
                        const int nx = c * MM + m + 1;
                        const T avn = nx < (D + 1) * MM ? sT[((nx / MM) * nq + qc) * NbfP + il[nx % MM]] : av;
#pragma unroll
                        for (int u = 0; u < Q; ++u) Y[m][u] = Mma<T>::mma(av, W[c][u], Y[m][u]);
                        av = avn;
                    }
                    if constexpr (VTAIL) {
#pragma unroll
                        for (int kt = 0; kt < NT; ++kt) {
                            const T a = sT[(c * nq + qc) * NbfP + iperm_f[min(16 * MM + kt, Nbf - 1)]];
#pragma unroll
                            for (int u = 0; u < Q; ++u) Yt[kt][u] += a * W[c][u];
                        }
                    }
                }
            };
#pragma unroll
            for (int t = 0; t < FT; ++t) {
                const int q = 4 * t + g;
                const int qc = min(q, nq - 1);
                const bool act = kind != 0 && q < nq;
                T fd[NFC];
#pragma unroll
                for (int c = 0; c < NFC; ++c) fd[c] = fdf[(qc * NFC + c) * 16];
                const T wq = a.w[qc];
                acc Co[Q], Cn[Q];
#pragma unroll
                for (int r = 0; r < Q; ++r) Co[r] = Cn[r] = acc{0, 0, 0, 0};
                const int ab = (tc * nq + min(4 * t + tq, nq - 1)) * NbfP;
                if constexpr (STAGE) {
#pragma unroll
                    for (int s = 0; s < KS; ++s) {
                        const int k = 4 * s + g;
                        const T av = sT[ab + min(k, Nbf - 1)];
#pragma unroll
                        for (int r = 0; r < Q; ++r) {
                            const bool kin = 4 * s + 3 < Nbf || k < Nbf;
                            const T v = Uj[(pp[s] & 0xffff) + r * Nbf];
                            const T vn = Nj[(pp[s] >> 16) + r * Nbf];
                            Co[r] = Mma<T>::mma(av, kin ? v : T(0), Co[r]);
                            Cn[r] = Mma<T>::mma(av, kin ? vn : T(0), Cn[r]);
                        }
                    }
                } else {
                    // LDS operands one K step ahead of the MFMAs that use them
                    T av = sT[ab + min(g, Nbf - 1)], bo[Q];
#pragma unroll
                    for (int r = 0; r < Q; ++r) bo[r] = Uj[pp[0] + r * Nbf];
#pragma unroll
                    for (int s = 0; s < KS; ++s) {
                        const int k = 4 * s + g;
                        const bool kin = 4 * s + 3 < Nbf || k < Nbf;
                        T avn = av, bon[Q];
#pragma unroll
                        for (int r = 0; r < Q; ++r) bon[r] = bo[r];
                        if (s + 1 < KS) {
                            avn = sT[ab + min(k + 4, Nbf - 1)];
#pragma unroll
                            for (int r = 0; r < Q; ++r) bon[r] = Uj[pp[s + 1] + r * Nbf];
                        }
#pragma unroll
                        for (int r = 0; r < Q; ++r) {
                            Co[r] = Mma<T>::mma(av, kin ? bo[r] : T(0), Co[r]);
                            Cn[r] = Mma<T>::mma(av, Nb[s][r], Cn[r]);
                        }
                        av = avn;
#pragma unroll
                        for (int r = 0; r < Q; ++r) bo[r] = bon[r];
                    }
                }
                // this lane: facet point q; register c = value (0) / d/dxi_e (1 + e)
                T const* G0 = fd;
                T const* G1 = fd + DD + 2;
                const T l0 = fd[DD], m0 = fd[DD + 1], l1 = fd[2 * DD + 2], m1 = fd[2 * DD + 3];
                T const* n = fd + 2 * DD + 4;
                T nn = T(0);
#pragma unroll
                for (int d = 0; d < D; ++d) nn += n[d] * n[d];
                const T inl = rsqrt_<T>(nn);
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
                T dl[Q], sig[Q][D];
                const bool skel = kind == 1;
#pragma unroll
                for (int i = 0; i < Q; ++i) {
                    dl[i] = skel ? T(0.5) * (u1[i] - u0[i]) : -u0[i];
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
                T dn = T(0);
#pragma unroll
                for (int r = 0; r < D; ++r) dn += dl[r] * n[r];
                T W[D + 1][Q];
#pragma unroll
                for (int u = 0; u < Q; ++u) {
                    T v = T(0);
#pragma unroll
                    for (int jj = 0; jj < D; ++jj) v += n[jj] * sig[u][jj];
                    W[0][u] = act ? -wq * v : T(0);
                }
#pragma unroll
                for (int e = 0; e < D; ++e)
#pragma unroll
                    for (int u = 0; u < Q; ++u) {
                        T v = T(0);
#pragma unroll
                        for (int jj = 0; jj < D; ++jj) {
                            T Cuj = m0 * (dl[u] * n[jj] + dl[jj] * n[u]);
                            if (u == jj) Cuj += l0 * dn;
                            v += G0[e + jj * D] * Cuj;
                        }
                        W[1 + e][u] = act ? wq * v : T(0);
                    }
                lift(t, W);
            }
        }
    }

    // ------------------------------------------------------------------ store through LDS
    if constexpr (VTAIL) {
        // sum the tail rows over the four lane groups (lanes j, j + 16, j + 32, j + 48)
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
    __syncthreads(); // every read of U is complete
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
    }
    __syncthreads();
    if (live) {
        T* dst = y + e0 * BS;
        const int n = nel * BS;
#pragma unroll
        for (int i0 = 0; i0 < 16 * BS; i0 += 64) {
            const int i = i0 + lane;
            if (i < n) dst[i] = U[(i / BS) * BSP + i % BS];
        }
    }
}

template <int D, int N, class T>
hipError_t run(MfmaArgs<T> const& a, T const* x, T* y, int ncols, int64_t ld, hipStream_t s) {
    const int64_t blocks = (a.num_groups + kWaves - 1) / kWaves;
    const int64_t total = (blocks + 7) / 8 * 8 * int64_t(ncols);
    if (total > int64_t(0x7fffffff)) return hipErrorInvalidValue;
    hipLaunchKernelGGL((elasticity_mfma_kernel<D, N, T>), dim3(unsigned(total)), dim3(64 * kWaves), 0, s, a, x, y, ld,
                       ncols);
    return hipGetLastError();
}

} // namespace

template <class T>
hipError_t launch_elasticity_mfma(int dim, int degree, MfmaArgs<T> const& a, T const* x, T* y, int ncols,
                                 int64_t ld, hipStream_t s) {
    if (dim == 3) {
        switch (degree) {
        case 2: return run<3, 2, T>(a, x, y, ncols, ld, s);
        case 3: return run<3, 3, T>(a, x, y, ncols, ld, s);
        case 4: return run<3, 4, T>(a, x, y, ncols, ld, s);
        }
    } else if (dim == 2) {
        switch (degree) {
        case 2: return run<2, 2, T>(a, x, y, ncols, ld, s);
        case 3: return run<2, 3, T>(a, x, y, ncols, ld, s);
        case 4: return run<2, 4, T>(a, x, y, ncols, ld, s);
        case 6: return run<2, 6, T>(a, x, y, ncols, ld, s);
        }
    }
    return hipErrorInvalidValue;
}

bool mfma_supported(int dim, int degree) {
    return (dim == 3 && degree >= 2 && degree <= 4) || (dim == 2 && (degree == 2 || degree == 3 || degree == 4 || degree == 6));
}

template hipError_t launch_elasticity_mfma<double>(int, int, MfmaArgs<double> const&, double const*, double*, int,
                                                  int64_t, hipStream_t);
template hipError_t launch_elasticity_mfma<float>(int, int, MfmaArgs<float> const&, float const*, float*, int,
                                                 int64_t, hipStream_t);

} // namespace tdg
