// Sensitivity of the elasticity bilinear form to the per-quadrature-point Lame fields:
//   g_lam(e, q) = d/d lam_Q(e, q) psi^T A(lam, mu) u,   g_mu likewise,
// the kernel of the adjoint gradient dJ/dm = -psi^T (dA/dm) u of a misfit J(u(m)) with
// A psi = dJ/du.  Strains of u and psi are recomputed from the element vectors, not stored.
//
// The bilinear form is the apply's (elasticity.py:201-234 summed over both elements of a facet):
//   volume   sum_q W|J| [lam' div u div psi + mu' (grad u + grad u^T) : grad psi]
//   facets   sum_q w [ -{sigma(psi)} n . [[u]] - {sigma(u)} n . [[psi]] + delta |n| [[psi]] . [[u]] ]
//            ({.} = average on skeleton facets, the element's own value on Dirichlet/Fault boundary
//            facets; natural boundary facets have no terms)
// with lam' = matE lam_nodal at the volume points, the facet traces matE_q lam_nodal of each side,
// delta = (p_0 + p_1)/4 (skeleton) or p_0 (boundary), p_s = C_s c1_s^2 / c0_s, c0 = min_i 2 mu_i,
// c1 = max_i (D lam_i + 2 mu_i) over the side's nodal coefficients, and
// lam_nodal = M^-1 matE^T diag(W|J|) lam_Q (Elasticity.cpp:104-292).  Each element computes every
// term that involves its own coefficients (its volume, its own facet traces, its own side's
// penalty), so no two workgroups write the same output.
//
// Layout: one workgroup of 256 lanes owns EPB elements, their u and psi blocks in LDS; lanes per
// volume point, per facet point, per nodal coefficient, as the thread-per-point apply.
#include "kernels.h"

namespace tdg {

namespace {

constexpr int binom_a(int n, int k) { return k == 0 ? 1 : binom_a(n - 1, k - 1) * n / k; }
constexpr int vol_points_a(int D, int N) {
    return D == 2 ? (N == 1 ? 4 : N == 2 ? 7 : N == 3 ? 15 : N == 4 ? 19 : N == 5 ? 28 : 37)
                  : (N == 1 ? 6 : N == 2 ? 14 : N == 3 ? 31 : 57);
}
constexpr int fct_points_a(int D, int N) { return D == 2 ? N + 1 : (N == 1 ? 4 : N == 2 ? 7 : N == 3 ? 15 : 19); }
constexpr int pow2_a(int n) {
    int p = 1;
    while (p < n) p *= 2;
    return p;
}

constexpr int kThreads = 256;

template <int D, int N, class T>
__global__ __launch_bounds__(kThreads) void material_gradient_kernel(ApplyArgs<T> a, GradArgs<T> g,
                                                                     T const* __restrict__ u, T const* __restrict__ psi,
                                                                     T* __restrict__ g_lam, T* __restrict__ g_mu) {
    constexpr int Nbf = binom_a(N + D, D), Nq = vol_points_a(D, N), nq = fct_points_a(D, N);
    constexpr int NF = D + 1, Q = D, BS = Nbf * Q, DD = D * D, C = D + 1;
    constexpr int QP = pow2_a(Nq), EPB = QP >= kThreads ? 1 : kThreads / QP;
    __shared__ T sU[EPB * BS], sP[EPB * BS];
    __shared__ T sV[EPB * Nq * 2];          // volume point sensitivities (lam', mu')
    __shared__ T sF[EPB * NF * nq * 3];     // facet point sensitivities (lam trace, mu trace, delta)
    __shared__ T sN[EPB * Nbf * 2];         // nodal sensitivities
    __shared__ T sY[EPB * Nbf * 2];         // M^-1 times the nodal sensitivities

    const int tid = threadIdx.x;
    const int64_t e0 = int64_t(blockIdx.x) * EPB;
    const int ne = int(min(int64_t(EPB), a.num_elements - e0));
    for (int i = tid; i < ne * BS; i += kThreads) {
        sU[i] = u[e0 * BS + i];
        sP[i] = psi[e0 * BS + i];
    }
    __syncthreads();

    // ---- volume points
    for (int item = tid; item < EPB * QP; item += kThreads) {
        const int el = item / QP, q = item % QP;
        if (el >= ne || q >= Nq) continue;
        const int64_t eg = e0 + el;
        T gu[D][Q], gp[D][Q]; // reference gradients
#pragma unroll
        for (int e = 0; e < D; ++e)
#pragma unroll
            for (int r = 0; r < Q; ++r) gu[e][r] = gp[e][r] = T(0);
        for (int l = 0; l < Nbf; ++l) {
#pragma unroll
            for (int e = 0; e < D; ++e) {
                const T t = a.tabV[(l * D + e) * Nq + q];
#pragma unroll
                for (int r = 0; r < Q; ++r) {
                    gu[e][r] += t * sU[el * BS + l + r * Nbf];
                    gp[e][r] += t * sP[el * BS + l + r * Nbf];
                }
            }
        }
        T G[DD];
#pragma unroll
        for (int i = 0; i < DD; ++i) G[i] = a.G[(eg * Nq + q) * DD + i];
        T Ju[Q][D], Jp[Q][D]; // d v_r / d x_s
#pragma unroll
        for (int r = 0; r < Q; ++r)
#pragma unroll
            for (int s = 0; s < D; ++s) {
                T vu = T(0), vp = T(0);
#pragma unroll
                for (int e = 0; e < D; ++e) {
                    vu += G[e + s * D] * gu[e][r];
                    vp += G[e + s * D] * gp[e][r];
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
        const T wj = g.jw[eg * Nq + q];
        sV[(el * Nq + q) * 2 + 0] = wj * du * dp;
        sV[(el * Nq + q) * 2 + 1] = wj * ss;
    }

    // ---- facet points (element's own traces; the jump with the neighbour's values)
    for (int item = tid; item < ne * NF * nq; item += kThreads) {
        const int el = item / (NF * nq), rem = item % (NF * nq), f = rem / nq, q = rem % nq;
        const int64_t eg = e0 + el;
        const int4 info = a.ef[eg * NF + f];
        T* out = &sF[((el * NF + f) * nq + q) * 3];
        const int kind = info.z & 3;
        if (kind == 0) {
            out[0] = out[1] = out[2] = T(0);
            continue;
        }
        const int side = (info.z >> 2) & 1, nbface = (info.z >> 4) & 3;
        const int64_t fct = info.x;
        const bool skel = kind == 1;
        // own value and reference gradient traces of u and psi (face-0 table, relabelled vector)
        T uo[Q], po[Q], gu[D][Q], gp[D][Q];
#pragma unroll
        for (int r = 0; r < Q; ++r) {
            uo[r] = po[r] = T(0);
#pragma unroll
            for (int e = 0; e < D; ++e) gu[e][r] = gp[e][r] = T(0);
        }
        for (int l = 0; l < Nbf; ++l) {
            const int k = a.perm[f * Nbf + l];
            T const* t = &a.tabF[(l * C) * nq + q];
#pragma unroll
            for (int r = 0; r < Q; ++r) {
                const T vu = sU[el * BS + k + r * Nbf], vp = sP[el * BS + k + r * Nbf];
                uo[r] += t[0] * vu;
                po[r] += t[0] * vp;
#pragma unroll
                for (int e = 0; e < D; ++e) {
                    gu[e][r] += t[(1 + e) * nq] * vu;
                    gp[e][r] += t[(1 + e) * nq] * vp;
                }
            }
        }
        // jumps u_K - u_nb (boundary: u_K)
        T ju[Q], jp[Q];
#pragma unroll
        for (int r = 0; r < Q; ++r) {
            ju[r] = uo[r];
            jp[r] = po[r];
        }
        if (skel) {
            T const* Ux = u + int64_t(info.y) * BS;
            T const* Px = psi + int64_t(info.y) * BS;
            for (int l = 0; l < Nbf; ++l) {
                const int k = a.perm[nbface * Nbf + l];
                const T t = a.tabF[(l * C) * nq + q];
#pragma unroll
                for (int r = 0; r < Q; ++r) {
                    ju[r] -= t * Ux[k + r * Nbf];
                    jp[r] -= t * Px[k + r * Nbf];
                }
            }
        }
        T Gp[DD];
#pragma unroll
        for (int i = 0; i < DD; ++i) Gp[i] = a.fG[((fct * 2 + side) * nq + q) * DD + i];
        T Ju[Q][D], Jp[Q][D];
#pragma unroll
        for (int r = 0; r < Q; ++r)
#pragma unroll
            for (int s = 0; s < D; ++s) {
                T vu = T(0), vp = T(0);
#pragma unroll
                for (int e = 0; e < D; ++e) {
                    vu += Gp[e + s * D] * gu[e][r];
                    vp += Gp[e + s * D] * gp[e][r];
                }
                Ju[r][s] = vu;
                Jp[r][s] = vp;
            }
        // the element's outward area-scaled normal
        T n[D], nn = T(0);
#pragma unroll
        for (int j = 0; j < D; ++j) {
            n[j] = side ? -a.fn[(fct * nq + q) * D + j] : a.fn[(fct * nq + q) * D + j];
            nn += n[j] * n[j];
        }
        T du = T(0), dp = T(0), jun = T(0), jpn = T(0), jj = T(0), su = T(0), sp = T(0);
#pragma unroll
        for (int r = 0; r < D; ++r) {
            du += Ju[r][r];
            dp += Jp[r][r];
            jun += ju[r] * n[r];
            jpn += jp[r] * n[r];
            jj += ju[r] * jp[r];
        }
#pragma unroll
        for (int r = 0; r < D; ++r)
#pragma unroll
            for (int j = 0; j < D; ++j) {
                su += jp[r] * (Ju[r][j] + Ju[j][r]) * n[j];
                sp += ju[r] * (Jp[r][j] + Jp[j][r]) * n[j];
            }
        const T w = a.w[q], h = skel ? T(0.5) : T(1);
        out[0] = w * h * (-dp * jun - du * jpn);
        out[1] = w * h * (-sp - su);
        out[2] = w * sqrt(nn) * jj;
    }
    __syncthreads();

    // ---- nodal sensitivities: volume and facet traces through matE, penalty at the extremal nodes
    for (int item = tid; item < ne * Nbf; item += kThreads) {
        const int el = item / Nbf, i = item % Nbf;
        const int64_t eg = e0 + el;
        T gl = T(0), gm = T(0);
        for (int q = 0; q < Nq; ++q) {
            const T t = g.tabE[i * Nq + q];
            gl += t * sV[(el * Nq + q) * 2 + 0];
            gm += t * sV[(el * Nq + q) * 2 + 1];
        }
        const int4 pi = g.pinfo[eg];          // argmin 2 mu, argmax (D lam + 2 mu)
        const T c0 = g.pc[eg * 2 + 0], c1 = g.pc[eg * 2 + 1];
        for (int f = 0; f < NF; ++f) {
            const int4 info = a.ef[eg * NF + f];
            const int kind = info.z & 3;
            if (kind == 0) continue;
            const int l = a.iperm[f * Nbf + i];
            T gd = T(0);
            for (int q = 0; q < nq; ++q) {
                const T t = a.tabF[(l * C) * nq + q];
                T const* s = &sF[((el * NF + f) * nq + q) * 3];
                gl += t * s[0];
                gm += t * s[1];
                gd += s[2];
            }
            const T gpen = gd * (kind == 1 ? T(0.25) : T(1));
            const T Cf = g.pcoef[eg * NF + f];
            if (i == pi.x) gm += gpen * (-Cf * c1 * c1 / (c0 * c0)) * T(2);
            if (i == pi.y) {
                gl += gpen * (T(2) * Cf * c1 / c0) * T(D);
                gm += gpen * (T(2) * Cf * c1 / c0) * T(2);
            }
        }
        sN[(el * Nbf + i) * 2 + 0] = gl;
        sN[(el * Nbf + i) * 2 + 1] = gm;
    }
    __syncthreads();
    // ---- through the projection: y = M^-1 g_nodal (M symmetric), then g_Q = W|J| matE y
    for (int item = tid; item < ne * Nbf; item += kThreads) {
        const int el = item / Nbf, k = item % Nbf;
        const int64_t eg = e0 + el;
        T const* Mi = g.Minv + eg * Nbf * Nbf;
        T yl = T(0), ym = T(0);
        for (int i = 0; i < Nbf; ++i) {
            const T m = Mi[k + i * Nbf];
            yl += m * sN[(el * Nbf + i) * 2 + 0];
            ym += m * sN[(el * Nbf + i) * 2 + 1];
        }
        sY[(el * Nbf + k) * 2 + 0] = yl;
        sY[(el * Nbf + k) * 2 + 1] = ym;
    }
    __syncthreads();
    for (int item = tid; item < ne * Nq; item += kThreads) {
        const int el = item / Nq, q = item % Nq;
        const int64_t eg = e0 + el;
        T vl = T(0), vm = T(0);
        for (int k = 0; k < Nbf; ++k) {
            const T t = g.tabE[k * Nq + q];
            vl += t * sY[(el * Nbf + k) * 2 + 0];
            vm += t * sY[(el * Nbf + k) * 2 + 1];
        }
        const T wj = g.jw[eg * Nq + q];
        g_lam[eg * Nq + q] = wj * vl;
        g_mu[eg * Nq + q] = wj * vm;
    }
}

template <int D, int N, class T>
hipError_t run_grad(ApplyArgs<T> const& a, GradArgs<T> const& g, T const* u, T const* psi, T* gl, T* gm,
                    hipStream_t s) {
    constexpr int QP = pow2_a(vol_points_a(D, N)), EPB = QP >= kThreads ? 1 : kThreads / QP;
    dim3 grid(unsigned((a.num_elements + EPB - 1) / EPB));
    hipLaunchKernelGGL((material_gradient_kernel<D, N, T>), grid, dim3(kThreads), 0, s, a, g, u, psi, gl, gm);
    return hipGetLastError();
}

} // namespace

template <class T>
hipError_t launch_material_gradient(int dim, int degree, ApplyArgs<T> const& a, GradArgs<T> const& g, T const* u,
                                    T const* psi, T* g_lam, T* g_mu, hipStream_t s) {
    if (dim == 3) {
        switch (degree) {
        case 1: return run_grad<3, 1, T>(a, g, u, psi, g_lam, g_mu, s);
        case 2: return run_grad<3, 2, T>(a, g, u, psi, g_lam, g_mu, s);
        case 3: return run_grad<3, 3, T>(a, g, u, psi, g_lam, g_mu, s);
        case 4: return run_grad<3, 4, T>(a, g, u, psi, g_lam, g_mu, s);
        }
    } else if (dim == 2) {
        switch (degree) {
        case 1: return run_grad<2, 1, T>(a, g, u, psi, g_lam, g_mu, s);
        case 2: return run_grad<2, 2, T>(a, g, u, psi, g_lam, g_mu, s);
        case 3: return run_grad<2, 3, T>(a, g, u, psi, g_lam, g_mu, s);
        case 4: return run_grad<2, 4, T>(a, g, u, psi, g_lam, g_mu, s);
        case 5: return run_grad<2, 5, T>(a, g, u, psi, g_lam, g_mu, s);
        case 6: return run_grad<2, 6, T>(a, g, u, psi, g_lam, g_mu, s);
        }
    }
    return hipErrorInvalidValue;
}

template hipError_t launch_material_gradient<double>(int, int, ApplyArgs<double> const&, GradArgs<double> const&,
                                                     double const*, double const*, double*, double*, hipStream_t);
template hipError_t launch_material_gradient<float>(int, int, ApplyArgs<float> const&, GradArgs<float> const&,
                                                    float const*, float const*, float*, float*, hipStream_t);

} // namespace tdg
