// Oracle: Elasticity and Poisson local operators (precompute, matrix-free apply, assembly, rhs).
// Test infrastructure only; restates app/localoperator/{Elasticity,Poisson}.cpp and the
// contractions of app/kernels/{elasticity,poisson}.py element by element.
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <limits>
#include <stdexcept>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace oracle {

// InverseInequality<D>::trace_constant (src/form/InverseInequality.h:27-29)
static double trace_constant(unsigned D, unsigned N) { return (N + 1) * (N + D) / static_cast<double>(D); }

// transpose helper: (a,b,p) -> (b,a,p) for D x D blocks
static void transpose_blocks(unsigned D, double const* A, double* B, std::size_t np) {
    for (std::size_t p = 0; p < np; ++p)
        for (unsigned i = 0; i < D; ++i)
            for (unsigned j = 0; j < D; ++j) B[p * D * D + i + j * D] = A[p * D * D + j + i * D];
}

// E(k,p) -> E_T(p,k)
static std::vector<double> transpose2(std::vector<double> const& E, std::size_t n0, std::size_t n1) {
    std::vector<double> T(E.size());
    for (std::size_t a = 0; a < n0; ++a)
        for (std::size_t b = 0; b < n1; ++b) T[b + a * n1] = E[a + b * n0];
    return T;
}
// G(k,e,q) -> G120(e,q,k)
static std::vector<double> perm120(std::vector<double> const& G, std::size_t nk, std::size_t ne,
                                   std::size_t nqp) {
    std::vector<double> T(G.size());
    for (std::size_t k = 0; k < nk; ++k)
        for (std::size_t e = 0; e < ne; ++e)
            for (std::size_t q = 0; q < nqp; ++q) T[e + ne * (q + nqp * k)] = G[k + nk * (e + ne * q)];
    return T;
}

// L2 projection of a coefficient to the nodal material space
// (project_material_lhs/rhs, elasticity.py:36-43; Elasticity.cpp:104-164).
static void project_coefficient(unsigned Nbf, unsigned Nq, std::vector<double> const& matE_Q_T,
                                double const* W, double const* J, double const* f_Q, double* out) {
    std::vector<double> M(Nbf * Nbf, 0.0);
    for (unsigned k = 0; k < Nbf; ++k)
        for (unsigned l = 0; l < Nbf; ++l) {
            double s = 0.0;
            for (unsigned q = 0; q < Nq; ++q)
                s += matE_Q_T[q + k * Nq] * W[q] * J[q] * matE_Q_T[q + l * Nq];
            M[k + l * Nbf] = s;
        }
    for (unsigned k = 0; k < Nbf; ++k) {
        double s = 0.0;
        for (unsigned q = 0; q < Nq; ++q) s += f_Q[q] * matE_Q_T[q + k * Nq] * W[q] * J[q];
        out[k] = s;
    }
    lu_solve(M, Nbf, out, 1);
}

// ============================================================================== Elasticity
// Elasticity ctor + prepare_* (Elasticity.cpp:27-64, 86-292)
void Elasticity::setup(Mesh const& mesh, unsigned N, std::vector<double> const* nodes,
                       double const* lam_Q, double const* mu_Q) {
    c.init(mesh, N, nodes);
    unsigned D = c.D;
    Q = D;
    degree = N;
    space = RefElement(D, N, false);
    matSpace = RefElement(D, N, false);
    Nbf = space.nbf;
    Nq = c.volRule.size();
    nq = c.fctRule.size();
    E_Q = space.eval_basis(c.volRule.pts);
    Dxi_Q = space.eval_grad(c.volRule.pts);
    Dxi_Q_120 = perm120(Dxi_Q, Nbf, D, Nq);
    matE_Q_T = transpose2(matSpace.eval_basis(c.volRule.pts), matSpace.nbf, Nq);
    E_q.clear();
    E_q_T.clear();
    Dxi_q.clear();
    Dxi_q_120.clear();
    matE_q_T.clear();
    for (unsigned f = 0; f <= D; ++f) {
        auto pts = c.cl.facet_param(f, c.fctRule.pts);
        E_q.push_back(space.eval_basis(pts));
        E_q_T.push_back(transpose2(E_q.back(), Nbf, nq));
        Dxi_q.push_back(space.eval_grad(pts));
        Dxi_q_120.push_back(perm120(Dxi_q.back(), Nbf, D, nq));
        matE_q_T.push_back(transpose2(matSpace.eval_basis(pts), matSpace.nbf, nq));
    }
    std::size_t ne = c.nE(), nf = c.nF();
    unsigned DD = D * D;
    lam.assign(ne * Nbf, 0.0);
    mu.assign(ne * Nbf, 0.0);
    JInvT.assign(ne * Nq * DD, 0.0);
#pragma omp parallel for schedule(static)
    for (std::size_t el = 0; el < ne; ++el) {
        project_coefficient(Nbf, Nq, matE_Q_T, c.volRule.wts.data(), &c.absDetJ[el * Nq],
                            &lam_Q[el * Nq], &lam[el * Nbf]);
        project_coefficient(Nbf, Nq, matE_Q_T, c.volRule.wts.data(), &c.absDetJ[el * Nq],
                            &mu_Q[el * Nq], &mu[el * Nbf]);
        transpose_blocks(D, &c.JInv[el * Nq * DD], &JInvT[el * Nq * DD], Nq);
    }
    lam_q0.assign(nf * nq, 0.0);
    mu_q0.assign(nf * nq, 0.0);
    lam_q1.assign(nf * nq, 0.0);
    mu_q1.assign(nf * nq, 0.0);
    JInvT0.assign(nf * nq * DD, 0.0);
    JInvT1.assign(nf * nq * DD, 0.0);
#pragma omp parallel for schedule(static)
    for (std::size_t f = 0; f < nf; ++f) {
        auto const& info = c.topo.fct[f];
        bool skel = info.up[0] != info.up[1];
        // precomputeSurface (elasticity.py:85-88)
        for (int side = 0; side < 2; ++side) {
            int s = skel ? side : 0; // boundary: side 1 mirrors side 0
            auto const& mT = matE_q_T[info.localNo[s]];
            for (unsigned q = 0; q < nq; ++q) {
                double a = 0.0, b = 0.0;
                for (unsigned t = 0; t < Nbf; ++t) {
                    a += mT[q + t * nq] * lam[info.up[s] * Nbf + t];
                    b += mT[q + t * nq] * mu[info.up[s] * Nbf + t];
                }
                (side == 0 ? lam_q0 : lam_q1)[f * nq + q] = a;
                (side == 0 ? mu_q0 : mu_q1)[f * nq + q] = b;
            }
        }
        transpose_blocks(D, &c.JInv0[f * nq * DD], &JInvT0[f * nq * DD], nq);
        transpose_blocks(D, skel ? &c.JInv1[f * nq * DD] : &c.JInv0[f * nq * DD],
                         &JInvT1[f * nq * DD], nq);
    }
    // prepare_volume_post_skeleton (precomputeVolume, elasticity.py:45-49)
    lamWJ.assign(ne * Nq, 0.0);
    muWJ.assign(ne * Nq, 0.0);
#pragma omp parallel for schedule(static)
    for (std::size_t el = 0; el < ne; ++el)
        for (unsigned q = 0; q < Nq; ++q) {
            double a = 0.0, b = 0.0;
            for (unsigned t = 0; t < Nbf; ++t) {
                a += matE_Q_T[q + t * Nq] * lam[el * Nbf + t];
                b += matE_Q_T[q + t * Nq] * mu[el * Nbf + t];
            }
            double wj = c.volRule.wts[q] * c.absDetJ[el * Nq + q];
            lamWJ[el * Nq + q] = a * wj;
            muWJ[el * Nq + q] = b * wj;
        }
    // prepare_penalty (Elasticity.cpp:255-292)
    penalty.assign(nf, 0.0);
    double cN1 = trace_constant(D, N - 1);
    auto p = [&](std::size_t f, int side) {
        auto el = c.topo.fct[f].up[side];
        double c0 = std::numeric_limits<double>::max(), c1 = std::numeric_limits<double>::lowest();
        for (unsigned i = 0; i < Nbf; ++i) {
            c0 = std::min(c0, 2.0 * mu[el * Nbf + i]);
            c1 = std::max(c1, D * lam[el * Nbf + i] + 2.0 * mu[el * Nbf + i]);
        }
        return (D + 1) * cN1 * (c.area[f] / c.volume[el]) * (c1 * c1 / c0);
    };
    for (std::size_t f = 0; f < nf; ++f) {
        auto const& info = c.topo.fct[f];
        penalty[f] = info.up[0] != info.up[1] ? (p(f, 0) + p(f, 1)) / 4.0 : p(f, 0);
    }
}

// Elasticity::apply_<false> (Elasticity.cpp:698-856) with the contractions of
// elasticity.py:201-234 (apply_volume, flux_u_*, flux_sigma_*, apply_facet).
void Elasticity::apply_element(std::size_t elNo, double const* x, double* y) const {
    unsigned D = c.D, DD = D * D, bs = Nbf * Q;
    double const* U = x + elNo * bs;
    double* Unew = y + elNo * bs;
    // ---- apply_volume: Ju_Q(q,s,r) = G_Q_T(s,e,q) Dxi_Q_120(e,q,l) U(l,r)
    std::vector<double> S(D * Q * Nq, 0.0); // S(e,u,q) = G(e,j,q) sigma_WJ(u,j,q)
    for (unsigned q = 0; q < Nq; ++q) {
        double gref[3][3]; // gref[e][r]
        for (unsigned e = 0; e < D; ++e)
            for (unsigned r = 0; r < Q; ++r) {
                double s = 0.0;
                for (unsigned l = 0; l < Nbf; ++l) s += Dxi_Q_120[e + D * (q + Nq * l)] * U[l + r * Nbf];
                gref[e][r] = s;
            }
        double const* GT = &JInvT[(elNo * Nq + q) * DD];
        double const* G = &c.JInv[(elNo * Nq + q) * DD];
        double Ju[3][3]; // Ju[s][r]
        for (unsigned s = 0; s < D; ++s)
            for (unsigned r = 0; r < Q; ++r) {
                double v = 0.0;
                for (unsigned e = 0; e < D; ++e) v += GT[s + e * D] * gref[e][r];
                Ju[s][r] = v;
            }
        double tr = 0.0;
        for (unsigned r = 0; r < D; ++r) tr += Ju[r][r];
        double lw = lamWJ[elNo * Nq + q], mw = muWJ[elNo * Nq + q];
        double P[3][3]; // P[u][j]
        for (unsigned u = 0; u < Q; ++u)
            for (unsigned j = 0; j < D; ++j) P[u][j] = (u == j ? lw * tr : 0.0) + mw * (Ju[u][j] + Ju[j][u]);
        for (unsigned e = 0; e < D; ++e)
            for (unsigned u = 0; u < Q; ++u) {
                double v = 0.0;
                for (unsigned j = 0; j < D; ++j) v += G[e + j * D] * P[u][j];
                S[e + D * (u + Q * q)] = v;
            }
    }
    for (unsigned k = 0; k < Nbf; ++k)
        for (unsigned u = 0; u < Q; ++u) {
            double v = 0.0;
            for (unsigned q = 0; q < Nq; ++q)
                for (unsigned e = 0; e < D; ++e) v += Dxi_Q[k + Nbf * (e + D * q)] * S[e + D * (u + Q * q)];
            Unew[k + u * Nbf] = v;
        }
    // ---- facets
    std::vector<double> uhmu(nq * Q), sig(DD * nq), nflip(nq * D), nuflip(nq * D);
    for (unsigned f = 0; f <= D; ++f) {
        auto const& si = c.topo.side[elNo * (D + 1) + f];
        bool skel = static_cast<std::size_t>(si.lid) != elNo;
        bool fd = si.bc == BC::Fault || si.bc == BC::Dirichlet;
        std::size_t fct = si.fctNo;
        double const* n = &c.normal[fct * nq * D];
        double const* nu = &c.unitNormal[fct * nq * D];
        double const* l0 = &lam_q0[fct * nq];
        double const* l1 = &lam_q1[fct * nq];
        double const* m0 = &mu_q0[fct * nq];
        double const* m1 = &mu_q1[fct * nq];
        double const* G0 = &JInvT0[fct * nq * DD];
        double const* G1 = &JInvT1[fct * nq * DD];
        if (skel && si.side == 1) {
            std::swap(l0, l1);
            std::swap(m0, m1);
            std::swap(G0, G1);
            for (unsigned i = 0; i < nq * D; ++i) {
                nflip[i] = -n[i];
                nuflip[i] = -nu[i];
            }
            n = nflip.data();
            nu = nuflip.data();
        }
        std::fill(uhmu.begin(), uhmu.end(), 0.0);
        std::fill(sig.begin(), sig.end(), 0.0);
        auto const& Ef = E_q_T[f];
        auto const& Df = Dxi_q_120[f];
        // own traces
        auto trace = [&](std::vector<double> const& ET, std::vector<double> const& D120,
                         double const* Ux, double const* GT, unsigned q, double* u, double Ju[3][3]) {
            for (unsigned i = 0; i < Q; ++i) {
                double s = 0.0;
                for (unsigned l = 0; l < Nbf; ++l) s += ET[q + nq * l] * Ux[l + i * Nbf];
                u[i] = s;
            }
            double gref[3][3];
            for (unsigned e = 0; e < D; ++e)
                for (unsigned r = 0; r < Q; ++r) {
                    double s = 0.0;
                    for (unsigned l = 0; l < Nbf; ++l) s += D120[e + D * (q + nq * l)] * Ux[l + r * Nbf];
                    gref[e][r] = s;
                }
            // Ju_q(q,r,s) = G_q_T(s,e,q) gref(e,r)
            for (unsigned r = 0; r < Q; ++r)
                for (unsigned s = 0; s < D; ++s) {
                    double v = 0.0;
                    for (unsigned e = 0; e < D; ++e) v += GT[q * DD + s + e * D] * gref[e][r];
                    Ju[r][s] = v;
                }
        };
        auto constitutive = [&](double Ju[3][3], double lq, double mq, double sg[3][3]) {
            double tr = 0.0;
            for (unsigned r = 0; r < D; ++r) tr += Ju[r][r];
            for (unsigned i = 0; i < D; ++i)
                for (unsigned j = 0; j < D; ++j) sg[i][j] = (i == j ? lq * tr : 0.0) + mq * (Ju[i][j] + Ju[j][i]);
        };
        if (si.bc == BC::None || (skel && fd)) {
            double const* Ue = x + si.lid * bs;
            auto const& Eo = E_q_T[si.localNo];
            auto const& Do = Dxi_q_120[si.localNo];
            double c00 = -penalty[fct];
            for (unsigned q = 0; q < nq; ++q) {
                double u0[3], u1[3], J0[3][3], J1[3][3], s0[3][3], s1[3][3];
                trace(Ef, Df, U, G0, q, u0, J0);
                trace(Eo, Do, Ue, G1, q, u1, J1);
                for (unsigned i = 0; i < Q; ++i) uhmu[q * Q + i] = 0.5 * (-u0[i] + u1[i]);
                constitutive(J0, l0[q], m0[q], s0);
                constitutive(J1, l1[q], m1[q], s1);
                for (unsigned i = 0; i < D; ++i)
                    for (unsigned j = 0; j < D; ++j)
                        sig[i + D * (j + D * q)] =
                            0.5 * (s0[i][j] + s1[i][j]) + c00 * (u0[i] - u1[i]) * nu[q * D + j];
            }
        } else if (fd) {
            double c00 = -penalty[fct];
            for (unsigned q = 0; q < nq; ++q) {
                double u0[3], J0[3][3], s0[3][3];
                trace(Ef, Df, U, G0, q, u0, J0);
                for (unsigned i = 0; i < Q; ++i) uhmu[q * Q + i] = -u0[i];
                constitutive(J0, l0[q], m0[q], s0);
                for (unsigned i = 0; i < D; ++i)
                    for (unsigned j = 0; j < D; ++j)
                        sig[i + D * (j + D * q)] = s0[i][j] + c00 * u0[i] * nu[q * D + j];
            }
        } else {
            continue;
        }
        // apply_facet
        auto const& Dq = Dxi_q[f];
        auto const& Eq = E_q[f];
        for (unsigned q = 0; q < nq; ++q) {
            double w = c.fctRule.wts[q];
            double dn = 0.0;
            for (unsigned r = 0; r < D; ++r) dn += uhmu[q * Q + r] * n[q * D + r];
            double C[3][3]; // C[u][j]
            for (unsigned u = 0; u < Q; ++u)
                for (unsigned j = 0; j < D; ++j)
                    C[u][j] = (u == j ? l0[q] * dn : 0.0) +
                              m0[q] * (uhmu[q * Q + u] * n[q * D + j] + uhmu[q * Q + j] * n[q * D + u]);
            // Z[e][u] = G_q_T0(j,e,q) C[u][j]
            double Z[3][3];
            for (unsigned e = 0; e < D; ++e)
                for (unsigned u = 0; u < Q; ++u) {
                    double v = 0.0;
                    for (unsigned j = 0; j < D; ++j) v += G0[q * DD + j + e * D] * C[u][j];
                    Z[e][u] = v;
                }
            double V[3];
            for (unsigned u = 0; u < Q; ++u) {
                double v = 0.0;
                for (unsigned j = 0; j < D; ++j) v += n[q * D + j] * sig[u + D * (j + D * q)];
                V[u] = v;
            }
            for (unsigned k = 0; k < Nbf; ++k)
                for (unsigned u = 0; u < Q; ++u) {
                    double v = 0.0;
                    for (unsigned e = 0; e < D; ++e) v += Dq[k + Nbf * (e + D * q)] * Z[e][u];
                    Unew[k + u * Nbf] += w * v + w * (-Eq[k + Nbf * q]) * V[u];
                }
        }
    }
}

void Elasticity::apply(double const* x, double* y, int nthreads) const {
    std::size_t ne = c.nE();
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
    for (std::size_t el = 0; el < ne; ++el) apply_element(el, x, y);
}

// Dx(k,i,p) = G(e,i,p) Dxi(k,e,p)  (Dx_Q / Dx_q, elasticity.py:53, 99-100)
static std::vector<double> physical_grad(unsigned D, unsigned Nbf, unsigned np, double const* G,
                                         std::vector<double> const& Dxi) {
    std::vector<double> Dx(Nbf * D * np, 0.0);
    for (unsigned p = 0; p < np; ++p)
        for (unsigned i = 0; i < D; ++i)
            for (unsigned k = 0; k < Nbf; ++k) {
                double s = 0.0;
                for (unsigned e = 0; e < D; ++e) s += G[p * D * D + e + i * D] * Dxi[k + Nbf * (e + D * p)];
                Dx[k + Nbf * (i + D * p)] = s;
            }
    return Dx;
}

// DGOperator::assemble with Elasticity::assemble_volume/skeleton/boundary (Elasticity.cpp:317-519;
// assembleVolume, assembleTractionOp, lift_ip, assembleSurface in elasticity.py:55-57, 123-149).
void Elasticity::assemble_dense(double* A) const {
    unsigned D = c.D, bs = Nbf * Q;
    std::size_t ne = c.nE(), n = bs * ne;
    std::fill(A, A + n * n, 0.0);
    auto at = [&](std::size_t re, std::size_t ce, unsigned k, unsigned p, unsigned l, unsigned u) -> double& {
        return A[(re * bs + k + p * Nbf) + n * (ce * bs + l + u * Nbf)];
    };
    for (std::size_t el = 0; el < ne; ++el) {
        auto Dx = physical_grad(D, Nbf, Nq, &c.JInv[el * Nq * D * D], Dxi_Q);
        for (unsigned q = 0; q < Nq; ++q) {
            double lw = lamWJ[el * Nq + q], mw = muWJ[el * Nq + q];
            for (unsigned k = 0; k < Nbf; ++k)
                for (unsigned p = 0; p < Q; ++p)
                    for (unsigned l = 0; l < Nbf; ++l)
                        for (unsigned u = 0; u < Q; ++u) {
                            double v = lw * Dx[l + Nbf * (u + D * q)] * Dx[k + Nbf * (p + D * q)];
                            double dd = 0.0;
                            for (unsigned j = 0; j < D; ++j)
                                dd += Dx[k + Nbf * (j + D * q)] *
                                      ((p == u ? Dx[l + Nbf * (j + D * q)] : 0.0) +
                                       (j == u ? Dx[l + Nbf * (p + D * q)] : 0.0));
                            at(el, el, k, p, l, u) += v + mw * dd;
                        }
        }
    }
    for (std::size_t f = 0; f < c.nF(); ++f) {
        auto const& info = c.topo.fct[f];
        bool skel = info.up[0] != info.up[1];
        if (!skel && (info.bc == BC::Natural)) continue;
        int ns = skel ? 2 : 1;
        std::vector<double> Dx[2], top[2];
        double const* lq[2] = {&lam_q0[f * nq], &lam_q1[f * nq]};
        double const* mq[2] = {&mu_q0[f * nq], &mu_q1[f * nq]};
        double const* n = &c.normal[f * nq * D];
        for (int s = 0; s < ns; ++s) {
            double const* g = s == 0 ? &c.JInv0[f * nq * D * D] : &c.JInv1[f * nq * D * D];
            Dx[s] = physical_grad(D, Nbf, nq, g, Dxi_q[info.localNo[s]]);
            // traction_op_q(k,p,u,q)
            top[s].assign(Nbf * D * D * nq, 0.0);
            for (unsigned q = 0; q < nq; ++q)
                for (unsigned k = 0; k < Nbf; ++k)
                    for (unsigned p = 0; p < D; ++p)
                        for (unsigned u = 0; u < D; ++u) {
                            double nd = 0.0;
                            for (unsigned j = 0; j < D; ++j) nd += n[q * D + j] * Dx[s][k + Nbf * (j + D * q)];
                            double v = lq[s][q] * Dx[s][k + Nbf * (p + D * q)] * n[q * D + u] +
                                       mq[s][q] * (nd * (p == u ? 1.0 : 0.0) +
                                                   Dx[s][k + Nbf * (u + D * q)] * n[q * D + p]);
                            top[s][k + Nbf * (p + D * (u + D * q))] = v;
                        }
        }
        double pen = penalty[f];
        double c0[2], c1[2], c2[2];
        if (skel) {
            c0[0] = -0.5; c0[1] = 0.5; c1[0] = -0.5; c1[1] = 0.5; c2[0] = pen; c2[1] = -pen;
        } else {
            c0[0] = -1.0; c1[0] = -1.0; c2[0] = pen;
        }
        for (int xs = 0; xs < ns; ++xs)
            for (int ys = 0; ys < ns; ++ys) {
                auto const& Ex = E_q[info.localNo[xs]];
                auto const& Ey = E_q[info.localNo[ys]];
                for (unsigned q = 0; q < nq; ++q) {
                    double w = c.fctRule.wts[q];
                    double nl = c.normalLength[f * nq + q];
                    for (unsigned k = 0; k < Nbf; ++k)
                        for (unsigned p = 0; p < Q; ++p)
                            for (unsigned l = 0; l < Nbf; ++l)
                                for (unsigned u = 0; u < Q; ++u) {
                                    double v = c0[xs] * Ex[k + Nbf * q] * w * top[ys][l + Nbf * (u + D * (p + D * q))] +
                                               c1[ys] * Ey[l + Nbf * q] * w * top[xs][k + Nbf * (p + D * (u + D * q))] +
                                               (p == u ? c2[std::abs(ys - xs)] * w * Ex[k + Nbf * q] *
                                                             Ey[l + Nbf * q] * nl
                                                       : 0.0);
                                    at(info.up[xs], info.up[ys], k, p, l, u) += v;
                                }
                }
            }
    }
}

// rhs_volume + rhs_skeleton + rhs_boundary (IP) (Elasticity.cpp:521-696; rhsVolume, rhs_lift_ip,
// rhsFacet in elasticity.py:151-180).  force_Q: E x Q x Nq.  facet_q: F x Q x nq, the Dirichlet
// data on Dirichlet facets and the slip on Fault facets (bc_skeleton / bc_boundary,
// Elasticity.cpp:545-573: a Fault on the boundary takes half the slip); other facets are ignored.
// Skeleton facets add to both sides with c10 = eps/2 and c20 = +-penalty, the side-0 normal and each
// side's own traces (Elasticity.cpp:575-643); boundary facets c10 = eps, c20 = penalty (:645-696).
void Elasticity::rhs(double const* force_Q, double const* facet_q, double* b) const {
    unsigned D = c.D, bs = Nbf * Q;
    std::size_t ne = c.nE();
    std::fill(b, b + bs * ne, 0.0);
    if (force_Q) {
        for (std::size_t el = 0; el < ne; ++el)
            for (unsigned k = 0; k < Nbf; ++k)
                for (unsigned p = 0; p < Q; ++p) {
                    double s = 0.0;
                    for (unsigned q = 0; q < Nq; ++q)
                        s += c.absDetJ[el * Nq + q] * c.volRule.wts[q] * E_Q[k + Nbf * q] *
                             force_Q[(el * Q + p) * Nq + q];
                    b[el * bs + k + p * Nbf] += s;
                }
    }
    if (!facet_q) return;
    std::vector<double> fq(Q * nq);
    for (std::size_t f = 0; f < c.nF(); ++f) {
        auto const& info = c.topo.fct[f];
        bool skel = info.up[0] != info.up[1];
        if (info.bc != BC::Dirichlet && info.bc != BC::Fault) continue;
        double scale = (!skel && info.bc == BC::Fault) ? 0.5 : 1.0;
        for (unsigned i = 0; i < Q * nq; ++i) fq[i] = scale * facet_q[f * Q * nq + i];
        double const* n = &c.normal[f * nq * D];
        for (int side = 0; side < (skel ? 2 : 1); ++side) {
            double const* g = side == 0 ? &c.JInv0[f * nq * D * D] : &c.JInv1[f * nq * D * D];
            auto Dx = physical_grad(D, Nbf, nq, g, Dxi_q[info.localNo[side]]);
            auto const& Eq = E_q[info.localNo[side]];
            double const* lam = side == 0 ? &lam_q0[f * nq] : &lam_q1[f * nq];
            double const* mu = side == 0 ? &mu_q0[f * nq] : &mu_q1[f * nq];
            double c10 = skel ? -0.5 : -1.0, c20 = side == 0 ? penalty[f] : -penalty[f];
            double* B = b + info.up[side] * bs;
            for (unsigned q = 0; q < nq; ++q) {
                double w = c.fctRule.wts[q];
                double lq = lam[q], mq = mu[q];
                double fn = 0.0;
                for (unsigned i = 0; i < D; ++i) fn += fq[i * nq + q] * n[q * D + i];
                for (unsigned k = 0; k < Nbf; ++k) {
                    double nd = 0.0;
                    for (unsigned j = 0; j < D; ++j) nd += Dx[k + Nbf * (j + D * q)] * n[q * D + j];
                    double dfd = 0.0;
                    for (unsigned i = 0; i < D; ++i) dfd += Dx[k + Nbf * (i + D * q)] * fq[i * nq + q];
                    for (unsigned p = 0; p < Q; ++p) {
                        double tt = lq * Dx[k + Nbf * (p + D * q)] * fn +
                                    mq * (nd * fq[p * nq + q] + dfd * n[q * D + p]);
                        B[k + p * Nbf] += c10 * tt * w +
                                          c20 * w * Eq[k + Nbf * q] * fq[p * nq + q] * c.normalLength[f * nq + q];
                    }
                }
            }
        }
    }
}

// compute_traction / compute_traction_bnd (elasticity.py:242-250; Elasticity.cpp:949-1018).  f_q
// (Q x nq, may be null) is the facet's data as for rhs: read on Dirichlet and Fault facets only
// (bc_skeleton / bc_boundary), halved on boundary Fault facets; zero elsewhere.
void Elasticity::traction(std::size_t f, double const* x, double const* f_q_in, double* t) const {
    unsigned D = c.D, bs = Nbf * Q;
    auto const& info = c.topo.fct[f];
    bool skel = info.up[0] != info.up[1];
    int ns = skel ? 2 : 1;
    double const* nu = &c.unitNormal[f * nq * D];
    double const* lq[2] = {&lam_q0[f * nq], &lam_q1[f * nq]};
    double const* mq[2] = {&mu_q0[f * nq], &mu_q1[f * nq]};
    std::vector<double> tr(D * nq, 0.0), uv(2 * D * nq, 0.0);
    for (int s = 0; s < ns; ++s) {
        double const* g = s == 0 ? &c.JInv0[f * nq * D * D] : &c.JInv1[f * nq * D * D];
        auto Dx = physical_grad(D, Nbf, nq, g, Dxi_q[info.localNo[s]]);
        auto const& Eq = E_q[info.localNo[s]];
        double const* u = x + info.up[s] * bs;
        for (unsigned q = 0; q < nq; ++q) {
            double J[3][3]; // J[l-comp][deriv]
            for (unsigned r = 0; r < D; ++r)
                for (unsigned j = 0; j < D; ++j) {
                    double v = 0.0;
                    for (unsigned l = 0; l < Nbf; ++l) v += Dx[l + Nbf * (j + D * q)] * u[l + r * Nbf];
                    J[r][j] = v;
                }
            double div = 0.0;
            for (unsigned r = 0; r < D; ++r) div += J[r][r];
            for (unsigned p = 0; p < D; ++p) {
                double v = lq[s][q] * div * nu[q * D + p];
                for (unsigned j = 0; j < D; ++j) v += mq[s][q] * (J[p][j] + J[j][p]) * nu[q * D + j];
                tr[p * nq + q] += (skel ? 0.5 : 1.0) * v;
                double ue = 0.0;
                for (unsigned l = 0; l < Nbf; ++l) ue += Eq[l + Nbf * q] * u[l + p * Nbf];
                uv[(s * D + p) * nq + q] = ue;
            }
        }
    }
    double c00 = -penalty[f];
    bool data = f_q_in && (info.bc == BC::Dirichlet || info.bc == BC::Fault);
    double scale = (!skel && info.bc == BC::Fault) ? 0.5 : 1.0;
    for (unsigned p = 0; p < D; ++p)
        for (unsigned q = 0; q < nq; ++q) {
            double fv = data ? scale * f_q_in[p * nq + q] : 0.0;
            double jump = uv[p * nq + q] - (skel ? uv[(D + p) * nq + q] : 0.0) - fv;
            t[p * nq + q] = tr[p * nq + q] + c00 * jump;
        }
}

// Sensitivity of a(u, psi) = psi^T A u to lam_Q, mu_Q.  The bilinear form of the apply is, per
// element, sum_q W|J| [lam' div u div psi + mu' (grad u + grad u^T) : grad psi] (apply_volume) and,
// per facet with side-0 area-scaled normal n, the symmetric interior penalty terms
//   skeleton: sum_q w [ -{sigma(psi)} n . [[u]] - {sigma(u)} n . [[psi]] + delta |n| [[psi]] . [[u]] ]
//   boundary: sum_q w [ -sigma(psi_0) n . u_0 - sigma(u_0) n . psi_0 + delta |n| psi_0 . u_0 ]
// (flux_u_*, flux_sigma_*, apply_facet of elasticity.py:201-234 summed over both elements).  lam',
// the traces lam_q and the penalty all derive from the nodal projection lam_nodal = P lam_Q
// (Elasticity.cpp:104-164, 189-225, 255-292); the penalty through min_i 2 mu_i and
// max_i (D lam_i + 2 mu_i) of each side, differentiated at the extremal node.
void Elasticity::material_gradient(double const* u, double const* psi, double* g_lam, double* g_mu) const {
    unsigned D = c.D, DD = D * D, bs = Nbf * Q;
    std::size_t ne = c.nE(), nf = c.nF();
    std::vector<double> gl(ne * Nbf, 0.0), gm(ne * Nbf, 0.0); // gradients w.r.t. nodal coefficients
    // gradient (J[r][j] = d v_r / d x_j) and value of a vector at points from a physical-gradient table
    auto grad_at = [&](double const* v, std::vector<double> const& Dx, unsigned np, unsigned p, double* J) {
        for (unsigned r = 0; r < D; ++r)
            for (unsigned j = 0; j < D; ++j) {
                double t = 0.0;
                for (unsigned k = 0; k < Nbf; ++k) t += Dx[k + Nbf * (j + D * p)] * v[k + r * Nbf];
                J[r * D + j] = t;
            }
        (void)np;
    };
    // volume
    for (std::size_t el = 0; el < ne; ++el) {
        auto Dx = physical_grad(D, Nbf, Nq, &c.JInv[el * Nq * DD], Dxi_Q);
        for (unsigned q = 0; q < Nq; ++q) {
            double Ju[9], Jp[9];
            grad_at(u + el * bs, Dx, Nq, q, Ju);
            grad_at(psi + el * bs, Dx, Nq, q, Jp);
            double du = 0.0, dp = 0.0, ss = 0.0;
            for (unsigned r = 0; r < D; ++r) {
                du += Ju[r * D + r];
                dp += Jp[r * D + r];
            }
            for (unsigned r = 0; r < D; ++r)
                for (unsigned j = 0; j < D; ++j) ss += (Ju[r * D + j] + Ju[j * D + r]) * Jp[r * D + j];
            double wj = c.volRule.wts[q] * c.absDetJ[el * Nq + q];
            for (unsigned i = 0; i < Nbf; ++i) {
                gl[el * Nbf + i] += matE_Q_T[q + i * Nq] * wj * du * dp;
                gm[el * Nbf + i] += matE_Q_T[q + i * Nq] * wj * ss;
            }
        }
    }
    // facets
    double cN1 = trace_constant(D, degree - 1);
    auto extremes = [&](std::size_t el, int* imin, int* imax, double* c0, double* c1) {
        *c0 = std::numeric_limits<double>::max();
        *c1 = std::numeric_limits<double>::lowest();
        for (unsigned i = 0; i < Nbf; ++i) {
            double a = 2.0 * mu[el * Nbf + i], b = D * lam[el * Nbf + i] + 2.0 * mu[el * Nbf + i];
            if (a < *c0) { *c0 = a; *imin = int(i); }
            if (b > *c1) { *c1 = b; *imax = int(i); }
        }
    };
    for (std::size_t f = 0; f < nf; ++f) {
        auto const& info = c.topo.fct[f];
        bool skel = info.up[0] != info.up[1];
        if (!skel && info.bc != BC::Dirichlet && info.bc != BC::Fault) continue; // natural: no terms
        int ns = skel ? 2 : 1;
        double h = skel ? 0.5 : 1.0;
        double const* lq[2] = {&lam_q0[f * nq], &lam_q1[f * nq]};
        double const* mq[2] = {&mu_q0[f * nq], &mu_q1[f * nq]};
        std::vector<double> Dx[2];
        std::vector<double> uv[2], pv[2];
        for (int s2 = 0; s2 < ns; ++s2) {
            Dx[s2] = physical_grad(D, Nbf, nq, s2 == 0 ? &c.JInv0[f * nq * DD] : &c.JInv1[f * nq * DD],
                                   Dxi_q[info.localNo[s2]]);
            uv[s2].assign(D * nq, 0.0);
            pv[s2].assign(D * nq, 0.0);
            auto const& Eq = E_q[info.localNo[s2]];
            for (unsigned q = 0; q < nq; ++q)
                for (unsigned r = 0; r < D; ++r)
                    for (unsigned k = 0; k < Nbf; ++k) {
                        uv[s2][r * nq + q] += Eq[k + Nbf * q] * u[info.up[s2] * bs + k + r * Nbf];
                        pv[s2][r * nq + q] += Eq[k + Nbf * q] * psi[info.up[s2] * bs + k + r * Nbf];
                    }
        }
        double gdelta = 0.0;
        for (unsigned q = 0; q < nq; ++q) {
            double w = c.fctRule.wts[q];
            double const* n = &c.normal[(f * nq + q) * D];
            double ju[3], jp[3];
            for (unsigned r = 0; r < D; ++r) {
                ju[r] = uv[0][r * nq + q] - (skel ? uv[1][r * nq + q] : 0.0);
                jp[r] = pv[0][r * nq + q] - (skel ? pv[1][r * nq + q] : 0.0);
            }
            double jun = 0.0, jpn = 0.0, jj = 0.0;
            for (unsigned r = 0; r < D; ++r) {
                jun += ju[r] * n[r];
                jpn += jp[r] * n[r];
                jj += ju[r] * jp[r];
            }
            gdelta += w * c.normalLength[f * nq + q] * jj;
            for (int s2 = 0; s2 < ns; ++s2) {
                double Ju[9], Jp[9];
                grad_at(u + info.up[s2] * bs, Dx[s2], nq, q, Ju);
                grad_at(psi + info.up[s2] * bs, Dx[s2], nq, q, Jp);
                double du = 0.0, dp = 0.0, su = 0.0, sp = 0.0;
                for (unsigned r = 0; r < D; ++r) {
                    du += Ju[r * D + r];
                    dp += Jp[r * D + r];
                }
                for (unsigned r = 0; r < D; ++r)
                    for (unsigned j = 0; j < D; ++j) {
                        su += jp[r] * (Ju[r * D + j] + Ju[j * D + r]) * n[j];
                        sp += ju[r] * (Jp[r * D + j] + Jp[j * D + r]) * n[j];
                    }
                double dl = w * h * (-dp * jun - du * jpn), dm = w * h * (-sp - su);
                auto const& mT = matE_q_T[info.localNo[s2]];
                std::size_t el = info.up[s2];
                for (unsigned i = 0; i < Nbf; ++i) {
                    gl[el * Nbf + i] += mT[q + i * nq] * dl;
                    gm[el * Nbf + i] += mT[q + i * nq] * dm;
                }
                (void)lq;
                (void)mq;
            }
        }
        // penalty: delta = (p_0 + p_1) / 4 (skeleton) or p_0 (boundary), p_s = C c1^2 / c0
        for (int s2 = 0; s2 < ns; ++s2) {
            std::size_t el = info.up[s2];
            int imin = 0, imax = 0;
            double c0 = 0.0, c1 = 0.0;
            extremes(el, &imin, &imax, &c0, &c1);
            double C = (D + 1) * cN1 * (c.area[f] / c.volume[el]);
            double gp = gdelta * (skel ? 0.25 : 1.0);
            double dp_dc0 = -C * c1 * c1 / (c0 * c0), dp_dc1 = 2.0 * C * c1 / c0;
            gm[el * Nbf + imin] += gp * dp_dc0 * 2.0;
            gl[el * Nbf + imax] += gp * dp_dc1 * D;
            gm[el * Nbf + imax] += gp * dp_dc1 * 2.0;
        }
    }
    // through the projection lam_nodal = M^-1 matE^T diag(W|J|) lam_Q: g_Q = diag(W|J|) matE M^-1 g
    for (std::size_t el = 0; el < ne; ++el) {
        std::vector<double> M(Nbf * Nbf, 0.0), rhs(2 * Nbf);
        for (unsigned k = 0; k < Nbf; ++k)
            for (unsigned l = 0; l < Nbf; ++l) {
                double t = 0.0;
                for (unsigned q = 0; q < Nq; ++q)
                    t += matE_Q_T[q + k * Nq] * c.volRule.wts[q] * c.absDetJ[el * Nq + q] * matE_Q_T[q + l * Nq];
                M[k + l * Nbf] = t;
            }
        for (unsigned k = 0; k < Nbf; ++k) {
            rhs[k] = gl[el * Nbf + k];
            rhs[Nbf + k] = gm[el * Nbf + k];
        }
        lu_solve(M, Nbf, rhs.data(), 2); // M symmetric: M^-T g = M^-1 g
        for (unsigned q = 0; q < Nq; ++q) {
            double wj = c.volRule.wts[q] * c.absDetJ[el * Nq + q], a = 0.0, b = 0.0;
            for (unsigned k = 0; k < Nbf; ++k) {
                a += matE_Q_T[q + k * Nq] * rhs[k];
                b += matE_Q_T[q + k * Nq] * rhs[Nbf + k];
            }
            g_lam[el * Nq + q] = wj * a;
            g_mu[el * Nq + q] = wj * b;
        }
    }
}

void Elasticity::inverse_mass(double const* x, double* y) const {
    unsigned bs = Nbf * Q;
    auto Minv = space.inverse_mass_matrix(); // Nbf x Nbf, column-major
    std::vector<double> t(Nbf), uq(Nq);
    for (std::size_t el = 0; el < c.nE(); ++el)
        for (unsigned p = 0; p < Q; ++p) {
            double const* xe = x + el * bs + p * Nbf;
            for (unsigned k = 0; k < Nbf; ++k) {
                double v = 0.0;
                for (unsigned l = 0; l < Nbf; ++l) v += Minv[k + l * Nbf] * xe[l];
                t[k] = v;
            }
            for (unsigned q = 0; q < Nq; ++q) {
                double v = 0.0;
                for (unsigned k = 0; k < Nbf; ++k) v += E_Q[k + Nbf * q] * t[k];
                uq[q] = -v * c.volRule.wts[q] / c.absDetJ[el * Nq + q];
            }
            for (unsigned k = 0; k < Nbf; ++k) {
                double v = 0.0;
                for (unsigned q = 0; q < Nq; ++q) v += E_Q[k + Nbf * q] * uq[q];
                t[k] = v;
            }
            for (unsigned k = 0; k < Nbf; ++k) {
                double v = 0.0;
                for (unsigned l = 0; l < Nbf; ++l) v += Minv[k + l * Nbf] * t[l];
                y[el * bs + p * Nbf + k] = v;
            }
        }
}

void Elasticity::wave_rhs(double const* x, double const* f_q, double* y) const {
    std::size_t n = block_size() * c.nE();
    std::vector<double> ax(n), b(n, 0.0);
    apply(x, ax.data(), 1);
    if (f_q) rhs(nullptr, f_q, b.data()); // the add_bc terms of apply_<true> are -rhs_facet
    for (std::size_t i = 0; i < n; ++i) ax[i] -= b[i];
    inverse_mass(ax.data(), y);
}

std::vector<double> nodal_interpolation(unsigned D, unsigned to_degree, unsigned from_degree) {
    RefElement to(D, to_degree, false), from(D, from_degree, false);
    std::vector<double> pts(to.nodes.begin(), to.nodes.end());
    return from.eval_basis(pts); // phi^from_l(x^to_k) stored (l, k): transpose to (k, l)
}

// ================================================================================= Poisson
// Poisson ctor + prepare_* (Poisson.cpp:27-61, 133-223); solution space is modal Dubiner.
void Poisson::setup(Mesh const& mesh, unsigned N, std::vector<double> const* nodes, double const* K_Q) {
    c.init(mesh, N, nodes);
    unsigned D = c.D;
    space = RefElement(D, N, true);
    matSpace = RefElement(D, N, false);
    Nbf = space.nbf;
    Nq = c.volRule.size();
    nq = c.fctRule.size();
    E_Q = space.eval_basis(c.volRule.pts);
    Dxi_Q = space.eval_grad(c.volRule.pts);
    matE_Q_T = transpose2(matSpace.eval_basis(c.volRule.pts), matSpace.nbf, Nq);
    E_q.clear();
    E_q_T.clear();
    Dxi_q.clear();
    Dxi_q_120.clear();
    matE_q_T.clear();
    for (unsigned f = 0; f <= D; ++f) {
        auto pts = c.cl.facet_param(f, c.fctRule.pts);
        E_q.push_back(space.eval_basis(pts));
        E_q_T.push_back(transpose2(E_q.back(), Nbf, nq));
        Dxi_q.push_back(space.eval_grad(pts));
        Dxi_q_120.push_back(perm120(Dxi_q.back(), Nbf, D, nq));
        matE_q_T.push_back(transpose2(matSpace.eval_basis(pts), matSpace.nbf, nq));
    }
    std::size_t ne = c.nE(), nf = c.nF();
    unsigned DD = D * D, Nm = matSpace.nbf;
    K.assign(ne * Nm, 0.0);
    for (std::size_t el = 0; el < ne; ++el)
        project_coefficient(Nm, Nq, matE_Q_T, c.volRule.wts.data(), &c.absDetJ[el * Nq],
                            &K_Q[el * Nq], &K[el * Nm]);
    // K_G_q (poisson.py:121): K_G_q(e,i,q) = K(m) matE_q_T(q,m) G_q(e,i,q)
    KG0.assign(nf * nq * DD, 0.0);
    KG1.assign(nf * nq * DD, 0.0);
    for (std::size_t f = 0; f < nf; ++f) {
        auto const& info = c.topo.fct[f];
        bool skel = info.up[0] != info.up[1];
        for (int side = 0; side < 2; ++side) {
            int s = skel ? side : 0;
            double const* G = s == 0 ? &c.JInv0[f * nq * DD] : &c.JInv1[f * nq * DD];
            auto const& mT = matE_q_T[info.localNo[s]];
            for (unsigned q = 0; q < nq; ++q) {
                double kq = 0.0;
                for (unsigned m = 0; m < Nm; ++m) kq += K[info.up[s] * Nm + m] * mT[q + m * nq];
                for (unsigned i = 0; i < DD; ++i) (side == 0 ? KG0 : KG1)[(f * nq + q) * DD + i] = kq * G[q * DD + i];
            }
        }
    }
    // J_W_K_Q (poisson.py:120)
    JWK.assign(ne * Nq, 0.0);
    for (std::size_t el = 0; el < ne; ++el)
        for (unsigned q = 0; q < Nq; ++q) {
            double kq = 0.0;
            for (unsigned m = 0; m < Nm; ++m) kq += K[el * Nm + m] * matE_Q_T[q + m * Nq];
            JWK[el * Nq + q] = c.absDetJ[el * Nq + q] * c.volRule.wts[q] * kq;
        }
    // prepare_penalty (Poisson.cpp:209-223)
    penalty.assign(nf, 0.0);
    double cN1 = trace_constant(D, N - 1);
    auto p = [&](std::size_t f, int side) {
        auto el = c.topo.fct[f].up[side];
        double k0 = *std::min_element(&K[el * Nm], &K[el * Nm] + Nm);
        double k1 = *std::max_element(&K[el * Nm], &K[el * Nm] + Nm);
        return (D + 1) * cN1 * (c.area[f] / c.volume[el]) * (k1 * k1 / k0);
    };
    for (std::size_t f = 0; f < nf; ++f) {
        auto const& info = c.topo.fct[f];
        penalty[f] = info.up[0] != info.up[1] ? (p(f, 0) + p(f, 1)) / 4.0 : p(f, 0);
    }
}

// Poisson::apply (Poisson.cpp:551-644; poisson.py:123-144)
void Poisson::apply_element(std::size_t elNo, double const* x, double* y) const {
    unsigned D = c.D, DD = D * D;
    double const* U = x + elNo * Nbf;
    double* Unew = y + elNo * Nbf;
    // apply_volume: U_new(k) = J_W_K_Q(q) Dx_Q(k,r,q) Dx_Q(l,r,q) U(l)
    auto Dx = physical_grad(D, Nbf, Nq, &c.JInv[elNo * Nq * DD], Dxi_Q);
    std::vector<double> g(D * Nq, 0.0);
    for (unsigned q = 0; q < Nq; ++q)
        for (unsigned r = 0; r < D; ++r) {
            double s = 0.0;
            for (unsigned l = 0; l < Nbf; ++l) s += Dx[l + Nbf * (r + D * q)] * U[l];
            g[r + D * q] = s * JWK[elNo * Nq + q];
        }
    for (unsigned k = 0; k < Nbf; ++k) {
        double s = 0.0;
        for (unsigned q = 0; q < Nq; ++q)
            for (unsigned r = 0; r < D; ++r) s += Dx[k + Nbf * (r + D * q)] * g[r + D * q];
        Unew[k] = s;
    }
    std::vector<double> nflip(nq * D), nuflip(nq * D), uh(nq), sig(D * nq);
    for (unsigned f = 0; f <= D; ++f) {
        auto const& si = c.topo.side[elNo * (D + 1) + f];
        bool skel = static_cast<std::size_t>(si.lid) != elNo;
        bool fd = si.bc == BC::Fault || si.bc == BC::Dirichlet;
        std::size_t fct = si.fctNo;
        double const* n = &c.normal[fct * nq * D];
        double const* nu = &c.unitNormal[fct * nq * D];
        double const* K0 = &KG0[fct * nq * DD];
        double const* K1 = &KG1[fct * nq * DD];
        if (skel && si.side == 1) {
            std::swap(K0, K1);
            for (unsigned i = 0; i < nq * D; ++i) {
                nflip[i] = -n[i];
                nuflip[i] = -nu[i];
            }
            n = nflip.data();
            nu = nuflip.data();
        }
        auto grad_trace = [&](std::vector<double> const& D120, double const* Ux, double const* KG,
                              unsigned q, unsigned p) {
            // K_G_q(e,p,q) Dxi_q_120(e,q,l) U(l)
            double v = 0.0;
            for (unsigned e = 0; e < D; ++e) {
                double s = 0.0;
                for (unsigned l = 0; l < Nbf; ++l) s += D120[e + D * (q + nq * l)] * Ux[l];
                v += KG[q * DD + e + p * D] * s;
            }
            return v;
        };
        auto val = [&](std::vector<double> const& ET, double const* Ux, unsigned q) {
            double s = 0.0;
            for (unsigned l = 0; l < Nbf; ++l) s += ET[q + nq * l] * Ux[l];
            return s;
        };
        if (si.bc == BC::None || (skel && fd)) {
            double const* Ue = x + si.lid * Nbf;
            double c00 = -penalty[fct];
            for (unsigned q = 0; q < nq; ++q) {
                double u0 = val(E_q_T[f], U, q), u1 = val(E_q_T[si.localNo], Ue, q);
                uh[q] = 0.5 * (-u0 + u1);
                for (unsigned p = 0; p < D; ++p)
                    sig[p + D * q] = 0.5 * (grad_trace(Dxi_q_120[f], U, K0, q, p) +
                                            grad_trace(Dxi_q_120[si.localNo], Ue, K1, q, p)) +
                                     c00 * (u0 - u1) * nu[q * D + p];
            }
        } else if (fd) {
            double c00 = -penalty[fct];
            for (unsigned q = 0; q < nq; ++q) {
                double u0 = val(E_q_T[f], U, q);
                uh[q] = -u0;
                for (unsigned p = 0; p < D; ++p)
                    sig[p + D * q] = grad_trace(Dxi_q_120[f], U, K0, q, p) + c00 * u0 * nu[q * D + p];
            }
        } else {
            continue;
        }
        // apply_facet: U_new(k) += w(q) n(r,q) (u_hat(q) K_G_q0(e,r,q) Dxi_q(k,e,q) - E_q(k,q) sigma(r,q))
        for (unsigned q = 0; q < nq; ++q) {
            double w = c.fctRule.wts[q];
            for (unsigned k = 0; k < Nbf; ++k) {
                double v = 0.0;
                for (unsigned r = 0; r < D; ++r) {
                    double kd = 0.0;
                    for (unsigned e = 0; e < D; ++e) kd += K0[q * DD + e + r * D] * Dxi_q[f][k + Nbf * (e + D * q)];
                    v += n[q * D + r] * (uh[q] * kd - E_q[f][k + Nbf * q] * sig[r + D * q]);
                }
                Unew[k] += w * v;
            }
        }
    }
}

void Poisson::apply(double const* x, double* y, int nthreads) const {
    std::size_t ne = c.nE();
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
    for (std::size_t el = 0; el < ne; ++el) apply_element(el, x, y);
}

// Poisson::assemble_volume/skeleton/boundary (Poisson.cpp:225-381; poisson.py:29-80)
void Poisson::assemble_dense(double* A) const {
    unsigned D = c.D, DD = D * D, Nm = matSpace.nbf;
    std::size_t ne = c.nE(), n = Nbf * ne;
    std::fill(A, A + n * n, 0.0);
    for (std::size_t el = 0; el < ne; ++el) {
        auto Dx = physical_grad(D, Nbf, Nq, &c.JInv[el * Nq * DD], Dxi_Q);
        for (unsigned q = 0; q < Nq; ++q) {
            for (unsigned k = 0; k < Nbf; ++k)
                for (unsigned l = 0; l < Nbf; ++l) {
                    double s = 0.0;
                    for (unsigned i = 0; i < D; ++i) s += Dx[k + Nbf * (i + D * q)] * Dx[l + Nbf * (i + D * q)];
                    A[(el * Nbf + k) + n * (el * Nbf + l)] += JWK[el * Nq + q] * s;
                }
        }
    }
    for (std::size_t f = 0; f < c.nF(); ++f) {
        auto const& info = c.topo.fct[f];
        bool skel = info.up[0] != info.up[1];
        if (!skel && info.bc == BC::Natural) continue;
        int ns = skel ? 2 : 1;
        std::vector<double> KDx[2];
        double const* nrm = &c.normal[f * nq * D];
        for (int s = 0; s < ns; ++s) {
            double const* G = s == 0 ? &c.JInv0[f * nq * DD] : &c.JInv1[f * nq * DD];
            auto Dx = physical_grad(D, Nbf, nq, G, Dxi_q[info.localNo[s]]);
            auto const& mT = matE_q_T[info.localNo[s]];
            KDx[s] = Dx;
            for (unsigned q = 0; q < nq; ++q) {
                double kq = 0.0;
                for (unsigned m = 0; m < Nm; ++m) kq += K[info.up[s] * Nm + m] * mT[q + m * nq];
                for (unsigned i = 0; i < D; ++i)
                    for (unsigned k = 0; k < Nbf; ++k) KDx[s][k + Nbf * (i + D * q)] *= kq;
            }
        }
        double pen = penalty[f];
        double c0[2], c1[2], c2[2];
        if (skel) {
            c0[0] = -0.5; c0[1] = 0.5; c1[0] = -0.5; c1[1] = 0.5; c2[0] = pen; c2[1] = -pen;
        } else {
            c0[0] = -1.0; c1[0] = -1.0; c2[0] = pen;
        }
        // a(x,y)(k,l) = c0[y] w KDx[x](k,i) n(i) E_q[y](l) + c1[x] w KDx[y](l,i) n(i) E_q[x](k)
        //             + c2[|y-x|] w E_q[x](k) E_q[y](l) nl
        for (int xs = 0; xs < ns; ++xs)
            for (int ys = 0; ys < ns; ++ys) {
                auto const& Ex = E_q[info.localNo[xs]];
                auto const& Ey = E_q[info.localNo[ys]];
                for (unsigned q = 0; q < nq; ++q) {
                    double w = c.fctRule.wts[q], nl = c.normalLength[f * nq + q];
                    for (unsigned k = 0; k < Nbf; ++k)
                        for (unsigned l = 0; l < Nbf; ++l) {
                            double a = 0.0, b = 0.0;
                            for (unsigned i = 0; i < D; ++i) {
                                a += KDx[xs][k + Nbf * (i + D * q)] * nrm[q * D + i];
                                b += KDx[ys][l + Nbf * (i + D * q)] * nrm[q * D + i];
                            }
                            double v = c0[ys] * w * a * Ey[l + Nbf * q] + c1[xs] * w * b * Ex[k + Nbf * q] +
                                       c2[std::abs(ys - xs)] * w * Ex[k + Nbf * q] * Ey[l + Nbf * q] * nl;
                            A[(info.up[xs] * Nbf + k) + n * (info.up[ys] * Nbf + l)] += v;
                        }
                }
            }
    }
}

// rhs_volume + rhs_skeleton + rhs_boundary (IP) (Poisson.cpp:383-540; poisson.py:79-103), with
// the facet data conventions of Elasticity::rhs.
void Poisson::rhs(double const* force_Q, double const* facet_q, double* b) const {
    unsigned D = c.D, DD = D * D, Nm = matSpace.nbf;
    std::size_t ne = c.nE();
    std::fill(b, b + Nbf * ne, 0.0);
    if (force_Q) {
        for (std::size_t el = 0; el < ne; ++el)
            for (unsigned k = 0; k < Nbf; ++k) {
                double s = 0.0;
                for (unsigned q = 0; q < Nq; ++q)
                    s += c.absDetJ[el * Nq + q] * c.volRule.wts[q] * E_Q[k + Nbf * q] * force_Q[el * Nq + q];
                b[el * Nbf + k] += s;
            }
    }
    if (!facet_q) return;
    for (std::size_t f = 0; f < c.nF(); ++f) {
        auto const& info = c.topo.fct[f];
        bool skel = info.up[0] != info.up[1];
        if (info.bc != BC::Dirichlet && info.bc != BC::Fault) continue;
        double scale = (!skel && info.bc == BC::Fault) ? 0.5 : 1.0;
        double const* fq = &facet_q[f * nq];
        double const* n = &c.normal[f * nq * D];
        for (int side = 0; side < (skel ? 2 : 1); ++side) {
            double const* g = side == 0 ? &c.JInv0[f * nq * DD] : &c.JInv1[f * nq * DD];
            auto Dx = physical_grad(D, Nbf, nq, g, Dxi_q[info.localNo[side]]);
            auto const& mT = matE_q_T[info.localNo[side]];
            auto const& Eq = E_q[info.localNo[side]];
            double c10 = skel ? -0.5 : -1.0, c20 = side == 0 ? penalty[f] : -penalty[f];
            std::size_t el = info.up[side];
            for (unsigned q = 0; q < nq; ++q) {
                double kq = 0.0;
                for (unsigned m = 0; m < Nm; ++m) kq += K[el * Nm + m] * mT[q + m * nq];
                double w = c.fctRule.wts[q];
                for (unsigned k = 0; k < Nbf; ++k) {
                    double kn = 0.0;
                    for (unsigned i = 0; i < D; ++i) kn += kq * Dx[k + Nbf * (i + D * q)] * n[q * D + i];
                    b[el * Nbf + k] += c10 * w * kn * scale * fq[q] +
                                       c20 * w * Eq[k + Nbf * q] * c.normalLength[f * nq + q] * scale * fq[q];
                }
            }
        }
    }
}

} // namespace oracle

namespace oracle {

// AdapterBase::AdapterBase / prepare (app/localoperator/AdapterBase.cpp:20-100) with
// Curvilinear::facetBasis (src/geometry/Curvilinear.cpp:281-316).
void FaultAdapter::setup(DGCommon const& c, double const* up, double const* ref_normal) {
    D = c.D;
    RefElement space(D - 1, c.N, false);
    nbf = space.nbf;
    nq = unsigned(c.fctRule.wts.size());
    e_q = space.eval_basis(c.fctRule.pts); // (l, q)
    w = c.fctRule.wts;
    facets.clear();
    for (std::size_t f = 0; f < c.nF(); ++f)
        if (c.topo.fct[f].bc == BC::Fault) facets.push_back(int64_t(f));
    const std::size_t nf = facets.size();
    nl.assign(nf * nq, 0.0);
    basis.assign(nf * nq * D * D, 0.0);
    minv.assign(nf * nbf * nbf, 0.0);
    const double tol = 10000.0 * std::numeric_limits<double>::epsilon();
    for (std::size_t m = 0; m < nf; ++m) {
        const std::size_t f = std::size_t(facets[m]);
        for (unsigned q = 0; q < nq; ++q) {
            double n[3] = {0, 0, 0}, len = 0.0, dot = 0.0;
            for (unsigned d = 0; d < D; ++d) {
                n[d] = c.normal[(f * nq + q) * D + d];
                len += n[d] * n[d];
                dot += ref_normal[d] * n[d];
            }
            len = std::sqrt(len);
            nl[m * nq + q] = len;
            if (std::fabs(dot) < tol) throw std::logic_error("Normal and reference normal are almost perpendicular.");
            const bool flipped = dot < 0.0;
            if (flipped)
                for (unsigned d = 0; d < D; ++d) n[d] = -n[d];
            double* b = &basis[(m * nq + q) * D * D]; // b[i D + j]: component i of vector j
            double nh[3] = {0, 0, 0};
            for (unsigned d = 0; d < D; ++d) nh[d] = n[d] / len;
            for (unsigned d = 0; d < D; ++d) b[d * D + 0] = nh[d];
            if (D == 2) {
                const double sg = up[0] * nh[1] - up[1] * nh[0];
                if (std::fabs(sg) < tol) throw std::logic_error("Up vector and normal are almost colinear.");
                const double s = sg > 0 ? 1.0 : -1.0;
                b[0 * D + 1] = -s * nh[1];
                b[1 * D + 1] = s * nh[0];
            } else {
                // s = normalize(up x n), d = normalize(s x n)
                double sv[3] = {up[1] * nh[2] - up[2] * nh[1], up[2] * nh[0] - up[0] * nh[2],
                                up[0] * nh[1] - up[1] * nh[0]};
                double sl = std::sqrt(sv[0] * sv[0] + sv[1] * sv[1] + sv[2] * sv[2]);
                if (sl < tol) throw std::logic_error("Up vector and normal are almost colinear.");
                for (double& v : sv) v /= sl;
                double dv[3] = {sv[1] * nh[2] - sv[2] * nh[1], sv[2] * nh[0] - sv[0] * nh[2],
                                sv[0] * nh[1] - sv[1] * nh[0]};
                double dl = std::sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
                for (unsigned d = 0; d < 3; ++d) {
                    b[d * D + 1] = dv[d] / dl;
                    b[d * D + 2] = sv[d];
                }
            }
            if (flipped)
                for (unsigned i = 0; i < D * D; ++i) b[i] = -b[i];
        }
        // m(i, j) = sum_q w_q nl_q e(i, q) e(j, q); minv = m^-1
        std::vector<double> M(nbf * nbf, 0.0), I(nbf * nbf, 0.0);
        for (unsigned i = 0; i < nbf; ++i)
            for (unsigned j = 0; j < nbf; ++j)
                for (unsigned q = 0; q < nq; ++q)
                    M[i + j * nbf] += w[q] * nl[m * nq + q] * e_q[i + nbf * q] * e_q[j + nbf * q];
        for (unsigned i = 0; i < nbf; ++i) I[i + i * nbf] = 1.0;
        lu_solve(M, int(nbf), I.data(), int(nbf));
        std::copy(I.begin(), I.end(), minv.begin() + m * nbf * nbf);
    }
}

// evaluate_traction: traction(k, p) = sum_l minv(l, k) sum_q e_q(l, q) w_q nl_q sum_o t(o, q) basis(o, p, q)
void FaultAdapter::traction(std::size_t m, double const* t_q, double* out) const {
    std::vector<double> s(nbf * D, 0.0);
    for (unsigned q = 0; q < nq; ++q) {
        double const* b = &basis[(m * nq + q) * D * D];
        for (unsigned p = 0; p < D; ++p) {
            double tb = 0.0;
            for (unsigned o = 0; o < D; ++o) tb += t_q[o * nq + q] * b[o * D + p];
            for (unsigned l = 0; l < nbf; ++l) s[l + p * nbf] += e_q[l + nbf * q] * w[q] * nl[m * nq + q] * tb;
        }
    }
    double const* Mi = &minv[m * nbf * nbf];
    for (unsigned p = 0; p < D; ++p)
        for (unsigned k = 0; k < nbf; ++k) {
            double v = 0.0;
            for (unsigned l = 0; l < nbf; ++l) v += Mi[l + k * nbf] * s[l + p * nbf];
            out[k + p * nbf] = v;
        }
}

// evaluate_slip: slip_q(p, q) = sum_l e_q(l, q) sum_n basis(p, n + 1, q) s(l, n)
void FaultAdapter::slip(std::size_t m, double const* s, double* slip_q) const {
    for (unsigned p = 0; p < D; ++p)
        for (unsigned q = 0; q < nq; ++q) {
            double const* b = &basis[(m * nq + q) * D * D];
            double v = 0.0;
            for (unsigned n = 0; n + 1 < D; ++n)
                for (unsigned l = 0; l < nbf; ++l) v += e_q[l + nbf * q] * b[p * D + n + 1] * s[l + n * nbf];
            slip_q[p * nq + q] = v;
        }
}

} // namespace oracle
