// Oracle: box meshes, DG topology and curvilinear geometry (test infrastructure).
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <map>
#include <stdexcept>

namespace oracle {

static std::vector<uint64_t> unflatten(uint64_t idx, std::vector<uint64_t> const& shape) {
    std::vector<uint64_t> a(shape.size());
    for (std::size_t d = 0; d < shape.size(); ++d) {
        a[d] = idx % shape[d];
        idx /= shape[d];
    }
    return a;
}
static uint64_t flatten(std::vector<uint64_t> const& idx, std::vector<uint64_t> const& shape) {
    uint64_t f = 0;
    for (int d = static_cast<int>(shape.size()) - 1; d >= 0; --d) f = f * shape[d] + idx[d];
    return f;
}

// Faces of a simplex in Choose<D>(D+1) order (src/mesh/Simplex.h:76-98).
static std::vector<std::vector<int>> simplex_faces(unsigned D) {
    if (D == 2) return {{0, 1}, {0, 2}, {1, 2}};
    return {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};
}

// GenMesh<D>::uniformMesh + tessellate + extractBoundaryMesh (src/mesh/GenMesh.cpp:21-143,
// 145-186) on [0,1]^D with N[d] intervals; tess = 1 swaps the 5-tet cube tessellation for the
// 6-tet Kuhn subdivision (cube split along its 000-111 diagonal).
Mesh gen_box_mesh(unsigned D, std::vector<uint64_t> const& N, int tess, std::vector<int> const& bcs) {
    Mesh m;
    m.D = D;
    std::vector<uint64_t> Np1(D);
    uint64_t nv = 1, nc = 1;
    for (unsigned d = 0; d < D; ++d) {
        Np1[d] = N[d] + 1;
        nv *= Np1[d];
        nc *= N[d];
    }
    m.verts.resize(nv * D);
    for (uint64_t v = 0; v < nv; ++v) {
        auto idx = unflatten(v, Np1);
        for (unsigned d = 0; d < D; ++d) {
            double h = 1.0 / static_cast<double>(N[d]);
            m.verts[v * D + d] = 0.0 + static_cast<double>(idx[d]) * h;
        }
    }
    std::vector<uint64_t> D2(D, 2);
    for (uint64_t c = 0; c < nc; ++c) {
        auto e = unflatten(c, N);
        std::vector<uint64_t> g(1u << D);
        for (uint64_t vf = 0; vf < (1u << D); ++vf) {
            auto o = unflatten(vf, D2);
            std::vector<uint64_t> v(D);
            for (unsigned d = 0; d < D; ++d) v[d] = e[d] + o[d];
            g[vf] = flatten(v, Np1);
        }
        bool odd = false;
        for (auto ee : e) odd ^= (ee % 2 == 1);
        std::vector<std::vector<uint64_t>> s;
        if (D == 2) {
            if (odd)
                s = {{g[0], g[1], g[3]}, {g[0], g[2], g[3]}};
            else
                s = {{g[0], g[1], g[2]}, {g[1], g[2], g[3]}};
        } else if (tess == 0) {
            if (odd)
                s = {{g[0], g[1], g[2], g[4]},
                     {g[1], g[2], g[3], g[7]},
                     {g[2], g[4], g[6], g[7]},
                     {g[1], g[4], g[5], g[7]},
                     {g[2], g[4], g[1], g[7]}};
            else
                s = {{g[0], g[2], g[3], g[6]},
                     {g[0], g[1], g[3], g[5]},
                     {g[0], g[4], g[5], g[6]},
                     {g[3], g[5], g[6], g[7]},
                     {g[0], g[3], g[5], g[6]}};
        } else {
            static const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2},
                                            {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
            for (auto const& p : perms) {
                int a = 1 << p[0];
                int b = a | (1 << p[1]);
                s.push_back({g[0], g[a], g[b], g[7]});
            }
        }
        for (auto& t : s) {
            std::sort(t.begin(), t.end());
            for (auto v : t) m.elems.push_back(static_cast<int64_t>(v));
        }
    }
    // boundary faces: all face vertices share coordinate index d which is 0 or N[d]
    auto faces = simplex_faces(D);
    std::size_t ne = m.elems.size() / (D + 1);
    for (std::size_t el = 0; el < ne; ++el) {
        for (auto const& f : faces) {
            std::vector<std::vector<uint64_t>> vs;
            for (int k : f) vs.push_back(unflatten(m.elems[el * (D + 1) + k], Np1));
            for (unsigned d = 0; d < D; ++d) {
                bool in = true;
                for (auto& v : vs) in = in && (v[d] == vs[0][d]);
                if (in && (vs[0][d] == 0 || vs[0][d] == N[d])) {
                    for (int k : f) m.bfacets.push_back(m.elems[el * (D + 1) + k]);
                    m.bbc.push_back(bcs[2 * d + (vs[0][d] == 0 ? 0 : 1)]);
                }
            }
        }
    }
    return m;
}

// DGOperatorTopo (src/form/DGOperatorTopo.cpp:13-109) on a single partition.
Topo build_topology(Mesh const& mesh) {
    unsigned D = mesh.D;
    auto faces = simplex_faces(D);
    std::map<std::vector<int64_t>, int> bc;
    for (std::size_t i = 0; i < mesh.bbc.size(); ++i) {
        std::vector<int64_t> key(mesh.bfacets.begin() + i * D, mesh.bfacets.begin() + (i + 1) * D);
        bc[key] = mesh.bbc[i];
    }
    Topo t;
    std::map<std::vector<int64_t>, int64_t> fid;
    std::size_t ne = mesh.num_elements();
    for (std::size_t el = 0; el < ne; ++el) {
        for (std::size_t lf = 0; lf < faces.size(); ++lf) {
            std::vector<int64_t> key;
            for (int k : faces[lf]) key.push_back(mesh.elems[el * (D + 1) + k]);
            auto it = fid.find(key);
            if (it == fid.end()) {
                FacetInfo fi;
                fi.up[0] = fi.up[1] = static_cast<int64_t>(el);
                fi.localNo[0] = fi.localNo[1] = static_cast<int>(lf);
                auto b = bc.find(key);
                fi.bc = b == bc.end() ? BC::None : static_cast<BC>(b->second);
                fid[key] = static_cast<int64_t>(t.fct.size());
                t.fct.push_back(fi);
            } else {
                auto& fi = t.fct[it->second];
                if (fi.up[1] != fi.up[0]) throw std::runtime_error("non-manifold facet");
                fi.up[1] = static_cast<int64_t>(el);
                fi.localNo[1] = static_cast<int>(lf);
            }
        }
    }
    t.side.resize(ne * (D + 1));
    for (std::size_t f = 0; f < t.fct.size(); ++f) {
        auto const& fi = t.fct[f];
        for (int i = 0; i < 2; ++i) {
            if (i == 1 && fi.up[0] == fi.up[1]) break;
            auto& si = t.side[fi.up[i] * (D + 1) + fi.localNo[i]];
            si.fctNo = static_cast<int64_t>(f);
            si.side = i;
            si.lid = fi.up[(i + 1) % 2];
            si.localNo = fi.localNo[(i + 1) % 2];
            si.bc = fi.bc;
        }
    }
    return t;
}

// RefPlexToGeneralPlex (src/geometry/Affine.h:14-29) applied to the WAB nodes of each element,
// vertices taken in ascending-id order (Curvilinear.cpp:101-106).
std::vector<double> affine_geometry_nodes(Mesh const& mesh, unsigned N) {
    unsigned D = mesh.D;
    auto ref = wab_nodes(D, N);
    std::size_t nbf = ref.size() / D, ne = mesh.num_elements();
    std::vector<double> out(ne * nbf * D);
    for (std::size_t el = 0; el < ne; ++el) {
        double const* v[4];
        for (unsigned k = 0; k <= D; ++k) v[k] = &mesh.verts[mesh.elems[el * (D + 1) + k] * D];
        for (std::size_t p = 0; p < nbf; ++p) {
            double s = 0.0;
            for (unsigned d = 0; d < D; ++d) s += ref[p * D + d];
            for (unsigned i = 0; i < D; ++i) {
                double x = (1.0 - s) * v[0][i];
                for (unsigned d = 0; d < D; ++d) x = x + ref[p * D + d] * v[d + 1][i];
                out[(el * nbf + p) * D + i] = x;
            }
        }
    }
    return out;
}

// Curvilinear ctor (src/geometry/Curvilinear.cpp:32-160): reference normals of the faces.
Curvilinear::Curvilinear(Mesh const& mesh, unsigned N_, std::vector<double> const* nodes_in)
    : D(mesh.D), N(N_), geo(mesh.D, N_, false) {
    nodes = nodes_in ? *nodes_in : affine_geometry_nodes(mesh, N);
    double rv[4][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    auto faces = simplex_faces(D);
    for (auto const& f : faces) {
        std::array<int, 3> fv{f[0], f[1], D == 3 ? f[2] : -1};
        f2v.push_back(fv);
        std::array<double, 3> n{0, 0, 0};
        if (D == 2) {
            double e[2] = {rv[f[1]][0] - rv[f[0]][0], rv[f[1]][1] - rv[f[0]][1]};
            n[0] = -e[1];
            n[1] = e[0];
        } else {
            double e1[3], e2[3];
            for (int d = 0; d < 3; ++d) {
                e1[d] = rv[f[1]][d] - rv[f[0]][d];
                e2[d] = rv[f[2]][d] - rv[f[0]][d];
            }
            n[0] = e1[1] * e2[2] - e1[2] * e2[1];
            n[1] = e1[2] * e2[0] - e1[0] * e2[2];
            n[2] = e1[0] * e2[1] - e1[1] * e2[0];
        }
        int missing = 0;
        for (int k = 0; k <= static_cast<int>(D); ++k)
            if (std::find(f.begin(), f.end(), k) == f.end()) missing = k;
        double dot = 0.0;
        for (unsigned d = 0; d < D; ++d) dot += n[d] * (rv[missing][d] - rv[f[0]][d]);
        if (dot > 0) {
            for (auto& x : n) x *= -1.0;
        }
        refNormals.push_back(n);
    }
}

std::size_t Curvilinear::num_elements() const { return nodes.size() / (geo.nbf * D); }

// Curvilinear::facetParam (Curvilinear.cpp:355-381)
std::vector<double> Curvilinear::facet_param(unsigned f, std::vector<double> const& chi) const {
    double rv[4][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    std::size_t np = chi.size() / (D - 1);
    std::vector<double> xi(np * D);
    for (std::size_t p = 0; p < np; ++p) {
        double cs = 0.0;
        for (unsigned d = 0; d < D - 1; ++d) cs += chi[p * (D - 1) + d];
        for (unsigned i = 0; i < D; ++i) {
            double x = (1.0 - cs) * rv[f2v[f][0]][i];
            for (unsigned d = 0; d < D - 1; ++d) x = x + chi[p * (D - 1) + d] * rv[f2v[f][d + 1]][i];
            xi[p * D + i] = x;
        }
    }
    return xi;
}

// Curvilinear::jacobian (Curvilinear.cpp:203-226)
void Curvilinear::jacobian(std::size_t el, std::vector<double> const& gradE, double* J) const {
    unsigned nbf = geo.nbf;
    std::size_t np = gradE.size() / (nbf * D);
    double const* X = &nodes[el * nbf * D];
    for (std::size_t p = 0; p < np; ++p)
        for (unsigned j = 0; j < D; ++j)
            for (unsigned d = 0; d < D; ++d) {
                double s = 0.0;
                for (unsigned k = 0; k < nbf; ++k) s += X[k * D + d] * gradE[k + nbf * (j + D * p)];
                J[d + D * (j + D * p)] = s;
            }
}

// Curvilinear::map (Curvilinear.cpp:178-191)
void Curvilinear::map(std::size_t el, std::vector<double> const& E, double* x) const {
    unsigned nbf = geo.nbf;
    std::size_t np = E.size() / nbf;
    double const* X = &nodes[el * nbf * D];
    for (std::size_t p = 0; p < np; ++p)
        for (unsigned d = 0; d < D; ++d) {
            double s = 0.0;
            for (unsigned k = 0; k < nbf; ++k) s += X[k * D + d] * E[k + nbf * p];
            x[d + D * p] = s;
        }
}

static double det(unsigned D, double const* J) {
    if (D == 2) return J[0] * J[3] - J[2] * J[1];
    return J[0] * (J[4] * J[8] - J[7] * J[5]) - J[3] * (J[1] * J[8] - J[7] * J[2]) +
           J[6] * (J[1] * J[5] - J[4] * J[2]);
}
static void inv(unsigned D, double const* J, double* R) {
    std::vector<double> A(J, J + D * D);
    auto I = mat_inverse(A, D);
    for (unsigned i = 0; i < D * D; ++i) R[i] = I[i];
}

// DGCurvilinearCommon ctor + prepare_volume + prepare_bndskl
// (src/form/DGCurvilinearCommon.cpp:13-101), MinQuadOrder = 2N+1 (app/config.h.in:8-11).
void DGCommon::init(Mesh const& m, unsigned N_, std::vector<double> const* nodes) {
    mesh = m;
    D = m.D;
    N = N_;
    topo = build_topology(mesh);
    cl = Curvilinear(mesh, N, nodes);
    unsigned mq = 2 * N + 1;
    fctRule = simplex_rule(D - 1, mq);
    volRule = simplex_rule(D, mq);
    geoE_Q = cl.geo.eval_basis(volRule.pts);
    geoDxi_Q = cl.geo.eval_grad(volRule.pts);
    for (unsigned f = 0; f <= D; ++f) {
        auto pts = cl.facet_param(f, fctRule.pts);
        geoE_q.push_back(cl.geo.eval_basis(pts));
        geoDxi_q.push_back(cl.geo.eval_grad(pts));
    }
    std::size_t ne = nE(), nf = nF(), Nq = volRule.size(), nq = fctRule.size();
    unsigned DD = D * D;
    absDetJ.assign(ne * Nq, 0.0);
    JInv.assign(ne * Nq * DD, 0.0);
    vcoords.assign(ne * Nq * D, 0.0);
    volume.assign(ne, 0.0);
#pragma omp parallel for schedule(static)
    for (std::size_t el = 0; el < ne; ++el) {
        std::vector<double> J(Nq * DD);
        cl.jacobian(el, geoDxi_Q, J.data());
        for (std::size_t q = 0; q < Nq; ++q) {
            absDetJ[el * Nq + q] = std::fabs(det(D, &J[q * DD]));
            inv(D, &J[q * DD], &JInv[(el * Nq + q) * DD]);
        }
        cl.map(el, geoE_Q, &vcoords[el * Nq * D]);
        double v = 0.0;
        for (std::size_t q = 0; q < Nq; ++q) v += volRule.wts[q] * absDetJ[el * Nq + q];
        volume[el] = v;
    }
    JInv0.assign(nf * nq * DD, 0.0);
    JInv1.assign(nf * nq * DD, 0.0);
    normal.assign(nf * nq * D, 0.0);
    unitNormal.assign(nf * nq * D, 0.0);
    normalLength.assign(nf * nq, 0.0);
    fcoords.assign(nf * nq * D, 0.0);
    area.assign(nf, 0.0);
#pragma omp parallel for schedule(static)
    for (std::size_t f = 0; f < nf; ++f) {
        std::vector<double> J(nq * DD);
        auto const& info = topo.fct[f];
        cl.jacobian(info.up[0], geoDxi_q[info.localNo[0]], J.data());
        auto const& rn = cl.refNormals[info.localNo[0]];
        for (std::size_t q = 0; q < nq; ++q) {
            double dj = det(D, &J[q * DD]);
            double* Ji = &JInv0[(f * nq + q) * DD];
            inv(D, &J[q * DD], Ji);
            double nrm2 = 0.0;
            for (unsigned i = 0; i < D; ++i) {
                double s = 0.0;
                for (unsigned j = 0; j < D; ++j) s += Ji[j + i * D] * rn[j];
                normal[(f * nq + q) * D + i] = std::fabs(dj) * s;
                nrm2 += normal[(f * nq + q) * D + i] * normal[(f * nq + q) * D + i];
            }
            double len = std::sqrt(nrm2);
            normalLength[f * nq + q] = len;
            for (unsigned i = 0; i < D; ++i)
                unitNormal[(f * nq + q) * D + i] = normal[(f * nq + q) * D + i] / len;
        }
        cl.map(info.up[0], geoE_q[info.localNo[0]], &fcoords[f * nq * D]);
        double a = 0.0;
        for (std::size_t q = 0; q < nq; ++q) a += fctRule.wts[q] * normalLength[f * nq + q];
        area[f] = a;
        cl.jacobian(info.up[1], geoDxi_q[info.localNo[1]], J.data());
        for (std::size_t q = 0; q < nq; ++q) inv(D, &J[q * DD], &JInv1[(f * nq + q) * DD]);
    }
}

} // namespace oracle
