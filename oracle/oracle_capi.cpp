// Oracle: extern "C" entry points for the test suite (ctypes).  Test infrastructure only.
#include "oracle.h"

#include <algorithm>
#include <cstring>
#include <exception>
#include <memory>
#include <stdexcept>
#include <string>

using namespace oracle;

namespace {
thread_local std::string g_err;

struct OpHandle {
    int kind; // 0 elasticity, 1 poisson
    std::unique_ptr<Elasticity> el;
    std::unique_ptr<Poisson> po;
    LocalOp* op() { return kind == 0 ? static_cast<LocalOp*>(el.get()) : po.get(); }
    DGCommon& c() { return kind == 0 ? el->c : po->c; }
};

template <class F> int guard(F&& f) {
    try {
        f();
        return 0;
    } catch (std::exception const& e) {
        g_err = e.what();
        return -1;
    }
}

Mesh make_mesh(int D, int64_t nv, double const* verts, int64_t ne, int64_t const* elems, int64_t nb,
               int64_t const* bf, int const* bbc) {
    Mesh m;
    m.D = D;
    m.verts.assign(verts, verts + nv * D);
    m.elems.assign(elems, elems + ne * (D + 1));
    for (int64_t e = 0; e < ne; ++e) std::sort(m.elems.begin() + e * (D + 1), m.elems.begin() + (e + 1) * (D + 1));
    m.bfacets.assign(bf, bf + nb * D);
    for (int64_t e = 0; e < nb; ++e) std::sort(m.bfacets.begin() + e * D, m.bfacets.begin() + (e + 1) * D);
    m.bbc.assign(bbc, bbc + nb);
    return m;
}
} // namespace

extern "C" {

char const* oracle_last_error() { return g_err.c_str(); }

int oracle_load_rules(char const* path) { return guard([&] { load_rule_tables(path); }); }

double oracle_jacobi_p(unsigned n, unsigned a, unsigned b, double x) { return jacobi_p(n, a, b, x); }
double oracle_sf_jacobi_p(unsigned n, unsigned a, unsigned b, double x, double y) {
    return sf_jacobi_p(n, a, b, x, y);
}
double oracle_dubiner(unsigned D, unsigned const* i, double const* xi) { return dubiner(D, i, xi); }
void oracle_grad_dubiner(unsigned D, unsigned const* i, double const* xi, double* g) {
    grad_dubiner(D, i, xi, g);
}
int oracle_all_integer_sums(unsigned D, unsigned maxSum, unsigned* out) {
    auto v = all_integer_sums(D, maxSum);
    if (out) std::memcpy(out, v.data(), v.size() * sizeof(unsigned));
    return static_cast<int>(v.size() / D);
}
void oracle_gauss_jacobi(unsigned n, unsigned a, unsigned b, double* pts, double* wts) {
    std::vector<double> p, w;
    gauss_jacobi(n, a, b, p, w);
    std::memcpy(pts, p.data(), n * sizeof(double));
    std::memcpy(wts, w.data(), n * sizeof(double));
}
void oracle_lgl_points(unsigned n, double* out) {
    auto p = lgl_points(n);
    std::memcpy(out, p.data(), n * sizeof(double));
}
// family: 0 auto, 1 tensor product, 2 WV, 3 JS.  Returns number of points (-1 if none); order out.
int oracle_rule(int family, unsigned D, unsigned minOrder, double* pts, double* wts, int* order) {
    Rule r;
    if (family == 0) r = simplex_rule(D, minOrder);
    else if (family == 1) r = tensor_product_rule(D, minOrder);
    else r = tabulated_rule(family == 2 ? "WV" : "JS", D, minOrder);
    if (r.size() == 0) return -1;
    if (pts) std::memcpy(pts, r.pts.data(), r.pts.size() * sizeof(double));
    if (wts) std::memcpy(wts, r.wts.data(), r.wts.size() * sizeof(double));
    if (order) *order = r.order;
    return static_cast<int>(r.size());
}
int oracle_wab_nodes(unsigned D, unsigned N, double* out) {
    auto v = wab_nodes(D, N);
    if (out) std::memcpy(out, v.data(), v.size() * sizeof(double));
    return static_cast<int>(v.size() / D);
}
// which: 0 mass, 1 inverse mass
int oracle_ref_mass(unsigned D, unsigned N, int modal, int which, double* out) {
    return guard([&] {
        RefElement r(D, N, modal != 0);
        auto M = which == 0 ? r.mass_matrix() : r.inverse_mass_matrix();
        std::memcpy(out, M.data(), M.size() * sizeof(double));
    });
}
int oracle_eval_basis(unsigned D, unsigned N, int modal, int np, double const* pts, double* E,
                      double* G) {
    return guard([&] {
        RefElement r(D, N, modal != 0);
        std::vector<double> p(pts, pts + np * D);
        if (E) {
            auto v = r.eval_basis(p);
            std::memcpy(E, v.data(), v.size() * sizeof(double));
        }
        if (G) {
            auto v = r.eval_grad(p);
            std::memcpy(G, v.data(), v.size() * sizeof(double));
        }
    });
}

// ----------------------------------------------------------------------------------- meshes
void* oracle_box_mesh(unsigned D, uint64_t const* N, int tess, int const* bcs) {
    try {
        std::vector<uint64_t> n(N, N + D);
        std::vector<int> b(bcs, bcs + 2 * D);
        return new Mesh(gen_box_mesh(D, n, tess, b));
    } catch (std::exception const& e) {
        g_err = e.what();
        return nullptr;
    }
}
void* oracle_mesh_new(int D, int64_t nv, double const* verts, int64_t ne, int64_t const* elems,
                      int64_t nb, int64_t const* bf, int const* bbc) {
    return new Mesh(make_mesh(D, nv, verts, ne, elems, nb, bf, bbc));
}
void oracle_mesh_sizes(void* h, int64_t* out) {
    auto m = static_cast<Mesh*>(h);
    out[0] = m->D;
    out[1] = static_cast<int64_t>(m->verts.size() / m->D);
    out[2] = static_cast<int64_t>(m->num_elements());
    out[3] = static_cast<int64_t>(m->bbc.size());
}
void oracle_mesh_arrays(void* h, double* verts, int64_t* elems, int64_t* bf, int* bbc) {
    auto m = static_cast<Mesh*>(h);
    std::memcpy(verts, m->verts.data(), m->verts.size() * sizeof(double));
    std::memcpy(elems, m->elems.data(), m->elems.size() * sizeof(int64_t));
    std::memcpy(bf, m->bfacets.data(), m->bfacets.size() * sizeof(int64_t));
    std::memcpy(bbc, m->bbc.data(), m->bbc.size() * sizeof(int));
}
void oracle_mesh_free(void* h) { delete static_cast<Mesh*>(h); }
void oracle_affine_nodes(void* h, unsigned N, double* out) {
    auto v = affine_geometry_nodes(*static_cast<Mesh*>(h), N);
    std::memcpy(out, v.data(), v.size() * sizeof(double));
}

// Geometry only: volume-quadrature-point coordinates (E x Nq x D) and facet-point coordinates
// (F x nq x D).  sizes out: Nq, nq, F.
int oracle_geometry_coords(void* h, unsigned N, double const* nodes, int64_t* sizes, double* vcoords,
                           double* fcoords) {
    return guard([&] {
        auto m = static_cast<Mesh*>(h);
        std::vector<double> nd;
        if (nodes) nd.assign(nodes, nodes + m->num_elements() * binom(N + m->D, m->D) * m->D);
        DGCommon c;
        c.init(*m, N, nodes ? &nd : nullptr);
        sizes[0] = c.volRule.size();
        sizes[1] = c.fctRule.size();
        sizes[2] = c.nF();
        if (vcoords) std::memcpy(vcoords, c.vcoords.data(), c.vcoords.size() * sizeof(double));
        if (fcoords) std::memcpy(fcoords, c.fcoords.data(), c.fcoords.size() * sizeof(double));
    });
}

// ------------------------------------------------------------------------------- operators
void* oracle_op_new(int kind, void* mesh, unsigned N, double const* nodes, double const* c0,
                    double const* c1) {
    try {
        auto m = static_cast<Mesh*>(mesh);
        std::vector<double> nd;
        if (nodes) nd.assign(nodes, nodes + m->num_elements() * binom(N + m->D, m->D) * m->D);
        auto h = new OpHandle{kind, nullptr, nullptr};
        if (kind == 0) {
            h->el = std::make_unique<Elasticity>();
            h->el->setup(*m, N, nodes ? &nd : nullptr, c0, c1);
        } else {
            h->po = std::make_unique<Poisson>();
            h->po->setup(*m, N, nodes ? &nd : nullptr, c0);
        }
        return h;
    } catch (std::exception const& e) {
        g_err = e.what();
        return nullptr;
    }
}
void oracle_op_free(void* h) { delete static_cast<OpHandle*>(h); }
// out: bs, E, F, Nq, nq, D
void oracle_op_info(void* h, int64_t* out) {
    auto o = static_cast<OpHandle*>(h);
    auto& c = o->c();
    out[0] = o->op()->block_size();
    out[1] = c.nE();
    out[2] = c.nF();
    out[3] = c.volRule.size();
    out[4] = c.fctRule.size();
    out[5] = c.D;
}
int oracle_op_apply(void* h, double const* x, double* y, int nthreads) {
    return guard([&] { static_cast<OpHandle*>(h)->op()->apply(x, y, nthreads); });
}
int oracle_op_assemble(void* h, double* A) {
    return guard([&] { static_cast<OpHandle*>(h)->op()->assemble_dense(A); });
}
int oracle_op_rhs(void* h, double const* force_Q, double const* dirichlet_q, double* b) {
    return guard([&] { static_cast<OpHandle*>(h)->op()->rhs(force_Q, dirichlet_q, b); });
}
int oracle_op_traction(void* h, int64_t fct, double const* x, double const* f_q, double* t) {
    return guard([&] {
        auto o = static_cast<OpHandle*>(h);
        if (o->kind != 0) throw std::runtime_error("traction: elasticity only");
        o->el->traction(fct, x, f_q, t);
    });
}
// Fault adapter: setup returns {fault facets, nbf, nq} in sizes
int oracle_op_fault_setup(void* h, double const* up, double const* ref_normal, int64_t* sizes) {
    return guard([&] {
        auto o = static_cast<OpHandle*>(h);
        if (o->kind != 0) throw std::runtime_error("fault adapter: elasticity only");
        o->el->fault.setup(o->el->c, up, ref_normal);
        sizes[0] = int64_t(o->el->fault.facets.size());
        sizes[1] = o->el->fault.nbf;
        sizes[2] = o->el->fault.nq;
    });
}
int oracle_op_fault_traction(void* h, int64_t m, double const* t_q, double* out) {
    return guard([&] { static_cast<OpHandle*>(h)->el->fault.traction(std::size_t(m), t_q, out); });
}
int oracle_op_fault_slip(void* h, int64_t m, double const* s, double* slip_q) {
    return guard([&] { static_cast<OpHandle*>(h)->el->fault.slip(std::size_t(m), s, slip_q); });
}
// facets (nf), basis (nf x nq x D x D), nl (nf x nq), e_q (nbf x nq)
int oracle_op_fault_arrays(void* h, int64_t* facets, double* basis, double* nl, double* e_q) {
    return guard([&] {
        auto const& a = static_cast<OpHandle*>(h)->el->fault;
        std::copy(a.facets.begin(), a.facets.end(), facets);
        std::copy(a.basis.begin(), a.basis.end(), basis);
        std::copy(a.nl.begin(), a.nl.end(), nl);
        std::copy(a.e_q.begin(), a.e_q.end(), e_q);
    });
}
// Discrete solution at the volume quadrature points: out (E x Q x Nq) = sum_k E_Q(k,q) x(k,p).
// Also returns the quadrature weights times |det J| (E x Nq) in wj when non-null.
int oracle_op_eval_volume(void* h, double const* x, double* out, double* wj) {
    return guard([&] {
        auto o = static_cast<OpHandle*>(h);
        auto& c = o->c();
        unsigned Nbf = o->kind == 0 ? o->el->Nbf : o->po->Nbf;
        unsigned Q = o->kind == 0 ? o->el->Q : 1;
        auto const& E = o->kind == 0 ? o->el->E_Q : o->po->E_Q;
        std::size_t Nq = c.volRule.size(), ne = c.nE(), bs = Nbf * Q;
        for (std::size_t e = 0; e < ne; ++e)
            for (unsigned p = 0; p < Q; ++p)
                for (std::size_t q = 0; q < Nq; ++q) {
                    double s = 0.0;
                    for (unsigned k = 0; k < Nbf; ++k) s += E[k + Nbf * q] * x[e * bs + k + p * Nbf];
                    out[(e * Q + p) * Nq + q] = s;
                    if (wj && p == 0) wj[e * Nq + q] = c.volRule.wts[q] * c.absDetJ[e * Nq + q];
                }
    });
}

// facet table: F x 5 (up0, up1, localNo0, localNo1, bc)
int oracle_op_wave_rhs(void* h, double const* x, double const* f_q, double* y) {
    return guard([&] {
        auto o = static_cast<OpHandle*>(h);
        if (o->kind != 0) throw std::invalid_argument("wave_rhs is defined for elasticity");
        o->el->wave_rhs(x, f_q, y);
    });
}

int oracle_op_inverse_mass(void* h, double const* x, double* y) {
    return guard([&] {
        auto o = static_cast<OpHandle*>(h);
        if (o->kind != 0) throw std::invalid_argument("inverse_mass is defined for elasticity");
        o->el->inverse_mass(x, y);
    });
}

// I (n_to x n_from, row-major for the caller): I[k][l] = phi^from_l(x^to_k)
int oracle_nodal_interpolation(int D, int to, int from, double* out) {
    return guard([&] {
        auto E = oracle::nodal_interpolation(unsigned(D), unsigned(to), unsigned(from));
        oracle::RefElement rt(unsigned(D), unsigned(to), false), rf(unsigned(D), unsigned(from), false);
        for (unsigned k = 0; k < rt.nbf; ++k)
            for (unsigned l = 0; l < rf.nbf; ++l) out[k * rf.nbf + l] = E[l + rf.nbf * k];
    });
}

int oracle_op_material_gradient(void* h, double const* u, double const* psi, double* g_lam, double* g_mu) {
    return guard([&] {
        auto o = static_cast<OpHandle*>(h);
        if (o->kind != 0) throw std::invalid_argument("material gradient is defined for elasticity");
        o->el->material_gradient(u, psi, g_lam, g_mu);
    });
}

void oracle_op_facets(void* h, int64_t* out) {
    auto& c = static_cast<OpHandle*>(h)->c();
    for (std::size_t f = 0; f < c.nF(); ++f) {
        auto const& fi = c.topo.fct[f];
        out[5 * f + 0] = fi.up[0];
        out[5 * f + 1] = fi.up[1];
        out[5 * f + 2] = fi.localNo[0];
        out[5 * f + 3] = fi.localNo[1];
        out[5 * f + 4] = static_cast<int>(fi.bc);
    }
}
// named arrays for cross-checks
int oracle_op_array(void* h, char const* name, double* out) {
    return guard([&] {
        auto o = static_cast<OpHandle*>(h);
        auto& c = o->c();
        std::string n(name);
        std::vector<double> const* v = nullptr;
        if (n == "absDetJ") v = &c.absDetJ;
        else if (n == "JInv") v = &c.JInv;
        else if (n == "vcoords") v = &c.vcoords;
        else if (n == "volume") v = &c.volume;
        else if (n == "JInv0") v = &c.JInv0;
        else if (n == "JInv1") v = &c.JInv1;
        else if (n == "normal") v = &c.normal;
        else if (n == "unitNormal") v = &c.unitNormal;
        else if (n == "normalLength") v = &c.normalLength;
        else if (n == "fcoords") v = &c.fcoords;
        else if (n == "area") v = &c.area;
        else if (n == "facetRuleWeights") v = &c.fctRule.wts;
        else if (o->kind == 0) {
            auto& e = *o->el;
            if (n == "lam") v = &e.lam;
            else if (n == "mu") v = &e.mu;
            else if (n == "lamWJ") v = &e.lamWJ;
            else if (n == "muWJ") v = &e.muWJ;
            else if (n == "lam_q0") v = &e.lam_q0;
            else if (n == "lam_q1") v = &e.lam_q1;
            else if (n == "mu_q0") v = &e.mu_q0;
            else if (n == "mu_q1") v = &e.mu_q1;
            else if (n == "penalty") v = &e.penalty;
        } else {
            auto& p = *o->po;
            if (n == "K") v = &p.K;
            else if (n == "JWK") v = &p.JWK;
            else if (n == "penalty") v = &p.penalty;
        }
        if (!v) throw std::runtime_error("unknown array " + n);
        std::memcpy(out, v->data(), v->size() * sizeof(double));
    });
}
int64_t oracle_op_array_size(void* h, char const* name) {
    // size query via a scratch copy is wasteful; compute from the known layouts instead
    auto o = static_cast<OpHandle*>(h);
    auto& c = o->c();
    std::string n(name);
    int64_t E = c.nE(), F = c.nF(), Nq = c.volRule.size(), nq = c.fctRule.size(), D = c.D;
    if (n == "absDetJ") return E * Nq;
    if (n == "JInv") return E * Nq * D * D;
    if (n == "vcoords") return E * Nq * D;
    if (n == "volume") return E;
    if (n == "JInv0" || n == "JInv1") return F * nq * D * D;
    if (n == "normal" || n == "unitNormal" || n == "fcoords") return F * nq * D;
    if (n == "normalLength" || n == "facetRuleWeights") return n == "normalLength" ? F * nq : nq;
    if (n == "area" || n == "penalty") return F;
    if (n == "lam" || n == "mu" || n == "K") return E * (o->kind == 0 ? o->el->Nbf : o->po->matSpace.nbf);
    if (n == "lamWJ" || n == "muWJ" || n == "JWK") return E * Nq;
    if (n == "lam_q0" || n == "lam_q1" || n == "mu_q0" || n == "mu_q1") return F * nq;
    return -1;
}

} // extern "C"
