// CPU FP64 restatement of tandem's matrix-free DG operator (the checker for the HIP path).
//
// TEST INFRASTRUCTURE ONLY.  Nothing in tandem_amd/ links, imports or calls this code; only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, and only as the thing
// the GPU results are compared against (or as the reported CPU baseline).
//
// Every function cites the reference file:line it restates.  The reference itself cannot be
// built here (PETSc, Eigen, Lua, METIS and the yateto/mneme submodules are absent), so this
// restatement is pinned by the reference's own golden values (test/basis.cpp, test/quadrules.cpp,
// test/form.cpp, test/geometry.cpp), by exactness identities and by operator invariants.
//
// Parity pins, per function group (see tests/test_oracle_*.py):
//   basis (Jacobi, Dubiner, WAB nodes)       : test/basis.cpp golden values, closed forms
//   quadrature (Gauss-Jacobi, tables, auto)  : test/quadrules.cpp Maple values, monomial exactness
//   reference element (mass, projection)     : test/form.cpp golden values
//   curvilinear geometry                     : test/geometry.cpp rotated square; volume/normal sums
//   elasticity / poisson apply               : assembled-matrix path (independent restatement),
//                                              symmetry, rigid-body null space, patch test, MMS
//                                              convergence orders
#pragma once

#include <array>
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace oracle {

enum class BC : int { None = 0, Natural = 1, Fault = 3, Dirichlet = 5 };

// ---------------------------------------------------------------- basis (src/basis/Functions.cpp)
double jacobi_p(unsigned n, unsigned a, unsigned b, double x);
double jacobi_p_derivative(unsigned n, unsigned a, unsigned b, double x);
double sf_jacobi_p(unsigned n, unsigned a, unsigned b, double x, double y);
std::array<double, 3> sf_jacobi_p_and_derivatives(unsigned n, unsigned a, unsigned b, double x,
                                                  double y);
double dubiner(unsigned D, unsigned const* i, double const* xi);
void grad_dubiner(unsigned D, unsigned const* i, double const* xi, double* grad);

// AllIntegerSums<D>(maxSum) ordering (src/util/Combinatorics.h:46-110); D entries per index.
std::vector<unsigned> all_integer_sums(unsigned D, unsigned maxSum);
std::size_t binom(unsigned n, unsigned k);

// ------------------------------------------------------------------ quadrature (src/quadrules)
struct Rule {
    unsigned D = 0;
    int order = -1;
    std::vector<double> pts; // D per point
    std::vector<double> wts;
    std::size_t size() const { return wts.size(); }
};
void gauss_jacobi(unsigned n, unsigned a, unsigned b, std::vector<double>& pts,
                  std::vector<double>& wts);
std::vector<double> lgl_points(unsigned n);
Rule tensor_product_rule(unsigned D, unsigned minQuadOrder);
// Tabulated families: "WV" (D=2,3) and "JS" (D=3).  Returns empty rule if not tabulated.
Rule tabulated_rule(std::string const& family, unsigned D, unsigned minQuadOrder);
Rule simplex_rule(unsigned D, unsigned minQuadOrder);
void load_rule_tables(std::string const& path);

// ------------------------------------------------------------------- nodes (WarpAndBlend.cpp)
double wab_alpha(unsigned D, unsigned degree);
std::vector<double> wab_nodes(unsigned D, unsigned degree); // D per node

// --------------------------------------------------------------- reference elements (RefElement)
struct RefElement {
    unsigned D = 0, N = 0, nbf = 0;
    bool modal = true;
    std::vector<double> nodes; // nodal only
    std::vector<double> V, Vinv; // nbf x nbf, column-major
    RefElement() = default;
    RefElement(unsigned D, unsigned N, bool modal);
    // E(k, p) = phi_k(pts_p), column-major nbf x npts
    std::vector<double> eval_basis(std::vector<double> const& pts) const;
    // G(k, d, p), column-major nbf x D x npts
    std::vector<double> eval_grad(std::vector<double> const& pts) const;
    std::vector<double> mass_matrix() const;
    std::vector<double> inverse_mass_matrix() const;
};

// dense helpers
std::vector<double> mat_inverse(std::vector<double> A, int n);      // column-major
void lu_solve(std::vector<double> A, int n, double* b, int nrhs); // column-major, b n x nrhs

// ------------------------------------------------------------------------------ mesh + topology
struct Mesh {
    unsigned D = 0;
    std::vector<double> verts;    // D per vertex
    std::vector<int64_t> elems;   // D+1 per element, ascending vertex ids
    std::vector<int64_t> bfacets; // D per boundary facet, ascending
    std::vector<int> bbc;         // BC per boundary facet
    std::size_t num_elements() const { return elems.size() / (D + 1); }
};
// GenMesh<D>::uniformMesh (src/mesh/GenMesh.cpp:21-143); tess 0 = GenMesh scheme (2 triangles
// per square, 5 tets per cube), tess 1 = 6-tet Kuhn subdivision of each cube (3D only).
// bcs[2*d + s] is the BC on the plane x_d = s.
Mesh gen_box_mesh(unsigned D, std::vector<uint64_t> const& N, int tess, std::vector<int> const& bcs);

struct FacetInfo {
    int64_t up[2];
    int localNo[2];
    BC bc;
};
struct SideInfo {
    int64_t fctNo;
    int side;
    int64_t lid;
    int localNo;
    BC bc;
};
struct Topo {
    std::vector<FacetInfo> fct;
    std::vector<SideInfo> side; // (D+1) per element
};
Topo build_topology(Mesh const& mesh); // DGOperatorTopo.cpp:13-109

// -------------------------------------------------------------------------- geometry (Curvilinear)
struct Curvilinear {
    unsigned D = 0, N = 0;
    RefElement geo;
    std::vector<double> nodes; // E x nbf x D  (per element, node-major)
    std::vector<std::array<int, 3>> f2v;
    std::vector<std::array<double, 3>> refNormals;
    Curvilinear() = default;
    Curvilinear(Mesh const& mesh, unsigned N, std::vector<double> const* nodes_in);
    std::size_t num_elements() const;
    std::vector<double> facet_param(unsigned f, std::vector<double> const& chi) const;
    // J(d, j, p) = sum_k x_k[d] gradE(k, j, p)
    void jacobian(std::size_t el, std::vector<double> const& gradE, double* J) const;
    void map(std::size_t el, std::vector<double> const& E, double* x) const;
};
std::vector<double> affine_geometry_nodes(Mesh const& mesh, unsigned N);

// ------------------------------------------------------------------------ DG curvilinear common
struct DGCommon {
    unsigned D = 0, N = 0;
    Mesh mesh;
    Topo topo;
    Curvilinear cl;
    Rule volRule, fctRule;
    std::vector<double> geoE_Q, geoDxi_Q;
    std::vector<std::vector<double>> geoE_q, geoDxi_q;
    // volume data
    std::vector<double> absDetJ, JInv, vcoords, volume;
    // facet data
    std::vector<double> JInv0, JInv1, normal, unitNormal, normalLength, fcoords, area;
    void init(Mesh const& mesh, unsigned N, std::vector<double> const* nodes);
    std::size_t nE() const { return mesh.num_elements(); }
    std::size_t nF() const { return topo.fct.size(); }
};

// Fault adapter (app/localoperator/AdapterBase.cpp:20-100, ElasticityAdapter.cpp:21-48, generated
// from app/kernels/elasticity_adapter.py:5-30): the fault facets (BC::Fault, facet order), on each
// the nodal P_N basis of the (D-1)-simplex at the facet rule's points, the normal length, the
// fault basis (unit normal, tangents from the up vector; signs fixed by the reference normal) and
// the inverse of the facet mass matrix weighted by the normal length.
struct FaultAdapter {
    unsigned D = 0, nbf = 0, nq = 0;
    std::vector<int64_t> facets;   // fault facet -> facet
    std::vector<double> e_q, w;    // nbf x nq (column-major), nq
    std::vector<double> nl;        // nf x nq
    std::vector<double> basis;     // nf x nq x D x D: basis[((m nq + q) D + i) D + j] = component i of vector j
    std::vector<double> minv;      // nf x nbf x nbf (column-major)
    void setup(DGCommon const& c, double const* up, double const* ref_normal);
    // traction(k, p) (nbf x D column-major) of facet-point traction t(o, q) (D x nq)
    void traction(std::size_t m, double const* t_q, double* out) const;
    // slip(o, q) (D x nq) at the facet points of the fault-basis slip s(l, n) (nbf x (D-1) column-major)
    void slip(std::size_t m, double const* s, double* slip_q) const;
};

struct LocalOp {
    virtual ~LocalOp() = default;
    virtual std::size_t block_size() const = 0;
    virtual void apply(double const* x, double* y, int nthreads) const = 0;
    virtual void assemble_dense(double* A) const = 0; // n x n column-major, n = bs*E
    virtual void rhs(double const* force_Q, double const* dirichlet_q, double* b) const = 0;
};

// app/localoperator/Elasticity.cpp
struct Elasticity : LocalOp {
    DGCommon c;
    RefElement space, matSpace;
    unsigned Nbf = 0, Nq = 0, nq = 0, Q = 0;
    std::vector<double> E_Q, Dxi_Q, Dxi_Q_120, matE_Q_T;
    std::vector<std::vector<double>> E_q, E_q_T, Dxi_q, Dxi_q_120, matE_q_T;
    std::vector<double> lam, mu;               // E x Nbf (P_N projected)
    std::vector<double> lamWJ, muWJ, JInvT;    // E x Nq (x D*D)
    std::vector<double> lam_q0, mu_q0, lam_q1, mu_q1, JInvT0, JInvT1; // F x nq (x D*D)
    std::vector<double> penalty;
    void setup(Mesh const& mesh, unsigned N, std::vector<double> const* nodes,
               double const* lam_Q, double const* mu_Q);
    std::size_t block_size() const override { return Nbf * Q; }
    void apply(double const* x, double* y, int nthreads) const override;
    void apply_element(std::size_t elNo, double const* x, double* y) const;
    void assemble_dense(double* A) const override;
    void rhs(double const* force_Q, double const* dirichlet_q, double* b) const override;
    // traction on a facet (compute_traction / compute_traction_bnd), t(p,q) D x nq
    void traction(std::size_t fctNo, double const* x, double const* f_q, double* t) const;
    // d/d(lam_Q, mu_Q) of psi^T A(lam, mu) u: the sensitivity of the bilinear form to the
    // per-quadrature-point Lame fields given to setup (through the L2 projection, the volume
    // weights, the facet traces and the penalty).  g_lam, g_mu: E x Nq.
    void material_gradient(double const* u, double const* psi, double* g_lam, double* g_mu) const;
    // apply_inverse_mass with density 1 (elasticity.py:236-238; weights -W/|J|, Elasticity.cpp:227-253):
    // y_e = Mhat^-1 E_Q diag(-W/|J|) E_Q^T Mhat^-1 x_e
    void inverse_mass(double const* x, double* y) const;
    // wave_rhs (Elasticity.cpp:865-880): apply_<true> with the facet data f_q (F x Q x nq, may be
    // null), then apply_inverse_mass
    void wave_rhs(double const* x, double const* f_q, double* y) const;
    unsigned degree = 0;
    FaultAdapter fault;
};

// NodalInterpolation::assemble (app/localoperator/NodalInterpolation.cpp:8-38): I (n_to x n_from,
// column-major) = V_to[:, :n_from] VInv_from, the from-degree nodal basis at the to-degree nodes.
std::vector<double> nodal_interpolation(unsigned D, unsigned to_degree, unsigned from_degree);

// app/localoperator/Poisson.cpp (modal Dubiner solution space)
struct Poisson : LocalOp {
    DGCommon c;
    RefElement space, matSpace;
    unsigned Nbf = 0, Nq = 0, nq = 0;
    std::vector<double> E_Q, Dxi_Q, matE_Q_T;
    std::vector<std::vector<double>> E_q, E_q_T, Dxi_q, Dxi_q_120, matE_q_T;
    std::vector<double> K;          // E x Nbf
    std::vector<double> JWK;        // E x Nq
    std::vector<double> KG0, KG1;   // F x nq x D*D
    std::vector<double> penalty;
    void setup(Mesh const& mesh, unsigned N, std::vector<double> const* nodes, double const* K_Q);
    std::size_t block_size() const override { return Nbf; }
    void apply(double const* x, double* y, int nthreads) const override;
    void apply_element(std::size_t elNo, double const* x, double* y) const;
    void assemble_dense(double* A) const override;
    void rhs(double const* force_Q, double const* dirichlet_q, double* b) const override;
};

} // namespace oracle
