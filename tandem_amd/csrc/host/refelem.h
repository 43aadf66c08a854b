// Reference-element machinery of the MI355X DG operator: orthogonal (Dubiner) polynomials,
// warp-and-blend nodes, simplex quadrature, and the nodal/modal evaluation tables the kernels use.
#pragma once

#include <array>
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace tdg {

// Multi-indices of total degree <= N in tandem's enumeration order (src/util/Combinatorics.h:46-110).
std::vector<std::array<unsigned, 3>> degree_indices(unsigned D, unsigned N);
std::size_t num_basis(unsigned D, unsigned N);

// Dubiner polynomial and gradient on the unit simplex (src/basis/Functions.cpp:88-149).
double dubiner(unsigned D, std::array<unsigned, 3> const& i, double const* xi);
void dubiner_grad(unsigned D, std::array<unsigned, 3> const& i, double const* xi, double* g);

struct QuadRule {
    unsigned dim = 0;
    std::vector<double> points; // dim per point
    std::vector<double> weights;
    std::size_t size() const { return weights.size(); }
};
// The smallest rule on the unit simplex exact to minOrder (src/quadrules/AutoRule.h:41-51).
QuadRule simplex_rule(unsigned D, unsigned minOrder);
// Collapsed Gauss-Jacobi rule with n^D points (src/quadrules/TensorProductRule.cpp).
QuadRule collapsed_rule(unsigned D, unsigned minOrder);
// Path of the tabulated rule data file; must be set before the first simplex_rule() call.
void set_rule_table_path(std::string const& path);

// Warp-and-blend nodes on the unit simplex (src/basis/WarpAndBlend.cpp:45-184).
std::vector<double> warp_blend_nodes(unsigned D, unsigned N);

// Nodal (Lagrange on warp-and-blend nodes) or modal (Dubiner) basis of degree N.
class Basis {
public:
    Basis(unsigned D, unsigned N, bool nodal);
    unsigned dim() const { return D_; }
    unsigned degree() const { return N_; }
    unsigned size() const { return nbf_; }
    bool nodal() const { return nodal_; }
    std::vector<double> const& nodes() const { return nodes_; }
    // values(k, p) column-major (nbf x np); grads(k, d, p) column-major (nbf x D x np)
    std::vector<double> values(std::vector<double> const& pts) const;
    std::vector<double> grads(std::vector<double> const& pts) const;

private:
    unsigned D_, N_, nbf_;
    bool nodal_;
    std::vector<double> nodes_;
    std::vector<double> vinv_; // inverse Vandermonde (nodal only), column-major
};

// Reference-simplex facet data: vertex lists in tandem's facet order (Simplex::downward),
// outward reference normals and the facet parametrisation xi = (1-sum chi) v0 + sum chi_d v_{d+1}.
struct RefSimplex {
    unsigned D;
    std::vector<std::array<int, 4>> faces; // face vertices [0..D-1], missing vertex at [D]
    std::vector<std::array<double, 3>> normals;
    explicit RefSimplex(unsigned D);
    std::vector<double> facet_points(unsigned f, std::vector<double> const& chi) const;
    // Affine map Phi_f of the reference simplex onto itself with Phi_f(facet 0 point) = facet f point;
    // returns its (constant) Jacobian, D x D column-major: columns 0 .. D-2 are the edges of facet f
    // from its first vertex (tangential), column D-1 points from it to the opposite vertex.
    std::array<double, 9> face_map_jacobian(unsigned f) const;
    // ... and the inverse of that Jacobian.
    std::array<double, 9> face_map_jacobian_inverse(unsigned f) const;
};

// Dense LU solve with partial pivoting (column-major), used by the host precompute.
void dense_solve(std::vector<double> A, int n, double* b, int nrhs);

} // namespace tdg
