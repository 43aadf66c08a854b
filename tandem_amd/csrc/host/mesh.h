// Simplex meshes and the DG facet topology of the MI355X operator.
#pragma once

#include <array>
#include <cstddef>
#include <cstdint>
#include <vector>

namespace tdg {

enum BCKind : int { BC_None = 0, BC_Natural = 1, BC_Fault = 3, BC_Dirichlet = 5 };

struct SimplexMesh {
    unsigned dim = 0;
    std::vector<double> vertices;      // dim per vertex
    std::vector<int64_t> elements;     // dim+1 per element, ascending vertex ids
    std::vector<int64_t> bnd_facets;   // dim per boundary facet, ascending
    std::vector<int> bnd_bc;
    std::size_t num_elements() const { return elements.size() / (dim + 1); }
};

// Box [0,1]^dim with n[d] cells per direction.  scheme 0: tandem's GenMesh tessellation (2 triangles
// per square, 5 alternating tets per cube, GenMesh.cpp:21-143); scheme 1: 6-tet Kuhn subdivision.
// bc[2d] / bc[2d+1] is the BC on the plane x_d = 0 / x_d = 1.
SimplexMesh box_mesh(unsigned dim, std::array<uint64_t, 3> const& n, int scheme, std::array<int, 6> const& bc);

// One facet of the mesh, seen from its two sides (FacetInfo.h:18-24); boundary facets have
// elem[1] == elem[0].
struct Facet {
    int64_t elem[2];
    int local[2];
    int bc;
};
// Per element and local face (SideInfo, FacetInfo.h:11-16).
struct Side {
    int64_t facet;
    int64_t nb;     // neighbour element (the element itself on a boundary facet)
    int side;       // which side of the facet this element is
    int nb_local;   // neighbour's local face number
    int bc;
};
struct Topology {
    std::vector<Facet> facets;
    std::vector<Side> sides; // (dim+1) per element
};
Topology build_topology(SimplexMesh const& m);

} // namespace tdg
