// Simplex meshes and DG facet topology (host).
#include "mesh.h"

#include <algorithm>
#include <stdexcept>
#include <unordered_map>

namespace tdg {

namespace {
struct KeyHash {
    std::size_t operator()(std::array<int64_t, 3> const& k) const noexcept {
        uint64_t h = 1469598103934665603ull;
        for (auto v : k) {
            h ^= static_cast<uint64_t>(v) + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
        }
        return static_cast<std::size_t>(h);
    }
};

// local faces of a simplex with ascending vertices, in tandem's order (Simplex.h:76-98)
const int kTriFaces[3][2] = {{0, 1}, {0, 2}, {1, 2}};
const int kTetFaces[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};

std::array<int64_t, 3> face_key(unsigned dim, int64_t const* el, int f) {
    std::array<int64_t, 3> k{-1, -1, -1};
    for (unsigned i = 0; i < dim; ++i) k[i] = el[dim == 2 ? kTriFaces[f][i] : kTetFaces[f][i]];
    return k;
}
} // namespace

SimplexMesh box_mesh(unsigned dim, std::array<uint64_t, 3> const& n, int scheme, std::array<int, 6> const& bc) {
    if (dim != 2 && dim != 3) throw std::runtime_error("box_mesh: dim must be 2 or 3");
    SimplexMesh m;
    m.dim = dim;
    uint64_t nx = n[0], ny = n[1], nz = dim == 3 ? n[2] : 1;
    uint64_t vx = nx + 1, vy = ny + 1, vz = dim == 3 ? nz + 1 : 1;
    auto vid = [&](uint64_t i, uint64_t j, uint64_t k) { return static_cast<int64_t>(i + vx * (j + vy * k)); };
    m.vertices.reserve(vx * vy * vz * dim);
    for (uint64_t k = 0; k < vz; ++k)
        for (uint64_t j = 0; j < vy; ++j)
            for (uint64_t i = 0; i < vx; ++i) {
                m.vertices.push_back(double(i) * (1.0 / double(nx)));
                m.vertices.push_back(double(j) * (1.0 / double(ny)));
                if (dim == 3) m.vertices.push_back(double(k) * (1.0 / double(nz)));
            }
    auto push = [&](std::initializer_list<int64_t> l) {
        std::array<int64_t, 4> s{};
        std::size_t c = 0;
        for (auto v : l) s[c++] = v;
        std::sort(s.begin(), s.begin() + c);
        m.elements.insert(m.elements.end(), s.begin(), s.begin() + c);
    };
    for (uint64_t k = 0; k < nz; ++k)
        for (uint64_t j = 0; j < ny; ++j)
            for (uint64_t i = 0; i < nx; ++i) {
                // corner c = bit0 (x) | bit1 (y) | bit2 (z)
                int64_t c[8];
                for (int b = 0; b < 8; ++b) c[b] = vid(i + (b & 1), j + ((b >> 1) & 1), k + ((b >> 2) & 1));
                bool odd = ((i & 1) ^ (j & 1) ^ (dim == 3 ? (k & 1) : 0)) != 0;
                if (dim == 2) {
                    if (odd) {
                        push({c[0], c[1], c[3]});
                        push({c[0], c[2], c[3]});
                    } else {
                        push({c[0], c[1], c[2]});
                        push({c[1], c[2], c[3]});
                    }
                } else if (scheme == 0) {
                    if (odd) {
                        push({c[0], c[1], c[2], c[4]});
                        push({c[1], c[2], c[3], c[7]});
                        push({c[2], c[4], c[6], c[7]});
                        push({c[1], c[4], c[5], c[7]});
                        push({c[2], c[4], c[1], c[7]});
                    } else {
                        push({c[0], c[2], c[3], c[6]});
                        push({c[0], c[1], c[3], c[5]});
                        push({c[0], c[4], c[5], c[6]});
                        push({c[3], c[5], c[6], c[7]});
                        push({c[0], c[3], c[5], c[6]});
                    }
                } else {
                    // Kuhn: one tet per axis permutation, walking 000 -> 111 along the axes
                    const int perm[6][2] = {{0, 1}, {0, 2}, {1, 0}, {1, 2}, {2, 0}, {2, 1}};
                    for (auto const& p : perm) {
                        int a = 1 << p[0];
                        int b = a | (1 << p[1]);
                        push({c[0], c[a], c[b], c[7]});
                    }
                }
            }
    // boundary facets: all vertices on one of the planes x_d = 0 or x_d = 1
    std::size_t ne = m.num_elements();
    std::vector<uint64_t> ncell{nx, ny, nz};
    std::vector<uint64_t> vdim{vx, vy, vz};
    auto grid = [&](int64_t v, unsigned d) {
        uint64_t u = static_cast<uint64_t>(v);
        if (d == 0) return u % vx;
        if (d == 1) return (u / vx) % vy;
        return u / (vx * vy);
    };
    unsigned nf = dim + 1;
    for (std::size_t e = 0; e < ne; ++e) {
        int64_t const* el = &m.elements[e * (dim + 1)];
        for (unsigned f = 0; f < nf; ++f) {
            auto key = face_key(dim, el, f);
            for (unsigned d = 0; d < dim; ++d) {
                uint64_t g0 = grid(key[0], d);
                bool same = true;
                for (unsigned i = 1; i < dim; ++i) same = same && grid(key[i], d) == g0;
                if (same && (g0 == 0 || g0 == ncell[d])) {
                    m.bnd_facets.insert(m.bnd_facets.end(), key.begin(), key.begin() + dim);
                    m.bnd_bc.push_back(bc[2 * d + (g0 == 0 ? 0 : 1)]);
                }
            }
        }
    }
    return m;
}

Topology build_topology(SimplexMesh const& m) {
    unsigned dim = m.dim, nf = dim + 1;
    std::size_t ne = m.num_elements();
    std::unordered_map<std::array<int64_t, 3>, int, KeyHash> bcmap;
    bcmap.reserve(m.bnd_bc.size() * 2);
    for (std::size_t b = 0; b < m.bnd_bc.size(); ++b) {
        std::array<int64_t, 3> k{-1, -1, -1};
        for (unsigned i = 0; i < dim; ++i) k[i] = m.bnd_facets[b * dim + i];
        std::sort(k.begin(), k.begin() + dim);
        bcmap[k] = m.bnd_bc[b];
    }
    Topology t;
    std::unordered_map<std::array<int64_t, 3>, int64_t, KeyHash> fid;
    fid.reserve(ne * nf);
    t.facets.reserve(ne * nf / 2 + 16);
    for (std::size_t e = 0; e < ne; ++e) {
        int64_t const* el = &m.elements[e * (dim + 1)];
        for (unsigned f = 0; f < nf; ++f) {
            auto key = face_key(dim, el, f);
            auto it = fid.find(key);
            if (it == fid.end()) {
                Facet fc;
                fc.elem[0] = fc.elem[1] = static_cast<int64_t>(e);
                fc.local[0] = fc.local[1] = static_cast<int>(f);
                auto b = bcmap.find(key);
                fc.bc = b == bcmap.end() ? BC_None : b->second;
                fid.emplace(key, static_cast<int64_t>(t.facets.size()));
                t.facets.push_back(fc);
            } else {
                Facet& fc = t.facets[it->second];
                if (fc.elem[1] != fc.elem[0]) throw std::runtime_error("mesh facet shared by more than two elements");
                fc.elem[1] = static_cast<int64_t>(e);
                fc.local[1] = static_cast<int>(f);
            }
        }
    }
    for (auto const& kv : bcmap)
        if (fid.find(kv.first) == fid.end())
            throw std::invalid_argument("a boundary facet is not a face of any element");
    t.sides.assign(ne * nf, Side{-1, -1, 0, 0, 0});
    for (std::size_t i = 0; i < t.facets.size(); ++i) {
        Facet const& fc = t.facets[i];
        int ns = fc.elem[0] == fc.elem[1] ? 1 : 2;
        for (int s = 0; s < ns; ++s) {
            Side& sd = t.sides[fc.elem[s] * nf + fc.local[s]];
            sd.facet = static_cast<int64_t>(i);
            sd.side = s;
            sd.nb = fc.elem[1 - s];
            sd.nb_local = fc.local[1 - s];
            sd.bc = fc.bc;
        }
    }
    return t;
}

} // namespace tdg
