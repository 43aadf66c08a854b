// Reference-element machinery (host, FP64).  See refelem.h.
#include "refelem.h"

#include <algorithm>
#include <cmath>
#include <fstream>
#include <limits>
#include <map>
#include <mutex>
#include <sstream>
#include <stdexcept>

namespace tdg {

// ------------------------------------------------------------------------ combinatorics
std::size_t num_basis(unsigned D, unsigned N) {
    std::size_t r = 1;
    for (unsigned i = 1; i <= D; ++i) r = r * (N + i) / i;
    return r;
}

// tandem enumerates all D-tuples with sum <= N by non-decreasing "markers" m_0 <= ... <= m_{D-1};
// tuple entry D-1-d is m_d - m_{d-1} (Combinatorics.h:46-110).
std::vector<std::array<unsigned, 3>> degree_indices(unsigned D, unsigned N) {
    std::vector<std::array<unsigned, 3>> out;
    std::array<unsigned, 3> m{0, 0, 0};
    for (;;) {
        std::array<unsigned, 3> t{0, 0, 0};
        unsigned prev = 0;
        for (unsigned d = 0; d < D; ++d) {
            t[D - 1 - d] = m[d] - prev;
            prev = m[d];
        }
        out.push_back(t);
        unsigned d = 0;
        while (d + 1 < D && m[d] == m[d + 1]) ++d;
        ++m[d];
        for (unsigned e = 0; e < d; ++e) m[e] = 0;
        if (m[D - 1] > N) return out;
    }
}

// ------------------------------------------------------------------------ Jacobi polynomials
// Three-term recurrence for P_n^{(a,b)} (Functions.cpp:6-23).
static double jacobi(unsigned n, unsigned a, unsigned b, double x) {
    double p0 = 1.0;
    if (n == 0) return p0;
    double p1 = 0.5 * a - 0.5 * b + (1.0 + 0.5 * (a + b)) * x;
    double ab2 = double(a * a) - double(b * b);
    for (unsigned m = 2; m <= n; ++m) {
        double c = 2.0 * m + a + b;
        double pn = ((c - 1.0) * (ab2 + c * (c - 2.0) * x) * p1 -
                     2.0 * (m + a - 1.0) * (m + b - 1.0) * c * p0) /
                    (2.0 * m * (m + a + b) * (c - 2.0));
        p0 = p1;
        p1 = pn;
    }
    return p1;
}
static double jacobi_d(unsigned n, unsigned a, unsigned b, double x) {
    return n == 0 ? 0.0 : 0.5 * (n + a + b + 1.0) * jacobi(n - 1, a + 1, b + 1, x);
}

// Homogenised Jacobi K(x, y) = P_n(x/y) y^n with its partial derivatives (Functions.cpp:29-86).
struct HJ {
    double v, dx, dy;
};
static HJ hjacobi(unsigned n, unsigned a, unsigned b, double x, double y) {
    if (n == 0) return {1.0, 0.0, 0.0};
    double v0 = 1.0, dx0 = 0.0, dy0 = 0.0;
    double v1 = (0.5 * a - 0.5 * b) * y + (1.0 + 0.5 * (a + b)) * x;
    double dx1 = 1.0 + 0.5 * (a + b);
    double dy1 = 0.5 * (double(a) - double(b));
    for (unsigned m = 2; m <= n; ++m) {
        double c0 = 2.0 * m + a + b;
        double f1 = c0 - 1.0, f2 = double(a * a) - double(b * b), f3 = c0 * (c0 - 2.0);
        double f4 = 2.0 * (m + a - 1.0) * (m + b - 1.0) * c0, f5 = 2.0 * m * (m + a + b) * (c0 - 2.0);
        double lin = f2 * y + f3 * x;
        double v2 = (f1 * lin * v1 - f4 * y * y * v0) / f5;
        double dx2 = (f1 * (f3 * v1 + lin * dx1) - f4 * y * y * dx0) / f5;
        double dy2 = (f1 * (f2 * v1 + lin * dy1) - f4 * (2.0 * y * v0 + y * y * dy0)) / f5;
        v0 = v1;
        dx0 = dx1;
        dy0 = dy1;
        v1 = v2;
        dx1 = dx2;
        dy1 = dy2;
    }
    return {v1, dx1, dy1};
}

double dubiner(unsigned D, std::array<unsigned, 3> const& i, double const* xi) {
    if (D == 1) return jacobi(i[0], 0, 0, 2.0 * xi[0] - 1.0);
    if (D == 2) {
        return hjacobi(i[0], 0, 0, 2.0 * xi[0] - 1.0 + xi[1], 1.0 - xi[1]).v *
               hjacobi(i[1], 2 * i[0] + 1, 0, 2.0 * xi[1] - 1.0, 1.0).v;
    }
    return hjacobi(i[0], 0, 0, 2.0 * xi[0] - 1.0 + xi[1] + xi[2], 1.0 - xi[1] - xi[2]).v *
           hjacobi(i[1], 2 * i[0] + 1, 0, 2.0 * xi[1] - 1.0 + xi[2], 1.0 - xi[2]).v *
           hjacobi(i[2], 2 * i[0] + 2 * i[1] + 2, 0, 2.0 * xi[2] - 1.0, 1.0).v;
}

void dubiner_grad(unsigned D, std::array<unsigned, 3> const& i, double const* xi, double* g) {
    if (D == 1) {
        g[0] = jacobi_d(i[0], 0, 0, 2.0 * xi[0] - 1.0);
        return;
    }
    if (D == 2) {
        HJ a = hjacobi(i[0], 0, 0, 2.0 * xi[0] - 1.0 + xi[1], 1.0 - xi[1]);
        HJ b = hjacobi(i[1], 2 * i[0] + 1, 0, 2.0 * xi[1] - 1.0, 1.0);
        // chain rule: x-argument of a depends on (2, 1), y-argument on (0, -1); b's x on (0, 2)
        g[0] = (a.dx * 2.0 + a.dy * 0.0) * b.v + a.v * b.dx * 0.0;
        g[1] = (a.dx * 1.0 + a.dy * -1.0) * b.v + a.v * b.dx * 2.0;
        return;
    }
    HJ a = hjacobi(i[0], 0, 0, 2.0 * xi[0] - 1.0 + xi[1] + xi[2], 1.0 - xi[1] - xi[2]);
    HJ b = hjacobi(i[1], 2 * i[0] + 1, 0, 2.0 * xi[1] - 1.0 + xi[2], 1.0 - xi[2]);
    HJ c = hjacobi(i[2], 2 * i[0] + 2 * i[1] + 2, 0, 2.0 * xi[2] - 1.0, 1.0);
    // d/dxi_d of each homogenised argument pair (x, y):
    //   a: x = 2xi0-1+xi1+xi2, y = 1-xi1-xi2;  b: x = 2xi1-1+xi2, y = 1-xi2;  c: x = 2xi2-1
    const double ax[3] = {2.0, 1.0, 1.0}, ay[3] = {0.0, -1.0, -1.0};
    const double bx[3] = {0.0, 2.0, 1.0}, by[3] = {0.0, 0.0, -1.0};
    const double cx[3] = {0.0, 0.0, 2.0};
    for (int d = 0; d < 3; ++d) {
        g[d] = (a.dx * ax[d] + a.dy * ay[d]) * b.v * c.v + a.v * (b.dx * bx[d] + b.dy * by[d]) * c.v +
               a.v * b.v * (c.dx * cx[d]);
    }
}

// ------------------------------------------------------------------------ dense algebra
void dense_solve(std::vector<double> A, int n, double* b, int nrhs) {
    for (int k = 0; k < n; ++k) {
        int p = k;
        double best = std::fabs(A[k + k * n]);
        for (int i = k + 1; i < n; ++i) {
            double v = std::fabs(A[i + k * n]);
            if (v > best) {
                best = v;
                p = i;
            }
        }
        if (best == 0.0) throw std::runtime_error("dense_solve: singular matrix");
        if (p != k) {
            for (int j = 0; j < n; ++j) std::swap(A[k + j * n], A[p + j * n]);
            for (int j = 0; j < nrhs; ++j) std::swap(b[k + j * n], b[p + j * n]);
        }
        double inv = A[k + k * n];
        for (int i = k + 1; i < n; ++i) {
            double l = A[i + k * n] / inv;
            if (l == 0.0) continue;
            for (int j = k + 1; j < n; ++j) A[i + j * n] -= l * A[k + j * n];
            for (int j = 0; j < nrhs; ++j) b[i + j * n] -= l * b[k + j * n];
        }
    }
    for (int j = 0; j < nrhs; ++j)
        for (int i = n - 1; i >= 0; --i) {
            double s = b[i + j * n];
            for (int k = i + 1; k < n; ++k) s -= A[i + k * n] * b[k + j * n];
            b[i + j * n] = s / A[i + i * n];
        }
}

// ------------------------------------------------------------------------ quadrature
// Gauss-Jacobi points/weights on [-1,1] for weight (1-x)^a (1+x)^b by Newton iteration from the
// cosine initial guess (GaussJacobi.cpp:15-39).
static void gauss_jacobi(unsigned n, unsigned a, unsigned b, std::vector<double>& x, std::vector<double>& w) {
    x.resize(n);
    w.resize(n);
    auto rp = [](uint64_t from, uint64_t to) {
        uint64_t p = 1;
        for (; from <= to; ++from) p *= from;
        return p;
    };
    double scale = -(2.0 * n + a + b + 2) * rp(n + 1, n + b) * (1 << (a + b)) /
                   ((n + a + b + 1.0) * rp(n + a + 1, n + a + b) * (n + 1));
    const double tol = 10.0 * std::numeric_limits<double>::epsilon();
    for (unsigned i = 1; i <= n; ++i) {
        double r = std::cos(M_PI * (0.5 * a + i - 0.25) / (0.5 * (1.0 + a + b) + n));
        double p = jacobi(n, a, b, r), dp = jacobi_d(n, a, b, r);
        for (unsigned it = 1; std::fabs(p) > tol && it < 100; ++it) {
            r -= p / dp;
            p = jacobi(n, a, b, r);
            dp = jacobi_d(n, a, b, r);
        }
        x[i - 1] = r;
        w[i - 1] = scale / (jacobi(n + 1, a, b, r) * dp);
    }
}

QuadRule collapsed_rule(unsigned D, unsigned minOrder) {
    unsigned n = 1 + minOrder / 2;
    std::vector<double> x[3], w[3];
    for (unsigned a = 0; a < D; ++a) {
        gauss_jacobi(n, a, 0, x[a], w[a]);
        for (auto& v : x[a]) v = 0.5 * v + 0.5;
        for (auto& v : w[a]) v *= 0.5;
    }
    QuadRule r;
    r.dim = D;
    if (D == 1) {
        r.points = x[0];
        r.weights = w[0];
    } else if (D == 2) {
        for (unsigned i = 0; i < n; ++i)
            for (unsigned j = 0; j < n; ++j) {
                r.points.insert(r.points.end(), {x[1][i], x[0][j] * (1.0 - x[1][i])});
                r.weights.push_back(0.5 * w[1][i] * w[0][j]);
            }
    } else {
        for (unsigned i = 0; i < n; ++i)
            for (unsigned j = 0; j < n; ++j)
                for (unsigned k = 0; k < n; ++k) {
                    r.points.insert(r.points.end(), {x[2][i], x[1][j] * (1 - x[2][i]),
                                                     x[0][k] * (1 - x[1][j]) * (1 - x[2][i])});
                    r.weights.push_back(0.125 * w[2][i] * w[1][j] * w[0][k]);
                }
    }
    return r;
}

char const* rule_table_text();

namespace {
std::mutex g_table_mutex;
std::string g_table_path;
bool g_tables_loaded = false;
// family ("WV"/"JS") + dim -> order -> rule in the published coordinates
std::map<std::pair<std::string, unsigned>, std::map<unsigned, QuadRule>> g_tables;

void load_tables() {
    if (g_tables_loaded) return;
    std::ifstream file;
    std::istringstream blob;
    std::istream* src = &blob;
    if (g_table_path.empty()) {
        blob.str(rule_table_text());
    } else {
        file.open(g_table_path);
        if (!file) throw std::runtime_error("cannot open quadrature table '" + g_table_path + "'");
        src = &file;
    }
    std::istream& f = *src;
    std::string line;
    while (std::getline(f, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream hs(line);
        std::string tag, fam;
        unsigned dim = 0, order = 0, n = 0;
        hs >> tag >> fam >> dim >> order >> n;
        if (tag != "rule") throw std::runtime_error("malformed quadrature table line");
        QuadRule r;
        r.dim = dim;
        for (unsigned p = 0; p < n; ++p) {
            if (!std::getline(f, line)) throw std::runtime_error("truncated quadrature table");
            std::istringstream ps(line);
            std::string tok;
            for (unsigned d = 0; d <= dim; ++d) {
                ps >> tok;
                double v = std::strtod(tok.c_str(), nullptr);
                if (d < dim) r.points.push_back(v);
                else r.weights.push_back(v);
            }
        }
        g_tables[{fam, dim}][order] = std::move(r);
    }
    g_tables_loaded = true;
}

// smallest tabulated rule of order >= o, mapped to the unit simplex (TabulatedRules.h:66-104)
std::size_t tab_size(std::string const& fam, unsigned D, unsigned o) {
    auto& t = g_tables[{fam, D}];
    auto it = t.lower_bound(o);
    return it == t.end() ? std::numeric_limits<std::size_t>::max() : it->second.size();
}
QuadRule tab_rule(std::string const& fam, unsigned D, unsigned o) {
    QuadRule r = g_tables[{fam, D}].lower_bound(o)->second;
    if (fam == "JS") {
        for (auto& w : r.weights) w /= 6.0;
    } else {
        for (auto& x : r.points) x = 0.5 * x + 0.5;
        for (auto& w : r.weights) w /= double(1u << D);
    }
    return r;
}
} // namespace

void set_rule_table_path(std::string const& path) {
    std::lock_guard<std::mutex> lk(g_table_mutex);
    if (path != g_table_path) {
        g_table_path = path;
        g_tables.clear();
        g_tables_loaded = false;
    }
}

QuadRule simplex_rule(unsigned D, unsigned o) {
    if (D < 2) return collapsed_rule(D, o);
    std::lock_guard<std::mutex> lk(g_table_mutex);
    load_tables();
    std::size_t tp = 1;
    for (unsigned d = 0; d < D; ++d) tp *= (1 + o / 2);
    // candidates in tandem's preference order; ties go to the earlier one
    std::vector<std::string> fams = D == 2 ? std::vector<std::string>{"WV"} : std::vector<std::string>{"JS", "WV"};
    std::string best;
    std::size_t bsize = std::numeric_limits<std::size_t>::max();
    for (auto const& fam : fams) {
        std::size_t s = tab_size(fam, D, o);
        if (best.empty() || s < bsize) {
            best = fam;
            bsize = s;
        }
    }
    if (bsize <= tp) {
        if (bsize == std::numeric_limits<std::size_t>::max())
            throw std::runtime_error("no quadrature rule of the requested order");
        return tab_rule(best, D, o);
    }
    if (D == 3 && o > 13)
        throw std::runtime_error("3D quadrature orders above 13 are not tabulated in this build");
    return collapsed_rule(D, o);
}

// ------------------------------------------------------------------------ warp & blend
static double wab_alpha(unsigned D, unsigned N) {
    static const double t2[] = {0.0000,  0.0000,  1.4152,  0.1001,  0.2751,  0.9800,  1.0999, 1.2832,
                                1.3648,  1.4773,  1.4959,  1.5743,  1.5770,  1.6223,  1.6258, 1.64831,
                                1.76303, 1.92727, 1.92962, 1.91251, 1.91289, 1.89479, 1.89255,
                                1.88016, 1.87555, 1.86613, 1.86106, 1.85368, 1.84874, 1.84319};
    static const double t3[] = {0.0,      0.0,      0.0,     0.1002,  1.1332,  1.5608,  1.3413,
                                1.2577,   1.1603,   1.10153, 0.6080,  0.4523,  0.8856,  0.8717,
                                0.9655,   0.982344, 0.827846, 0.700081, 0.725887};
    if (D == 2) return N <= 30 ? t2[N - 1] : 5.0 / 3.0;
    return N <= 19 ? t3[N - 1] : 5.0 / 3.0;
}

static std::vector<double> lgl(unsigned n) {
    std::vector<double> x, w;
    gauss_jacobi(n - 2, 1, 1, x, w);
    std::vector<double> out{-1.0};
    out.insert(out.end(), x.rbegin(), x.rend());
    out.push_back(1.0);
    return out;
}

// 1D warp: interpolate (LGL - equidistant) displacement, divided by 1 - r^2 away from the ends.
static std::vector<double> warp1d(unsigned N, std::vector<double> const& r) {
    unsigned m = N + 1;
    std::vector<double> eq(m), gl = lgl(m), VT(m * m), L(m * r.size());
    for (unsigned i = 0; i < m; ++i) eq[i] = 2 * i / double(N) - 1.0;
    for (unsigned i = 0; i < m; ++i)
        for (unsigned j = 0; j < m; ++j) VT[j + i * m] = jacobi(j, 0, 0, eq[i]);
    for (std::size_t p = 0; p < r.size(); ++p)
        for (unsigned i = 0; i < m; ++i) L[i + p * m] = jacobi(i, 0, 0, r[p]);
    dense_solve(VT, m, L.data(), r.size());
    std::vector<double> out(r.size(), 0.0);
    for (std::size_t p = 0; p < r.size(); ++p) {
        for (unsigned i = 0; i < m; ++i) out[p] += L[i + p * m] * (gl[i] - eq[i]);
        if (std::fabs(r[p]) < 1.0 - std::numeric_limits<double>::epsilon()) out[p] /= 1.0 - r[p] * r[p];
    }
    return out;
}

namespace {
const double kS3 = std::sqrt(3.0), kS6 = std::sqrt(6.0);
const double kTri[3][2] = {{-1.0, -1.0 / kS3}, {1.0, -1.0 / kS3}, {0.0, 2.0 / kS3}};
const double kTet[4][3] = {{-1.0, -1.0 / kS3, -1.0 / kS6},
                           {1.0, -1.0 / kS3, -1.0 / kS6},
                           {0.0, 2.0 / kS3, -1.0 / kS6},
                           {0.0, 0.0, 3.0 / kS6}};
} // namespace

// Triangle warp in equilateral coordinates from barycentric columns b[0..2] (n points each).
static void tri_warp(unsigned N, double alpha, std::vector<double> const* b, std::size_t n,
                     std::vector<double>& wx, std::vector<double>& wy) {
    wx.assign(n, 0.0);
    wy.assign(n, 0.0);
    const int edge[3][3] = {{0, 1, 2}, {0, 2, 1}, {1, 2, 0}};
    for (auto const& e : edge) {
        std::vector<double> r(n);
        for (std::size_t p = 0; p < n; ++p) r[p] = b[e[1]][p] - b[e[0]][p];
        auto w = warp1d(N, r);
        double tx = kTri[e[1]][0] - kTri[e[0]][0], ty = kTri[e[1]][1] - kTri[e[0]][1];
        double tl = std::sqrt(tx * tx + ty * ty);
        tx /= tl;
        ty /= tl;
        for (std::size_t p = 0; p < n; ++p) {
            double bl = 4.0 * b[e[0]][p] * b[e[1]][p];
            bl = bl * (1.0 + (alpha * alpha) * b[e[2]][p] * b[e[2]][p]);
            wx[p] += w[p] * bl * tx;
            wy[p] += w[p] * bl * ty;
        }
    }
}

// solve x = v0 + A xi for xi, A columns v_{d+1} - v0
static void equi_to_unit(unsigned D, double const (*v)[3], double const* x, double* xi) {
    std::vector<double> A(D * D);
    for (unsigned i = 0; i < D; ++i)
        for (unsigned j = 0; j < D; ++j) A[i + j * D] = v[j + 1][i] - v[0][i];
    for (unsigned i = 0; i < D; ++i) xi[i] = x[i] - v[0][i];
    dense_solve(A, D, xi, 1);
}

std::vector<double> warp_blend_nodes(unsigned D, unsigned N) {
    std::size_t n = num_basis(D, N);
    std::vector<double> out(n * D, 0.0);
    if (N == 0) {
        for (unsigned d = 0; d < D; ++d) out[d] = D == 1 ? 0.5 : 1.0 / (D + 1);
        return out;
    }
    if (D == 1) {
        auto g = lgl(N + 1);
        for (std::size_t p = 0; p < n; ++p) out[p] = (g[p] + 1.0) / 2.0;
        return out;
    }
    auto idx = degree_indices(D, N);
    double alpha = wab_alpha(D, N);
    std::vector<double> bary[4];
    for (unsigned k = 0; k <= D; ++k) bary[k].resize(n);
    for (std::size_t p = 0; p < n; ++p) {
        double s = 0.0;
        for (unsigned d = 0; d < D; ++d) {
            bary[d + 1][p] = idx[p][d] / double(N);
            s += bary[d + 1][p];
        }
        bary[0][p] = 1.0 - s;
    }
    std::vector<double> X(n * 3, 0.0);
    if (D == 2) {
        std::vector<double> wx, wy;
        tri_warp(N, alpha, bary, n, wx, wy);
        double V[3][3] = {{kTri[0][0], kTri[0][1], 0}, {kTri[1][0], kTri[1][1], 0}, {kTri[2][0], kTri[2][1], 0}};
        for (std::size_t p = 0; p < n; ++p) {
            double x[2];
            for (int d = 0; d < 2; ++d) {
                double lin = 0.0;
                for (int k = 0; k < 3; ++k) lin += bary[k][p] * kTri[k][d];
                x[d] = lin + (d == 0 ? wx[p] : wy[p]);
            }
            equi_to_unit(2, V, x, &out[2 * p]);
        }
        return out;
    }
    // tetrahedron: blend the face warps (Hesthaven & Warburton, Nodal DG Methods, sec. 10.1)
    const int fc[4][4] = {{0, 1, 2, 3}, {0, 1, 3, 2}, {0, 2, 3, 1}, {1, 2, 3, 0}};
    for (auto const& f : fc) {
        double t0[3], t1[3], l0 = 0, l1 = 0;
        for (int d = 0; d < 3; ++d) {
            t0[d] = kTet[f[1]][d] - kTet[f[0]][d];
            t1[d] = kTet[f[2]][d] - 0.5 * (kTet[f[1]][d] + kTet[f[0]][d]);
            l0 += t0[d] * t0[d];
            l1 += t1[d] * t1[d];
        }
        l0 = std::sqrt(l0);
        l1 = std::sqrt(l1);
        std::vector<double> fb[3] = {bary[f[0]], bary[f[1]], bary[f[2]]};
        std::vector<double> wx, wy;
        tri_warp(N, alpha, fb, n, wx, wy);
        for (std::size_t p = 0; p < n; ++p) {
            double lm = bary[f[3]][p];
            double blend = fb[0][p] * fb[1][p] * fb[2][p];
            double denom = (fb[0][p] + 0.5 * lm) * (fb[1][p] + 0.5 * lm) * (fb[2][p] + 0.5 * lm);
            double keep = 1.0;
            if (std::fabs(denom) < std::numeric_limits<double>::epsilon()) {
                blend = 1.0;
                keep = 0.0;
            } else {
                blend = (1.0 + alpha * alpha * lm * lm) * blend / denom;
            }
            for (int d = 0; d < 3; ++d)
                X[3 * p + d] = keep * X[3 * p + d] + blend * (wx[p] * t0[d] / l0 + wy[p] * t1[d] / l1);
        }
    }
    for (std::size_t p = 0; p < n; ++p) {
        double x[3];
        for (int d = 0; d < 3; ++d) {
            double lin = 0.0;
            for (int k = 0; k < 4; ++k) lin += bary[k][p] * kTet[k][d];
            x[d] = X[3 * p + d] + lin;
        }
        equi_to_unit(3, kTet, x, &out[3 * p]);
    }
    return out;
}

// ------------------------------------------------------------------------ bases
Basis::Basis(unsigned D, unsigned N, bool nodal) : D_(D), N_(N), nbf_(num_basis(D, N)), nodal_(nodal) {
    if (!nodal_) return;
    nodes_ = warp_blend_nodes(D, N);
    auto idx = degree_indices(D, N);
    std::vector<double> V(nbf_ * nbf_);
    for (unsigned i = 0; i < nbf_; ++i)
        for (unsigned j = 0; j < nbf_; ++j) V[i + j * nbf_] = dubiner(D, idx[j], &nodes_[i * D]);
    vinv_.assign(nbf_ * nbf_, 0.0);
    for (unsigned i = 0; i < nbf_; ++i) vinv_[i + i * nbf_] = 1.0;
    dense_solve(V, nbf_, vinv_.data(), nbf_);
}

std::vector<double> Basis::values(std::vector<double> const& pts) const {
    std::size_t np = pts.size() / D_;
    auto idx = degree_indices(D_, N_);
    std::vector<double> m(nbf_ * np);
    for (std::size_t p = 0; p < np; ++p)
        for (unsigned j = 0; j < nbf_; ++j) m[j + p * nbf_] = dubiner(D_, idx[j], &pts[p * D_]);
    if (!nodal_) return m;
    // nodal phi_i = sum_j Vinv(j, i) P_j
    std::vector<double> out(nbf_ * np, 0.0);
    for (std::size_t p = 0; p < np; ++p)
        for (unsigned i = 0; i < nbf_; ++i) {
            double s = 0.0;
            for (unsigned j = 0; j < nbf_; ++j) s += vinv_[j + i * nbf_] * m[j + p * nbf_];
            out[i + p * nbf_] = s;
        }
    return out;
}

std::vector<double> Basis::grads(std::vector<double> const& pts) const {
    std::size_t np = pts.size() / D_;
    auto idx = degree_indices(D_, N_);
    std::vector<double> m(nbf_ * D_ * np);
    double g[3];
    for (std::size_t p = 0; p < np; ++p)
        for (unsigned j = 0; j < nbf_; ++j) {
            dubiner_grad(D_, idx[j], &pts[p * D_], g);
            for (unsigned d = 0; d < D_; ++d) m[j + nbf_ * (d + D_ * p)] = g[d];
        }
    if (!nodal_) return m;
    std::vector<double> out(m.size(), 0.0);
    for (std::size_t c = 0; c < D_ * np; ++c)
        for (unsigned i = 0; i < nbf_; ++i) {
            double s = 0.0;
            for (unsigned j = 0; j < nbf_; ++j) s += vinv_[j + i * nbf_] * m[j + c * nbf_];
            out[i + c * nbf_] = s;
        }
    return out;
}

// ------------------------------------------------------------------------ reference simplex
RefSimplex::RefSimplex(unsigned D_) : D(D_) {
    const double rv[4][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    if (D == 2) faces = {{0, 1, 2, -1}, {0, 2, 1, -1}, {1, 2, 0, -1}};
    else faces = {{0, 1, 2, 3}, {0, 1, 3, 2}, {0, 2, 3, 1}, {1, 2, 3, 0}};
    for (auto const& f : faces) {
        std::array<double, 3> n{0, 0, 0};
        int miss = f[D];
        if (D == 2) {
            n[0] = -(rv[f[1]][1] - rv[f[0]][1]);
            n[1] = rv[f[1]][0] - rv[f[0]][0];
        } else {
            double a[3], b[3];
            for (int d = 0; d < 3; ++d) {
                a[d] = rv[f[1]][d] - rv[f[0]][d];
                b[d] = rv[f[2]][d] - rv[f[0]][d];
            }
            n = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
        }
        double dot = 0.0;
        for (unsigned d = 0; d < D; ++d) dot += n[d] * (rv[miss][d] - rv[f[0]][d]);
        if (dot > 0)
            for (auto& v : n) v = -v;
        normals.push_back(n);
    }
}

std::vector<double> RefSimplex::facet_points(unsigned f, std::vector<double> const& chi) const {
    const double rv[4][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    std::size_t np = chi.size() / (D - 1);
    std::vector<double> xi(np * D);
    for (std::size_t p = 0; p < np; ++p) {
        double s = 0.0;
        for (unsigned d = 0; d + 1 < D; ++d) s += chi[p * (D - 1) + d];
        for (unsigned i = 0; i < D; ++i) {
            double v = (1.0 - s) * rv[faces[f][0]][i];
            for (unsigned d = 0; d + 1 < D; ++d) v = v + chi[p * (D - 1) + d] * rv[faces[f][d + 1]][i];
            xi[p * D + i] = v;
        }
    }
    return xi;
}

std::array<double, 9> RefSimplex::face_map_jacobian(unsigned f) const {
    const double rv[4][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    std::array<double, 9> J{};
    for (unsigned d = 0; d < D; ++d)
        for (unsigned i = 0; i < D; ++i) J[i + d * D] = rv[faces[f][d + 1]][i] - rv[faces[f][0]][i];
    return J;
}

std::array<double, 9> RefSimplex::face_map_jacobian_inverse(unsigned f) const {
    const double rv[4][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    // vertex map sigma: 0 -> faces[f][0], 1 -> faces[f][1], ..., D -> missing vertex
    std::vector<double> J(D * D);
    for (unsigned d = 0; d < D; ++d)
        for (unsigned i = 0; i < D; ++i) J[i + d * D] = rv[faces[f][d + 1]][i] - rv[faces[f][0]][i];
    std::vector<double> I(D * D, 0.0);
    for (unsigned i = 0; i < D; ++i) I[i + i * D] = 1.0;
    dense_solve(J, D, I.data(), D);
    std::array<double, 9> out{};
    for (unsigned i = 0; i < D * D; ++i) out[i] = I[i];
    return out;
}

} // namespace tdg
