// Oracle: polynomials, quadrature, warp-and-blend nodes and reference elements (test infrastructure).
#include "oracle.h"

#include <algorithm>
#include <cassert>
#include <cmath>
#include <fstream>
#include <limits>
#include <map>
#include <sstream>
#include <stdexcept>

namespace oracle {

// src/basis/Functions.cpp:6-23
double jacobi_p(unsigned n, unsigned a, unsigned b, double x) {
    if (n == 0) return 1.0;
    double Pm_2, Pm_1 = 1.0;
    double Pm = 0.5 * a - 0.5 * b + (1.0 + 0.5 * (a + b)) * x;
    double a2_b2 = static_cast<double>(a * a) - static_cast<double>(b * b);
    for (unsigned m = 2; m <= n; ++m) {
        Pm_2 = Pm_1;
        Pm_1 = Pm;
        Pm = ((2.0 * m + a + b - 1.0) * (a2_b2 + (2.0 * m + a + b) * (2.0 * m + a + b - 2.0) * x) *
                  Pm_1 -
              2.0 * (m + a - 1.0) * (m + b - 1.0) * (2.0 * m + a + b) * Pm_2) /
             (2.0 * m * (m + a + b) * (2.0 * m + a + b - 2.0));
    }
    return Pm;
}

// src/basis/Functions.cpp:25-27
double jacobi_p_derivative(unsigned n, unsigned a, unsigned b, double x) {
    return (n == 0) ? 0.0 : 0.5 * (n + a + b + 1.0) * jacobi_p(n - 1, a + 1, b + 1, x);
}

// src/basis/Functions.cpp:29-43 (recursion factors)
static std::array<double, 5> sf_factors(unsigned m, unsigned a, unsigned b) {
    double c0 = 2.0 * m + a + b;
    double c1 = c0 - 1.0;
    double c2 = static_cast<double>(a * a) - static_cast<double>(b * b);
    double c3 = c0 * (c0 - 2.0);
    double c4 = 2.0 * (m + a - 1.0) * (m + b - 1.0) * c0;
    double c5 = 2.0 * m * (m + a + b) * (c0 - 2.0);
    return {c1, c2, c3, c4, c5};
}

// src/basis/Functions.cpp:45-58: K^n_{a,b}(x, y) = P_n^{(a,b)}(x/y) y^n
double sf_jacobi_p(unsigned n, unsigned a, unsigned b, double x, double y) {
    if (n == 0) return 1.0;
    double Pm_2, Pm_1 = 1.0;
    double Pm = (0.5 * a - 0.5 * b) * y + (1.0 + 0.5 * (a + b)) * x;
    for (unsigned m = 2; m <= n; ++m) {
        Pm_2 = Pm_1;
        Pm_1 = Pm;
        auto c = sf_factors(m, a, b);
        Pm = (c[0] * (c[1] * y + c[2] * x) * Pm_1 - c[3] * y * y * Pm_2) / c[4];
    }
    return Pm;
}

// src/basis/Functions.cpp:60-86
std::array<double, 3> sf_jacobi_p_and_derivatives(unsigned n, unsigned a, unsigned b, double x,
                                                  double y) {
    if (n == 0) return {1.0, 0.0, 0.0};
    double Pm_2, dxPm_2, dyPm_2;
    double Pm_1 = 1.0, dxPm_1 = 0.0, dyPm_1 = 0.0;
    double Pm = sf_jacobi_p(1, a, b, x, y);
    double dxPm = 1.0 + 0.5 * (a + b);
    double dyPm = 0.5 * (static_cast<double>(a) - static_cast<double>(b));
    for (unsigned m = 2; m <= n; ++m) {
        Pm_2 = Pm_1;
        Pm_1 = Pm;
        dxPm_2 = dxPm_1;
        dxPm_1 = dxPm;
        dyPm_2 = dyPm_1;
        dyPm_1 = dyPm;
        auto c = sf_factors(m, a, b);
        Pm = (c[0] * (c[1] * y + c[2] * x) * Pm_1 - c[3] * y * y * Pm_2) / c[4];
        dxPm = (c[0] * (c[2] * Pm_1 + (c[1] * y + c[2] * x) * dxPm_1) - c[3] * y * y * dxPm_2) / c[4];
        dyPm = (c[0] * (c[1] * Pm_1 + (c[1] * y + c[2] * x) * dyPm_1) -
                c[3] * (2.0 * y * Pm_2 + y * y * dyPm_2)) /
               c[4];
    }
    return {Pm, dxPm, dyPm};
}

// Dubiner basis, src/basis/Functions.cpp:88-149 and Functions.h:135-165
double dubiner(unsigned D, unsigned const* i, double const* xi) {
    if (D == 1) {
        return jacobi_p(i[0], 0, 0, 2.0 * xi[0] - 1.0);
    } else if (D == 2) {
        double r_num = 2.0 * xi[0] - 1.0 + xi[1];
        double s = 2.0 * xi[1] - 1.0;
        double theta = 1.0 - xi[1];
        return sf_jacobi_p(i[0], 0, 0, r_num, theta) * sf_jacobi_p(i[1], 2 * i[0] + 1, 0, s, 1.0);
    }
    double r_num = 2.0 * xi[0] - 1.0 + xi[1] + xi[2];
    double s_num = 2.0 * xi[1] - 1.0 + xi[2];
    double t = 2.0 * xi[2] - 1.0;
    double sigmatheta = 1.0 - xi[1] - xi[2];
    double theta = 1.0 - xi[2];
    return sf_jacobi_p(i[0], 0, 0, r_num, sigmatheta) *
           sf_jacobi_p(i[1], 2 * i[0] + 1, 0, s_num, theta) *
           sf_jacobi_p(i[2], 2 * i[0] + 2 * i[1] + 2, 0, t, 1.0);
}

void grad_dubiner(unsigned D, unsigned const* i, double const* xi, double* grad) {
    if (D == 1) {
        grad[0] = jacobi_p_derivative(i[0], 0, 0, 2.0 * xi[0] - 1.0);
        return;
    } else if (D == 2) {
        double r_num = 2.0 * xi[0] - 1.0 + xi[1];
        double s = 2.0 * xi[1] - 1.0;
        double theta = 1.0 - xi[1];
        auto ti = sf_jacobi_p_and_derivatives(i[0], 0, 0, r_num, theta);
        auto tij = sf_jacobi_p_and_derivatives(i[1], 2 * i[0] + 1, 0, s, 1.0);
        auto dd = [&](double dr_num, double dtheta, double dt) {
            return (ti[1] * dr_num + ti[2] * dtheta) * tij[0] + ti[0] * tij[1] * dt;
        };
        grad[0] = dd(2.0, 0.0, 0.0);
        grad[1] = dd(1.0, -1.0, 2.0);
        return;
    }
    double r_num = 2.0 * xi[0] - 1.0 + xi[1] + xi[2];
    double s_num = 2.0 * xi[1] - 1.0 + xi[2];
    double t = 2.0 * xi[2] - 1.0;
    double sigmatheta = 1.0 - xi[1] - xi[2];
    double theta = 1.0 - xi[2];
    auto ti = sf_jacobi_p_and_derivatives(i[0], 0, 0, r_num, sigmatheta);
    auto tij = sf_jacobi_p_and_derivatives(i[1], 2 * i[0] + 1, 0, s_num, theta);
    auto tijk = sf_jacobi_p_and_derivatives(i[2], 2 * i[0] + 2 * i[1] + 2, 0, t, 1.0);
    auto dd = [&](double dr, double dst, double ds, double dth, double dt) {
        return (ti[1] * dr + ti[2] * dst) * tij[0] * tijk[0] +
               ti[0] * (tij[1] * ds + tij[2] * dth) * tijk[0] + ti[0] * tij[0] * (tijk[1] * dt);
    };
    grad[0] = dd(2.0, 0.0, 0.0, 0.0, 0.0);
    grad[1] = dd(1.0, -1.0, 2.0, 0.0, 0.0);
    grad[2] = dd(1.0, -1.0, 1.0, -1.0, 2.0);
}

// src/util/Combinatorics.h:46-110: markers m_0 <= ... <= m_{D-1}; the index is the reversed
// adjacent difference of the markers.
std::vector<unsigned> all_integer_sums(unsigned D, unsigned maxSum) {
    std::vector<unsigned> out;
    std::vector<unsigned> m(D, 0);
    while (true) {
        std::vector<unsigned> idx(D);
        idx[D - 1] = m[0];
        for (unsigned d = 1; d < D; ++d) idx[D - 1 - d] = m[d] - m[d - 1];
        out.insert(out.end(), idx.begin(), idx.end());
        unsigned d = 0;
        while (d < D - 1 && m[d] == m[d + 1]) ++d;
        ++m[d];
        for (unsigned e = 0; e < d; ++e) m[e] = 0;
        if (m[D - 1] > maxSum) break;
    }
    return out;
}

std::size_t binom(unsigned n, unsigned k) {
    if (k > n) return 0;
    std::size_t r = 1;
    for (unsigned i = 1; i <= k; ++i) r = r * (n - k + i) / i;
    return r;
}

static uint64_t range_product(uint64_t from, uint64_t to) {
    uint64_t p = 1;
    for (; from <= to; ++from) p *= from;
    return p;
}

// src/quadrules/GaussJacobi.cpp:15-39
void gauss_jacobi(unsigned n, unsigned a, unsigned b, std::vector<double>& pts,
                  std::vector<double>& wts) {
    pts.assign(n, 0.0);
    wts.assign(n, 0.0);
    const double tol = 10.0 * std::numeric_limits<double>::epsilon();
    double wf = -(2.0 * n + a + b + 2) * range_product(n + 1, n + b) * (1 << (a + b)) /
                ((n + a + b + 1.0) * range_product(n + a + 1, n + a + b) * (n + 1));
    for (unsigned i = 1; i <= n; ++i) {
        double x = std::cos(M_PI * (0.5 * a + i - 0.25) / (0.5 * (1.0 + a + b) + n));
        double Pn = 0.0, dPn = 1.0;
        unsigned it = 0;
        do {
            x -= Pn / dPn;
            Pn = jacobi_p(n, a, b, x);
            dPn = jacobi_p_derivative(n, a, b, x);
        } while (std::fabs(Pn) > tol && ++it < 100);
        pts[i - 1] = x;
        wts[i - 1] = wf / (jacobi_p(n + 1, a, b, x) * dPn);
    }
}

// src/basis/Nodal.cpp:18-28
std::vector<double> lgl_points(unsigned n) {
    std::vector<double> p, w;
    gauss_jacobi(n - 2, 1, 1, p, w);
    std::vector<double> out;
    out.push_back(-1.0);
    for (auto it = p.rbegin(); it != p.rend(); ++it) out.push_back(*it);
    out.push_back(1.0);
    return out;
}

// IntervalQuadratureRule::changeInterval(0, 1) (IntervalQuadratureRule.cpp:5-18)
static void to_unit(std::vector<double>& p, std::vector<double>& w) {
    for (auto& x : p) x = 0.5 * x + 0.5;
    for (auto& x : w) x *= 0.5;
}

// src/quadrules/TensorProductRule.cpp:10-76
Rule tensor_product_rule(unsigned D, unsigned minQuadOrder) {
    unsigned n = 1 + minQuadOrder / 2;
    Rule r;
    r.D = D;
    r.order = 2 * n - 1;
    std::vector<double> p0, w0, p1, w1, p2, w2;
    gauss_jacobi(n, 0, 0, p0, w0);
    to_unit(p0, w0);
    if (D == 1) {
        r.pts = p0;
        r.wts = w0;
        return r;
    }
    gauss_jacobi(n, 1, 0, p1, w1);
    to_unit(p1, w1);
    if (D == 2) {
        for (unsigned i = 0; i < n; ++i)
            for (unsigned j = 0; j < n; ++j) {
                r.pts.push_back(p1[i]);
                r.pts.push_back(p0[j] * (1.0 - p1[i]));
                r.wts.push_back(0.5 * w1[i] * w0[j]);
            }
        return r;
    }
    gauss_jacobi(n, 2, 0, p2, w2);
    to_unit(p2, w2);
    for (unsigned i = 0; i < n; ++i)
        for (unsigned j = 0; j < n; ++j)
            for (unsigned k = 0; k < n; ++k) {
                r.pts.push_back(p2[i]);
                r.pts.push_back(p1[j] * (1 - p2[i]));
                r.pts.push_back(p0[k] * (1 - p1[j]) * (1 - p2[i]));
                r.wts.push_back(0.125 * w2[i] * w1[j] * w0[k]);
            }
    return r;
}

// tabulated rules: family -> dim -> order -> raw (pts, wts)
static std::map<std::string, std::map<unsigned, std::map<unsigned, Rule>>> g_tables;

void load_rule_tables(std::string const& path) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("cannot open rule table " + path);
    g_tables.clear();
    std::string line;
    while (std::getline(f, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream s(line);
        std::string tag, fam;
        unsigned D, order, n;
        s >> tag >> fam >> D >> order >> n;
        if (tag != "rule") throw std::runtime_error("bad rule table line: " + line);
        Rule r;
        r.D = D;
        r.order = order;
        for (unsigned i = 0; i < n; ++i) {
            std::getline(f, line);
            std::istringstream ps(line);
            for (unsigned d = 0; d < D; ++d) {
                std::string tok;
                ps >> tok;
                r.pts.push_back(std::strtod(tok.c_str(), nullptr));
            }
            std::string tok;
            ps >> tok;
            r.wts.push_back(std::strtod(tok.c_str(), nullptr));
        }
        g_tables[fam][D][order] = r;
    }
}

// src/quadrules/TabulatedRules.h:33-104 (lower_bound on order, then the family's transform)
Rule tabulated_rule(std::string const& fam, unsigned D, unsigned minQuadOrder) {
    auto& t = g_tables[fam][D];
    auto it = t.lower_bound(minQuadOrder);
    if (it == t.end()) return Rule{};
    Rule r = it->second;
    if (fam == "JS") {
        for (auto& w : r.wts) w /= 6.0;
    } else {
        for (auto& x : r.pts) x = 0.5 * x + 0.5;
        for (auto& w : r.wts) w /= (1 << D);
    }
    return r;
}

static std::size_t tab_size(std::string const& fam, unsigned D, unsigned o) {
    auto& t = g_tables[fam][D];
    auto it = t.lower_bound(o);
    return it == t.end() ? std::numeric_limits<std::size_t>::max() : it->second.size();
}
static std::size_t tp_size(unsigned D, unsigned o) {
    std::size_t n = 1 + o / 2, s = 1;
    for (unsigned d = 0; d < D; ++d) s *= n;
    return s;
}

// src/quadrules/AutoRule.h:13-51
Rule simplex_rule(unsigned D, unsigned o) {
    if (D == 2) {
        if (tab_size("WV", 2, o) <= tp_size(2, o)) return tabulated_rule("WV", 2, o);
        return tensor_product_rule(2, o);
    } else if (D == 3) {
        bool js = tab_size("JS", 3, o) <= tab_size("WV", 3, o);
        std::size_t best = js ? tab_size("JS", 3, o) : tab_size("WV", 3, o);
        if (best <= tp_size(3, o)) return tabulated_rule(js ? "JS" : "WV", 3, o);
        return tensor_product_rule(3, o);
    }
    return tensor_product_rule(D, o);
}

// ------------------------------------------------------------------------------ dense helpers
// Gaussian elimination with partial pivoting; A column-major n x n, b column-major n x nrhs.
void lu_solve(std::vector<double> A, int n, double* b, int nrhs) {
    std::vector<int> piv(n);
    for (int k = 0; k < n; ++k) {
        int p = k;
        for (int i = k + 1; i < n; ++i)
            if (std::fabs(A[i + k * n]) > std::fabs(A[p + k * n])) p = i;
        piv[k] = p;
        if (p != k) {
            for (int j = 0; j < n; ++j) std::swap(A[k + j * n], A[p + j * n]);
            for (int j = 0; j < nrhs; ++j) std::swap(b[k + j * n], b[p + j * n]);
        }
        double d = A[k + k * n];
        if (d == 0.0) throw std::runtime_error("singular matrix");
        for (int i = k + 1; i < n; ++i) {
            double l = A[i + k * n] / d;
            A[i + k * n] = l;
            for (int j = k + 1; j < n; ++j) A[i + j * n] -= l * A[k + j * n];
            for (int j = 0; j < nrhs; ++j) b[i + j * n] -= l * b[k + j * n];
        }
    }
    for (int j = 0; j < nrhs; ++j) {
        for (int i = n - 1; i >= 0; --i) {
            double s = b[i + j * n];
            for (int k = i + 1; k < n; ++k) s -= A[i + k * n] * b[k + j * n];
            b[i + j * n] = s / A[i + i * n];
        }
    }
}

std::vector<double> mat_inverse(std::vector<double> A, int n) {
    std::vector<double> I(n * n, 0.0);
    for (int i = 0; i < n; ++i) I[i + i * n] = 1.0;
    lu_solve(std::move(A), n, I.data(), n);
    return I;
}

// ------------------------------------------------------------------ warp and blend nodes
// src/basis/WarpAndBlend.cpp:25-43
double wab_alpha(unsigned D, unsigned degree) {
    static const double a2[] = {0.0000,  0.0000,  1.4152,  0.1001,  0.2751,  0.9800,  1.0999,
                                1.2832,  1.3648,  1.4773,  1.4959,  1.5743,  1.5770,  1.6223,
                                1.6258,  1.64831, 1.76303, 1.92727, 1.92962, 1.91251, 1.91289,
                                1.89479, 1.89255, 1.88016, 1.87555, 1.86613, 1.86106, 1.85368,
                                1.84874, 1.84319};
    static const double a3[] = {0.0,      0.0,      0.0,      0.1002,   1.1332,
                                1.5608,   1.3413,   1.2577,   1.1603,   1.10153,
                                0.6080,   0.4523,   0.8856,   0.8717,   0.9655,
                                0.982344, 0.827846, 0.700081, 0.725887};
    if (D == 2) return degree <= 30 ? a2[degree - 1] : 5.0 / 3.0;
    return degree <= 19 ? a3[degree - 1] : 5.0 / 3.0;
}

// Warpfactor (WarpAndBlend.cpp:186-222): warp(r) = sum_i l_i(r) (gll_i - eq_i) / (1 - r^2)
static std::vector<double> warpfactor(unsigned N, std::vector<double> const& r) {
    std::vector<double> eq(N + 1);
    for (unsigned i = 0; i <= N; ++i) eq[i] = 2 * i / static_cast<double>(N) - 1.0;
    std::vector<double> VT((N + 1) * (N + 1));
    for (unsigned i = 0; i <= N; ++i)
        for (unsigned j = 0; j <= N; ++j) VT[j + i * (N + 1)] = jacobi_p(j, 0, 0, eq[i]);
    auto gl = lgl_points(N + 1);
    std::vector<double> phi((N + 1) * r.size());
    for (unsigned i = 0; i <= N; ++i)
        for (std::size_t j = 0; j < r.size(); ++j) phi[i + j * (N + 1)] = jacobi_p(i, 0, 0, r[j]);
    lu_solve(VT, N + 1, phi.data(), r.size());
    std::vector<double> w(r.size(), 0.0);
    for (std::size_t j = 0; j < r.size(); ++j) {
        for (unsigned i = 0; i <= N; ++i) w[j] += phi[i + j * (N + 1)] * (gl[i] - eq[i]);
        if (std::fabs(r[j]) < 1.0 - std::numeric_limits<double>::epsilon()) w[j] /= 1.0 - r[j] * r[j];
    }
    return w;
}

static const double kEqTri[3][2] = {
    {-1.0, -1.0 / std::sqrt(3.0)}, {1.0, -1.0 / std::sqrt(3.0)}, {0.0, 2.0 / std::sqrt(3.0)}};
static const double kEqTet[4][3] = {{-1.0, -1.0 / std::sqrt(3.0), -1.0 / std::sqrt(6.0)},
                                    {1.0, -1.0 / std::sqrt(3.0), -1.0 / std::sqrt(6.0)},
                                    {0.0, 2.0 / std::sqrt(3.0), -1.0 / std::sqrt(6.0)},
                                    {0.0, 0.0, 3.0 / std::sqrt(6.0)}};

// warpAndBlendTriangle (WarpAndBlend.cpp:224-250); L is n x 3 column-major, result n x 2.
static std::vector<double> wab_triangle(unsigned N, double alpha, std::vector<double> const& L,
                                        std::size_t n) {
    std::vector<double> res(2 * n, 0.0);
    static const int edges[3][3] = {{0, 1, 2}, {0, 2, 1}, {1, 2, 0}}; // f0, f1, missing
    for (auto const& e : edges) {
        std::vector<double> r(n);
        for (std::size_t i = 0; i < n; ++i) r[i] = L[i + e[1] * n] - L[i + e[0] * n];
        auto warp = warpfactor(N, r);
        double t[2] = {kEqTri[e[1]][0] - kEqTri[e[0]][0], kEqTri[e[1]][1] - kEqTri[e[0]][1]};
        double tn = std::sqrt(t[0] * t[0] + t[1] * t[1]);
        t[0] /= tn;
        t[1] /= tn;
        for (std::size_t i = 0; i < n; ++i) {
            double blend = 4.0 * L[i + e[0] * n] * L[i + e[1] * n];
            double lm = L[i + e[2] * n];
            blend = blend * (1.0 + (alpha * alpha) * lm * lm);
            for (int d = 0; d < 2; ++d) res[i + d * n] += warp[i] * blend * t[d];
        }
    }
    return res;
}

// x -> reference coordinates of the simplex with vertices v (GeneralPlexToRefPlex, Affine.h:31-55)
static void to_ref(unsigned D, double const* verts, double const* x, double* xi) {
    std::vector<double> A(D * D);
    for (unsigned i = 0; i < D; ++i)
        for (unsigned j = 0; j < D; ++j) A[i + j * D] = verts[(j + 1) * D + i] - verts[i];
    for (unsigned i = 0; i < D; ++i) xi[i] = x[i] - verts[i];
    lu_solve(A, D, xi, 1);
}

// WarpAndBlendFactory<D>::operator() (WarpAndBlend.cpp:45-184)
std::vector<double> wab_nodes(unsigned D, unsigned N) {
    std::size_t n = binom(N + D, D);
    std::vector<double> out(n * D, 0.0);
    if (D == 1) {
        if (N == 0) {
            out[0] = 0.5;
            return out;
        }
        auto gl = lgl_points(N + 1);
        for (std::size_t i = 0; i < n; ++i) out[i] = (gl[i] + 1.0) / 2.0;
        return out;
    }
    if (N == 0) {
        for (unsigned d = 0; d < D; ++d) out[d] = 1.0 / (D + 1);
        return out;
    }
    auto idx = all_integer_sums(D, N);
    double alpha = wab_alpha(D, N);
    std::vector<double> L(n * (D + 1));
    for (std::size_t p = 0; p < n; ++p) {
        double s = 0.0;
        for (unsigned d = 0; d < D; ++d) {
            L[p + (d + 1) * n] = idx[p * D + d] / static_cast<double>(N);
            s += L[p + (d + 1) * n];
        }
        L[p] = 1.0 - s;
    }
    std::vector<double> X(n * D, 0.0); // row-major n x D in equilateral coordinates
    if (D == 2) {
        auto w = wab_triangle(N, alpha, L, n);
        for (std::size_t p = 0; p < n; ++p)
            for (unsigned d = 0; d < 2; ++d) {
                double v = 0.0;
                for (unsigned k = 0; k < 3; ++k) v += L[p + k * n] * kEqTri[k][d];
                X[p * 2 + d] = v + w[p + d * n];
            }
        double verts[6] = {kEqTri[0][0], kEqTri[0][1], kEqTri[1][0],
                           kEqTri[1][1], kEqTri[2][0], kEqTri[2][1]};
        for (std::size_t p = 0; p < n; ++p) to_ref(2, verts, &X[p * 2], &out[p * 2]);
        return out;
    }
    static const int faces[4][4] = {{0, 1, 2, 3}, {0, 1, 3, 2}, {0, 2, 3, 1}, {1, 2, 3, 0}};
    for (auto const& f : faces) {
        double t0[3], t1[3];
        for (int d = 0; d < 3; ++d) {
            t0[d] = kEqTet[f[1]][d] - kEqTet[f[0]][d];
            t1[d] = kEqTet[f[2]][d] - 0.5 * (kEqTet[f[1]][d] + kEqTet[f[0]][d]);
        }
        double n0 = std::sqrt(t0[0] * t0[0] + t0[1] * t0[1] + t0[2] * t0[2]);
        double n1 = std::sqrt(t1[0] * t1[0] + t1[1] * t1[1] + t1[2] * t1[2]);
        for (int d = 0; d < 3; ++d) {
            t0[d] /= n0;
            t1[d] /= n1;
        }
        std::vector<double> Lf(n * 3);
        for (std::size_t p = 0; p < n; ++p)
            for (int k = 0; k < 3; ++k) Lf[p + k * n] = L[p + f[k] * n];
        auto w2 = wab_triangle(N, alpha, Lf, n);
        for (std::size_t p = 0; p < n; ++p) {
            double warp[3];
            for (int d = 0; d < 3; ++d) warp[d] = w2[p] * t0[d] + w2[p + n] * t1[d];
            double lm = L[p + f[3] * n];
            double blend = Lf[p] * Lf[p + n] * Lf[p + 2 * n];
            double denom = (Lf[p] + 0.5 * lm) * (Lf[p + n] + 0.5 * lm) * (Lf[p + 2 * n] + 0.5 * lm);
            double factor = 1.0 + alpha * alpha * lm * lm;
            double mask;
            if (std::fabs(denom) < std::numeric_limits<double>::epsilon()) {
                blend = 1.0;
                mask = 0.0;
            } else {
                blend = factor * blend / denom;
                mask = 1.0;
            }
            for (int d = 0; d < 3; ++d) X[p * 3 + d] = mask * X[p * 3 + d] + blend * warp[d];
        }
    }
    for (std::size_t p = 0; p < n; ++p)
        for (int d = 0; d < 3; ++d) {
            double v = 0.0;
            for (int k = 0; k < 4; ++k) v += L[p + k * n] * kEqTet[k][d];
            X[p * 3 + d] += v;
        }
    double verts[12];
    for (int k = 0; k < 4; ++k)
        for (int d = 0; d < 3; ++d) verts[k * 3 + d] = kEqTet[k][d];
    for (std::size_t p = 0; p < n; ++p) to_ref(3, verts, &X[p * 3], &out[p * 3]);
    return out;
}

// --------------------------------------------------------------------- reference elements
// src/form/RefElement.cpp:86-170 (nodal) and :20-83 (modal)
RefElement::RefElement(unsigned D_, unsigned N_, bool modal_) : D(D_), N(N_), modal(modal_) {
    nbf = binom(N + D, D);
    if (!modal) {
        nodes = wab_nodes(D, N);
        auto idx = all_integer_sums(D, N);
        V.assign(nbf * nbf, 0.0);
        for (unsigned i = 0; i < nbf; ++i)
            for (unsigned bf = 0; bf < nbf; ++bf)
                V[i + bf * nbf] = dubiner(D, &idx[bf * D], &nodes[i * D]);
        Vinv = mat_inverse(V, nbf);
    }
}

std::vector<double> RefElement::eval_basis(std::vector<double> const& pts) const {
    std::size_t np = pts.size() / D;
    auto idx = all_integer_sums(D, N);
    std::vector<double> M(nbf * np);
    for (std::size_t p = 0; p < np; ++p)
        for (unsigned bf = 0; bf < nbf; ++bf) M[bf + p * nbf] = dubiner(D, &idx[bf * D], &pts[p * D]);
    if (modal) return M;
    std::vector<double> E(nbf * np, 0.0);
    for (std::size_t p = 0; p < np; ++p)
        for (unsigned i = 0; i < nbf; ++i) {
            double s = 0.0;
            for (unsigned j = 0; j < nbf; ++j) s += Vinv[j + i * nbf] * M[j + p * nbf];
            E[i + p * nbf] = s;
        }
    return E;
}

std::vector<double> RefElement::eval_grad(std::vector<double> const& pts) const {
    std::size_t np = pts.size() / D;
    auto idx = all_integer_sums(D, N);
    std::vector<double> M(nbf * D * np);
    double g[3];
    for (std::size_t p = 0; p < np; ++p)
        for (unsigned bf = 0; bf < nbf; ++bf) {
            grad_dubiner(D, &idx[bf * D], &pts[p * D], g);
            for (unsigned d = 0; d < D; ++d) M[bf + nbf * (d + D * p)] = g[d];
        }
    if (modal) return M;
    std::vector<double> G(M.size(), 0.0);
    for (std::size_t c = 0; c < D * np; ++c)
        for (unsigned i = 0; i < nbf; ++i) {
            double s = 0.0;
            for (unsigned j = 0; j < nbf; ++j) s += Vinv[j + i * nbf] * M[j + c * nbf];
            G[i + c * nbf] = s;
        }
    return G;
}

// ModalRefElement::massMatrix (RefElement.cpp:20-38): diagonal with rule of order 2N
static std::vector<double> modal_mass(unsigned D, unsigned N) {
    RefElement m(D, N, true);
    auto rule = simplex_rule(D, 2 * N);
    auto E = m.eval_basis(rule.pts);
    std::vector<double> M(m.nbf * m.nbf, 0.0);
    for (unsigned i = 0; i < m.nbf; ++i)
        for (std::size_t q = 0; q < rule.size(); ++q)
            M[i + i * m.nbf] += rule.wts[q] * E[i + q * m.nbf] * E[i + q * m.nbf];
    return M;
}

std::vector<double> RefElement::mass_matrix() const {
    auto Mm = modal_mass(D, N);
    if (modal) return Mm;
    // Vinv^T Mm Vinv
    std::vector<double> M(nbf * nbf, 0.0);
    for (unsigned i = 0; i < nbf; ++i)
        for (unsigned j = 0; j < nbf; ++j) {
            double s = 0.0;
            for (unsigned k = 0; k < nbf; ++k) s += Vinv[k + i * nbf] * Mm[k + k * nbf] * Vinv[k + j * nbf];
            M[i + j * nbf] = s;
        }
    return M;
}

std::vector<double> RefElement::inverse_mass_matrix() const {
    auto Mm = modal_mass(D, N);
    for (unsigned i = 0; i < nbf; ++i) Mm[i + i * nbf] = 1.0 / Mm[i + i * nbf];
    if (modal) return Mm;
    std::vector<double> M(nbf * nbf, 0.0);
    for (unsigned i = 0; i < nbf; ++i)
        for (unsigned j = 0; j < nbf; ++j) {
            double s = 0.0;
            for (unsigned k = 0; k < nbf; ++k) s += V[i + k * nbf] * Mm[k + k * nbf] * V[j + k * nbf];
            M[i + j * nbf] = s;
        }
    return M;
}

} // namespace oracle
