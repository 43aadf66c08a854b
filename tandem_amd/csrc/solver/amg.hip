// Smoothed-aggregation AMG for the p-multigrid coarse level (set-up on the host in FP64, V-cycle on
// the device in FP32).  See amg.h for the method.
#include "amg.h"

#include "../host/operator.h"
#include "../host/refelem.h"
#include "block_jacobi.h"
#include "multi_kernels.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <random>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

namespace tdg {

namespace {

// ------------------------------------------------------------------------------ host matrices
// Block sparse rows: n block rows, blocks br x bc (row-major), columns ci, row starts rp.
struct Bsr {
    int64_t n = 0, ncol = 0;
    int br = 0, bc = 0;
    std::vector<int64_t> rp;
    std::vector<int32_t> ci;
    std::vector<double> v;
    double* blk(int64_t k) { return v.data() + k * br * bc; }
    double const* blk(int64_t k) const { return v.data() + k * br * bc; }
};

// In-place inverse of a dense SPD matrix (row-major, n x n) by Gauss-Jordan without pivoting;
// false if a pivot is not positive (the matrix is then left partly eliminated).
bool dense_inverse(std::vector<double>& A, int n) {
    std::vector<double> col(n);
    for (int k = 0; k < n; ++k) {
        for (int i = 0; i < n; ++i) col[i] = A[std::size_t(i) * n + k];
        if (!(col[k] > 0.0)) return false;
        const double p = 1.0 / col[k];
        for (int j = 0; j < n; ++j) A[std::size_t(k) * n + j] = j == k ? p : A[std::size_t(k) * n + j] * p;
        for (int i = 0; i < n; ++i) {
            if (i == k) continue;
            double* ri = &A[std::size_t(i) * n];
            double const* rk = &A[std::size_t(k) * n];
            const double f = col[i];
            for (int j = 0; j < n; ++j) ri[j] = j == k ? -f * p : ri[j] - f * rk[j];
        }
    }
    for (int i = 0; i < n; ++i) // exactly symmetric
        for (int j = i + 1; j < n; ++j) {
            const double s = 0.5 * (A[std::size_t(i) * n + j] + A[std::size_t(j) * n + i]);
            A[std::size_t(i) * n + j] = s;
            A[std::size_t(j) * n + i] = s;
        }
    return true;
}

// Inverses of the diagonal blocks (n x bs x bs).
std::vector<double> diag_inverse(Bsr const& A) {
    const int bs = A.br;
    std::vector<double> D(std::size_t(A.n) * bs * bs);
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(+ : bad)
    for (int64_t i = 0; i < A.n; ++i) {
        std::vector<double> b(std::size_t(bs) * bs, 0.0);
        for (int64_t k = A.rp[i]; k < A.rp[i + 1]; ++k)
            if (A.ci[k] == i) std::copy(A.blk(k), A.blk(k) + bs * bs, b.begin());
        for (int r = 0; r < bs; ++r)
            for (int c = r + 1; c < bs; ++c) {
                const double s = 0.5 * (b[r * bs + c] + b[c * bs + r]);
                b[r * bs + c] = b[c * bs + r] = s;
            }
        if (!dense_inverse(b, bs)) ++bad;
        std::copy(b.begin(), b.end(), D.begin() + i * bs * bs);
    }
    if (bad) throw std::runtime_error("AMG: a diagonal block is not positive definite");
    return D;
}

// y = A x (square Bsr, vectors n bs)
void spmv(Bsr const& A, std::vector<double> const& x, std::vector<double>& y) {
    const int bs = A.br;
    y.assign(std::size_t(A.n) * bs, 0.0);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < A.n; ++i)
        for (int64_t k = A.rp[i]; k < A.rp[i + 1]; ++k) {
            double const* b = A.blk(k);
            double const* xv = &x[std::size_t(A.ci[k]) * bs];
            for (int r = 0; r < bs; ++r) {
                double s = 0.0;
                for (int c = 0; c < bs; ++c) s += b[r * bs + c] * xv[c];
                y[std::size_t(i) * bs + r] += s;
            }
        }
}

// lambda_max(D^-1 A) by power iteration v <- D^-1 A v / |.| (the eigenvalues of D^-1 A are real and
// positive for SPD A and D); 30 steps from a seeded random vector, raised by 5% for the Chebyshev
// interval's upper end since the iterates approach it from below
double power_rho(Bsr const& A, std::vector<double> const& Dinv) {
    const int bs = A.br;
    const std::size_t N = std::size_t(A.n) * bs;
    std::vector<double> v(N), w, z(N);
    std::mt19937_64 g(7);
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    for (auto& e : v) e = u(g);
    double lam = 0.0;
    for (int it = 0; it < 30; ++it) {
        double nv = 0.0;
        for (double e : v) nv += e * e;
        nv = std::sqrt(nv);
        for (auto& e : v) e /= nv;
        spmv(A, v, w);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < A.n; ++i)
            for (int r = 0; r < bs; ++r) {
                double s = 0.0;
                for (int c = 0; c < bs; ++c) s += Dinv[(std::size_t(i) * bs + r) * bs + c] * w[std::size_t(i) * bs + c];
                z[std::size_t(i) * bs + r] = s;
            }
        double num = 0.0;
        for (std::size_t i = 0; i < N; ++i) num += v[i] * z[i];
        lam = num;
        v.swap(z);
    }
    return lam * 1.05;
}

// Aggregates: a node whose neighbours are all free starts an aggregate with them; free nodes then
// join a neighbouring aggregate; the rest group with their free neighbours.
std::vector<int32_t> aggregate(Bsr const& A, int64_t* nagg) {
    std::vector<int32_t> agg(A.n, -1);
    int64_t na = 0;
    for (int64_t i = 0; i < A.n; ++i) {
        if (agg[i] >= 0) continue;
        bool free = true;
        for (int64_t k = A.rp[i]; k < A.rp[i + 1] && free; ++k) free = agg[A.ci[k]] < 0;
        if (!free) continue;
        for (int64_t k = A.rp[i]; k < A.rp[i + 1]; ++k) agg[A.ci[k]] = int32_t(na);
        agg[i] = int32_t(na);
        ++na;
    }
    std::vector<int32_t> join(A.n, -1);
    for (int64_t i = 0; i < A.n; ++i) {
        if (agg[i] >= 0) continue;
        for (int64_t k = A.rp[i]; k < A.rp[i + 1]; ++k)
            if (agg[A.ci[k]] >= 0) {
                join[i] = agg[A.ci[k]];
                break;
            }
    }
    for (int64_t i = 0; i < A.n; ++i)
        if (agg[i] < 0 && join[i] >= 0) agg[i] = join[i];
    for (int64_t i = 0; i < A.n; ++i) {
        if (agg[i] >= 0) continue;
        agg[i] = int32_t(na);
        for (int64_t k = A.rp[i]; k < A.rp[i + 1]; ++k)
            if (agg[A.ci[k]] < 0) agg[A.ci[k]] = int32_t(na);
        ++na;
    }
    *nagg = na;
    return agg;
}

// Tentative prolongator: per aggregate the stacked near-null blocks B_i (bs x m) orthonormalised
// (modified Gram-Schmidt, twice); P_tent(i, agg(i)) = Q rows of node i, next near-null space = R.
void tentative(Bsr const& A, std::vector<int32_t> const& agg, int64_t na, std::vector<double> const& B, int m,
               Bsr& P, std::vector<double>& Bc) {
    const int bs = A.br;
    std::vector<std::vector<int64_t>> members(na);
    for (int64_t i = 0; i < A.n; ++i) members[agg[i]].push_back(i);
    P.n = A.n;
    P.ncol = na;
    P.br = bs;
    P.bc = m;
    P.rp.resize(A.n + 1);
    P.ci.resize(A.n);
    P.v.assign(std::size_t(A.n) * bs * m, 0.0);
    for (int64_t i = 0; i <= A.n; ++i) P.rp[i] = i;
    for (int64_t i = 0; i < A.n; ++i) P.ci[i] = agg[i];
    Bc.assign(std::size_t(na) * m * m, 0.0);
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : bad)
    for (int64_t a = 0; a < na; ++a) {
        auto const& mem = members[a];
        const std::size_t rows = mem.size() * bs;
        std::vector<double> Q(rows * m); // column-major rows x m
        for (std::size_t t = 0; t < mem.size(); ++t)
            for (int r = 0; r < bs; ++r)
                for (int c = 0; c < m; ++c) Q[c * rows + t * bs + r] = B[(std::size_t(mem[t]) * bs + r) * m + c];
        std::vector<double> R(std::size_t(m) * m, 0.0);
        for (int c = 0; c < m; ++c) {
            double* qc = &Q[c * rows];
            double n0 = 0.0;
            for (std::size_t r = 0; r < rows; ++r) n0 += qc[r] * qc[r];
            for (int pass = 0; pass < 2; ++pass)
                for (int p = 0; p < c; ++p) {
                    double const* qp = &Q[p * rows];
                    double d = 0.0;
                    for (std::size_t r = 0; r < rows; ++r) d += qp[r] * qc[r];
                    for (std::size_t r = 0; r < rows; ++r) qc[r] -= d * qp[r];
                    R[p * m + c] += d;
                }
            double nn = 0.0;
            for (std::size_t r = 0; r < rows; ++r) nn += qc[r] * qc[r];
            if (!(nn > 1e-24 * n0) || !(nn > 0.0)) {
                ++bad;
                nn = 1.0;
            }
            nn = std::sqrt(nn);
            for (std::size_t r = 0; r < rows; ++r) qc[r] /= nn;
            R[c * m + c] = nn;
        }
        for (std::size_t t = 0; t < mem.size(); ++t)
            for (int r = 0; r < bs; ++r)
                for (int c = 0; c < m; ++c) P.v[(std::size_t(mem[t]) * bs + r) * m + c] = Q[c * rows + t * bs + r];
        std::copy(R.begin(), R.end(), Bc.begin() + a * m * m);
    }
    if (bad) throw std::runtime_error("AMG: an aggregate does not carry the near-null space");
}

// C = A B for block-compatible Bsr (A: n x k nodes, B: k x l nodes), row by row with a map
Bsr multiply(Bsr const& A, Bsr const& B) {
    Bsr C;
    C.n = A.n;
    C.ncol = B.ncol;
    C.br = A.br;
    C.bc = B.bc;
    C.rp.assign(A.n + 1, 0);
    const int br = A.br, bk = A.bc, bc = B.bc;
    std::vector<std::vector<int32_t>> cols(A.n);
    std::vector<std::vector<double>> vals(A.n);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < A.n; ++i) {
        std::unordered_map<int32_t, int> pos;
        auto& ci = cols[i];
        auto& cv = vals[i];
        for (int64_t ka = A.rp[i]; ka < A.rp[i + 1]; ++ka) {
            double const* a = A.blk(ka);
            const int64_t j = A.ci[ka];
            for (int64_t kb = B.rp[j]; kb < B.rp[j + 1]; ++kb) {
                double const* b = B.blk(kb);
                auto it = pos.find(B.ci[kb]);
                int p;
                if (it == pos.end()) {
                    p = int(ci.size());
                    pos.emplace(B.ci[kb], p);
                    ci.push_back(B.ci[kb]);
                    cv.resize(cv.size() + std::size_t(br) * bc, 0.0);
                } else {
                    p = it->second;
                }
                double* c = &cv[std::size_t(p) * br * bc];
                for (int r = 0; r < br; ++r)
                    for (int t = 0; t < bk; ++t) {
                        const double ar = a[r * bk + t];
                        for (int s = 0; s < bc; ++s) c[r * bc + s] += ar * b[t * bc + s];
                    }
            }
        }
    }
    for (int64_t i = 0; i < A.n; ++i) C.rp[i + 1] = C.rp[i] + int64_t(cols[i].size());
    C.ci.resize(C.rp[A.n]);
    C.v.resize(std::size_t(C.rp[A.n]) * br * bc);
    for (int64_t i = 0; i < A.n; ++i) {
        std::copy(cols[i].begin(), cols[i].end(), C.ci.begin() + C.rp[i]);
        std::copy(vals[i].begin(), vals[i].end(), C.v.begin() + C.rp[i] * br * bc);
    }
    return C;
}

Bsr transpose(Bsr const& P) {
    Bsr T;
    T.n = P.ncol;
    T.ncol = P.n;
    T.br = P.bc;
    T.bc = P.br;
    T.rp.assign(T.n + 1, 0);
    for (int32_t c : P.ci) ++T.rp[c + 1];
    for (int64_t i = 0; i < T.n; ++i) T.rp[i + 1] += T.rp[i];
    T.ci.resize(P.ci.size());
    T.v.resize(P.v.size());
    std::vector<int64_t> fill(T.rp.begin(), T.rp.end() - 1);
    for (int64_t i = 0; i < P.n; ++i)
        for (int64_t k = P.rp[i]; k < P.rp[i + 1]; ++k) {
            const int64_t o = fill[P.ci[k]]++;
            T.ci[o] = int32_t(i);
            double const* s = P.blk(k);
            double* d = T.blk(o);
            for (int r = 0; r < P.br; ++r)
                for (int c = 0; c < P.bc; ++c) d[c * P.br + r] = s[r * P.bc + c];
        }
    return T;
}

// P = P_tent - omega D^-1 (A P_tent)
Bsr smooth_prolongator(Bsr const& A, std::vector<double> const& Dinv, double omega, Bsr const& Pt) {
    Bsr AP = multiply(A, Pt);
    const int bs = A.br, m = Pt.bc;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < AP.n; ++i) {
        double const* Di = &Dinv[std::size_t(i) * bs * bs];
        std::vector<double> t(std::size_t(bs) * m);
        for (int64_t k = AP.rp[i]; k < AP.rp[i + 1]; ++k) {
            double* b = AP.blk(k);
            for (int r = 0; r < bs; ++r)
                for (int c = 0; c < m; ++c) {
                    double s = 0.0;
                    for (int q = 0; q < bs; ++q) s += Di[r * bs + q] * b[q * m + c];
                    t[r * m + c] = -omega * s;
                }
            if (AP.ci[k] == Pt.ci[i]) // the tentative block of row i
                for (int q = 0; q < bs * m; ++q) t[q] += Pt.v[std::size_t(i) * bs * m + q];
            std::copy(t.begin(), t.end(), b);
        }
    }
    return AP;
}

// ------------------------------------------------------------------------------ device kernels
constexpr int kT = 256;

// One lane per scalar row (block row i, row r of its blocks) and column: BR lanes share a block row
// (the same x blocks, consecutive block rows), so a level of a few thousand block rows still fills
// the chip for a single column.
template <class T, int BR, int BC, int MODE> // MODE 0: y = A x; 1: y = b - A x; 2: y += A x
__global__ __launch_bounds__(kT) void bsr_kernel(int64_t n, int64_t const* __restrict__ rp,
                                                 int32_t const* __restrict__ ci, T const* __restrict__ v,
                                                 T const* __restrict__ xs, int64_t ldx, T* __restrict__ ys,
                                                 int64_t ldy, T const* __restrict__ bs, int64_t ldb) {
    const int64_t t = int64_t(blockIdx.x) * kT + threadIdx.x;
    if (t >= n * BR) return;
    const int64_t i = t / BR;
    const int r = int(t - i * BR);
    const int64_t c = blockIdx.y;
    T const* x = xs + c * ldx;
    T acc = T(0);
    for (int64_t k = rp[i]; k < rp[i + 1]; ++k) {
        T const* b = v + k * (BR * BC) + r * BC;
        T const* xv = x + int64_t(ci[k]) * BC;
#pragma unroll
        for (int s = 0; s < BC; ++s) acc += b[s] * xv[s];
    }
    T* y = ys + c * ldy + t;
    if constexpr (MODE == 0) {
        *y = acc;
    } else if constexpr (MODE == 1) {
        *y = bs[c * ldb + t] - acc;
    } else {
        *y += acc;
    }
}

// x = Ainv b (dense N x N, row-major), ncols columns
template <class T>
__global__ __launch_bounds__(kT) void dense_kernel(int64_t N, T const* __restrict__ Ainv, T const* __restrict__ bs,
                                                   int64_t ldb, T* __restrict__ xs, int64_t ldx) {
    const int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x;
    if (i >= N) return;
    T const* b = bs + int64_t(blockIdx.y) * ldb;
    T const* a = Ainv + i * N;
    T s = T(0);
    for (int64_t j = 0; j < N; ++j) s += a[j] * b[j];
    xs[int64_t(blockIdx.y) * ldx + i] = s;
}

// d = z / theta; x = d (zero guess) or x += d
template <class T>
__global__ __launch_bounds__(kT) void cheb_first_kernel(int64_t N, T* __restrict__ x, int64_t ldx, T* __restrict__ d,
                                                        int64_t ldd, T const* __restrict__ z, int64_t ldz, T inv_theta,
                                                        int zero) {
    const int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x;
    if (i >= N) return;
    const int64_t c = blockIdx.y;
    const T dv = z[c * ldz + i] * inv_theta;
    d[c * ldd + i] = dv;
    x[c * ldx + i] = zero ? dv : x[c * ldx + i] + dv;
}

// d = c1 d + c2 z; x += d
template <class T>
__global__ __launch_bounds__(kT) void cheb_next_kernel(int64_t N, T* __restrict__ x, int64_t ldx, T* __restrict__ d,
                                                       int64_t ldd, T const* __restrict__ z, int64_t ldz, T c1, T c2) {
    const int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x;
    if (i >= N) return;
    const int64_t c = blockIdx.y;
    const T dv = c1 * d[c * ldd + i] + c2 * z[c * ldz + i];
    d[c * ldd + i] = dv;
    x[c * ldx + i] += dv;
}

// r = b - t
template <class T>
__global__ __launch_bounds__(kT) void sub_kernel(int64_t N, T* __restrict__ r, int64_t ldr, T const* __restrict__ b,
                                                 int64_t ldb, T const* __restrict__ t, int64_t ldt) {
    const int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x;
    if (i >= N) return;
    const int64_t c = blockIdx.y;
    r[c * ldr + i] = b[c * ldb + i] - t[c * ldt + i];
}

unsigned blocks_for(int64_t n) { return unsigned((n + kT - 1) / kT); }

template <class T> struct Dev {
    T* p = nullptr;
    std::size_t n = 0;
    Dev() = default;
    Dev(Dev const&) = delete;
    Dev& operator=(Dev const&) = delete;
    ~Dev() {
        if (p) (void)hipFree(p);
    }
    void upload(std::vector<T> const& h) {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = h.size();
        if (n == 0) return;
        MHIP(hipMalloc(reinterpret_cast<void**>(&p), n * sizeof(T)));
        MHIP(hipMemcpy(p, h.data(), n * sizeof(T), hipMemcpyHostToDevice));
    }
    void ensure(std::size_t m) {
        if (n >= m) return;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        MHIP(hipMalloc(reinterpret_cast<void**>(&p), m * sizeof(T)));
        n = m;
    }
};

// A Bsr on the device in both precisions (structure shared).
struct DevBsr {
    int64_t n = 0;
    int br = 0, bc = 0;
    Dev<int64_t> rp;
    Dev<int32_t> ci;
    Dev<float> v32;
    Dev<double> v64;
    void upload(Bsr const& h) {
        n = h.n;
        br = h.br;
        bc = h.bc;
        rp.upload(h.rp);
        ci.upload(h.ci);
        std::vector<float> f(h.v.begin(), h.v.end());
        v32.upload(f);
        v64.upload(h.v);
    }
    template <class T> T const* values() const {
        if constexpr (sizeof(T) == 4) return v32.p;
        else return v64.p;
    }
};

template <int MODE, class T>
void bsr_launch(DevBsr const& A, T const* x, int64_t ldx, T* y, int64_t ldy, T const* b, int64_t ldb, int ncols,
                hipStream_t s) {
    const dim3 g(blocks_for(A.n * A.br), unsigned(ncols)), t(kT);
    T const* v = A.values<T>();
#define TDG_BSR(R, C)                                                                                        \
    if (A.br == R && A.bc == C) {                                                                            \
        hipLaunchKernelGGL((bsr_kernel<T, R, C, MODE>), g, t, 0, s, A.n, A.rp.p, A.ci.p, v, x, ldx, y, ldy, b,   \
                           ldb);                                                                             \
        MHIP(hipGetLastError());                                                                             \
        return;                                                                                              \
    }
    TDG_BSR(3, 3)
    TDG_BSR(6, 3)
    TDG_BSR(3, 6)
    TDG_BSR(12, 3)
    TDG_BSR(3, 12)
    TDG_BSR(6, 6)
    TDG_BSR(12, 6)
    TDG_BSR(6, 12)
#undef TDG_BSR
    throw std::logic_error("AMG: no kernel for this block shape");
}

// Work vectors of one level in one precision.
template <class T> struct LevelVectors {
    Dev<T> x, b, r, d, z, t;
};

} // namespace

// Device data of one level: the operator (levels >= 1), block-diagonal inverse as a diagonal Bsr,
// the prolongator from the next level and its transpose (both precisions), and work vectors.
struct AMGLevelDev {
    int64_t n = 0; // block rows
    int bs = 0;
    DevBsr A, Dinv, P, R;
    Dev<float> dense32; // coarsest level: dense inverse
    Dev<double> dense64;
    LevelVectors<float> w32;
    LevelVectors<double> w64;
    int64_t ld = 0; // stride of this level's work vectors
    template <class T> LevelVectors<T>& w() {
        if constexpr (sizeof(T) == 4) return w32;
        else return w64;
    }
    template <class T> T const* dense() const {
        if constexpr (sizeof(T) == 4) return dense32.p;
        else return dense64.p;
    }
};

AMG::AMG(Operator& op, AMGOptions const& o) : op_(op), o_(o) {
    if (op.kind() != Kind_Elasticity) throw std::invalid_argument("AMG: elasticity operators only");
    if (!op.assembled_fast()) throw std::invalid_argument("AMG: the operator's block size has no assembled kernel");
    const unsigned D = op.dim(), Nbf = op.nbf();
    const int bs = int(op.block_size());
    const int m = D == 2 ? 3 : 6;
    const int64_t E = int64_t(op.num_elements());
    if (!((D == 2 && (bs == 6 || bs == 12)) || (D == 3 && bs == 12)))
        throw std::invalid_argument("AMG: elasticity P1 / P2 (2D) and P1 (3D) coarse levels only");
    op.assemble();
    std::vector<double> Ah;
    std::vector<int32_t> nb;
    op.assembled_host(Ah, nb);
    // level 0 Bsr from the block rows: own block, then the face neighbours
    Bsr A;
    A.n = A.ncol = E;
    A.br = A.bc = bs;
    A.rp.assign(E + 1, 0);
    const unsigned NF = D + 1;
    for (int64_t e = 0; e < E; ++e) {
        int cnt = 1;
        for (unsigned f = 0; f < NF; ++f)
            if (nb[std::size_t(f) * E + e] != e) ++cnt;
        A.rp[e + 1] = A.rp[e] + cnt;
    }
    A.ci.resize(A.rp[E]);
    A.v.resize(std::size_t(A.rp[E]) * bs * bs);
    for (int64_t e = 0; e < E; ++e) {
        int64_t k = A.rp[e];
        for (unsigned s = 0; s <= NF; ++s) {
            int64_t col = e;
            if (s > 0) {
                col = nb[std::size_t(s - 1) * E + e];
                if (col == e) continue;
            }
            A.ci[k] = int32_t(col);
            double* b = A.blk(k);
            for (int i = 0; i < bs; ++i)
                for (int j = 0; j < bs; ++j) b[i * bs + j] = Ah[((std::size_t(s) * bs + i) * bs + j) * E + e];
            ++k;
        }
    }
    Ah.clear();
    Ah.shrink_to_fit();
    // near-null space: rigid-body modes at the element's nodes (affine image of the reference nodes)
    SimplexMesh const& mesh = op.mesh();
    std::vector<double> rn = warp_blend_nodes(D, op.degree());
    std::vector<double> B(std::size_t(E) * bs * m, 0.0);
    for (int64_t e = 0; e < E; ++e) {
        int64_t const* el = &mesh.elements[std::size_t(e) * NF];
        double cen[3] = {0, 0, 0};
        for (unsigned v = 0; v < NF; ++v)
            for (unsigned i = 0; i < D; ++i) cen[i] += mesh.vertices[el[v] * D + i] / NF;
        for (unsigned k = 0; k < Nbf; ++k) {
            double s = 0.0;
            for (unsigned d = 0; d < D; ++d) s += rn[k * D + d];
            double X[3] = {0, 0, 0};
            for (unsigned i = 0; i < D; ++i) {
                double v = (1.0 - s) * mesh.vertices[el[0] * D + i];
                for (unsigned d = 0; d < D; ++d) v += rn[k * D + d] * mesh.vertices[el[d + 1] * D + i];
                X[i] = v - cen[i];
            }
            auto row = [&](unsigned p) -> double* { return &B[(std::size_t(e) * bs + k + p * Nbf) * m]; };
            for (unsigned p = 0; p < D; ++p) row(p)[p] = 1.0; // translations
            if (D == 2) {
                row(0)[2] = -X[1];
                row(1)[2] = X[0];
            } else {
                row(0)[3] = -X[1]; // rotation about z
                row(1)[3] = X[0];
                row(1)[4] = -X[2]; // about x
                row(2)[4] = X[1];
                row(0)[5] = X[2]; // about y
                row(2)[5] = -X[0];
            }
        }
    }
    hipStream_t st = op.stream();
    const double nnz0 = double(A.v.size());
    double nnz = nnz0;
    int64_t dof = E * bs;
    for (int l = 0;; ++l) {
        auto lv = std::make_unique<AMGLevelDev>();
        lv->n = A.n;
        lv->bs = A.br;
        dofs_.push_back(A.n * A.br);
        std::vector<double> Dinv = diag_inverse(A);
        const double rho = power_rho(A, Dinv);
        rho_.push_back(rho);
        if (l > 0) {
            lv->A.upload(A);
            Bsr Dm;
            Dm.n = Dm.ncol = A.n;
            Dm.br = Dm.bc = A.br;
            Dm.rp.resize(A.n + 1);
            Dm.ci.resize(A.n);
            for (int64_t i = 0; i <= A.n; ++i) Dm.rp[i] = i;
            for (int64_t i = 0; i < A.n; ++i) Dm.ci[i] = int32_t(i);
            Dm.v = Dinv;
            lv->Dinv.upload(Dm);
        }
        const bool last = dof <= o_.coarse_dofs || l + 1 >= o_.max_levels;
        if (last) {
            // dense inverse of the coarsest operator (the operator itself if it is that small)
            const int N = int(A.n * A.br);
            std::vector<double> Dn(std::size_t(N) * N, 0.0);
            for (int64_t i = 0; i < A.n; ++i)
                for (int64_t k = A.rp[i]; k < A.rp[i + 1]; ++k)
                    for (int r = 0; r < A.br; ++r)
                        for (int c = 0; c < A.bc; ++c)
                            Dn[(std::size_t(i) * A.br + r) * N + std::size_t(A.ci[k]) * A.bc + c] =
                                A.blk(k)[r * A.bc + c];
            if (!dense_inverse(Dn, N)) throw std::runtime_error("AMG: coarse matrix is not positive definite");
            std::vector<float> f(Dn.begin(), Dn.end());
            lv->dense32.upload(f);
            lv->dense64.upload(Dn);
            lv_.push_back(std::move(lv));
            break;
        }
        int64_t na = 0;
        std::vector<int32_t> agg = aggregate(A, &na);
        Bsr Pt;
        std::vector<double> Bc;
        tentative(A, agg, na, B, m, Pt, Bc);
        Bsr P = smooth_prolongator(A, Dinv, 4.0 / (3.0 * rho), Pt);
        if (na >= A.n) throw std::runtime_error("AMG: aggregation does not coarsen");
        Bsr R = transpose(P);
        Bsr Ac = multiply(R, multiply(A, P));
        lv->P.upload(P);
        lv->R.upload(R);
        lv_.push_back(std::move(lv));
        A = std::move(Ac);
        B = std::move(Bc);
        dof = A.n * A.br;
        nnz += double(A.v.size());
    }
    complexity_ = nnz / nnz0;
    MHIP(hipStreamSynchronize(st));
}

AMG::~AMG() = default;

// r = b - A_l x
template <class T>
void AMG::residual(int l, T const* b, int64_t ldb, T const* x, int64_t ldx, T* r, int64_t ldr, int ncols) {
    AMGLevelDev& v = *lv_[l];
    hipStream_t st = op_.stream();
    if (l == 0) {
        T* t = v.w<T>().t.p;
        op_.apply_batched(x, t, ncols, v.ld);
        const int64_t N = v.n * v.bs;
        hipLaunchKernelGGL((sub_kernel<T>), dim3(blocks_for(N), unsigned(ncols)), dim3(kT), 0, st, N, r, ldr, b, ldb,
                           static_cast<T const*>(t), v.ld);
        MHIP(hipGetLastError());
    } else {
        bsr_launch<1>(v.A, x, ldx, r, ldr, b, ldb, ncols, st);
    }
}

// Chebyshev steps on D^-1 A from x (or from zero)
template <class T>
void AMG::smooth(int l, T const* b, int64_t ldb, T* x, int64_t ldx, int ncols, bool zero_guess) {
    AMGLevelDev& v = *lv_[l];
    LevelVectors<T>& w = v.w<T>();
    hipStream_t st = op_.stream();
    const int64_t N = v.n * v.bs;
    const dim3 g(blocks_for(N), unsigned(ncols)), t(kT);
    const double lo = o_.eig_lo * rho_[l], hi = o_.eig_hi * rho_[l];
    const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo), sigma = theta / delta;
    double rho0 = 1.0 / sigma;
    auto precond = [&](T const* r, T* z) {
        if (l == 0) op_.block_jacobi().apply(r, z, st, ncols, v.ld);
        else bsr_launch<0>(v.Dinv, r, v.ld, z, v.ld, static_cast<T const*>(nullptr), 0, ncols, st);
    };
    if (zero_guess) {
        precond(b, w.z.p);
    } else {
        residual<T>(l, b, ldb, x, ldx, w.r.p, v.ld, ncols);
        precond(w.r.p, w.z.p);
    }
    hipLaunchKernelGGL((cheb_first_kernel<T>), g, t, 0, st, N, x, ldx, w.d.p, v.ld, static_cast<T const*>(w.z.p),
                       v.ld, T(1.0 / theta), zero_guess ? 1 : 0);
    MHIP(hipGetLastError());
    for (int k = 1; k < o_.smooth; ++k) {
        residual<T>(l, b, ldb, x, ldx, w.r.p, v.ld, ncols);
        precond(w.r.p, w.z.p);
        const double rho1 = 1.0 / (2.0 * sigma - rho0);
        hipLaunchKernelGGL((cheb_next_kernel<T>), g, t, 0, st, N, x, ldx, w.d.p, v.ld, static_cast<T const*>(w.z.p),
                           v.ld, T(rho1 * rho0), T(2.0 * rho1 / delta));
        MHIP(hipGetLastError());
        rho0 = rho1;
    }
}

template <class T> void AMG::cycle(int l, T const* b, int64_t ldb, T* x, int64_t ldx, int ncols) {
    AMGLevelDev& v = *lv_[l];
    hipStream_t st = op_.stream();
    if (l + 1 == int(lv_.size())) {
        const int64_t N = v.n * v.bs;
        hipLaunchKernelGGL((dense_kernel<T>), dim3(blocks_for(N), unsigned(ncols)), dim3(kT), 0, st, N, v.dense<T>(), b,
                           ldb, x, ldx);
        MHIP(hipGetLastError());
        return;
    }
    AMGLevelDev& c = *lv_[l + 1];
    LevelVectors<T>& w = v.w<T>();
    LevelVectors<T>& wc = c.w<T>();
    smooth<T>(l, b, ldb, x, ldx, ncols, true);
    residual<T>(l, b, ldb, x, ldx, w.r.p, v.ld, ncols);
    bsr_launch<0>(v.R, static_cast<T const*>(w.r.p), v.ld, wc.b.p, c.ld, static_cast<T const*>(nullptr), 0, ncols, st);
    cycle<T>(l + 1, wc.b.p, c.ld, wc.x.p, c.ld, ncols);
    bsr_launch<2>(v.P, static_cast<T const*>(wc.x.p), c.ld, x, ldx, static_cast<T const*>(nullptr), 0, ncols, st);
    smooth<T>(l, b, ldb, x, ldx, ncols, false);
}

template <class T> void AMG::run(T const* r, T* z, int ncols, int64_t ld) {
    if (ncols < 1 || ld < dofs_[0] || (ncols > 65535)) throw std::invalid_argument("AMG: bad column layout");
    const int p = sizeof(T) == 4 ? 0 : 1;
    if (ncols > cap_cols_[p] || ld != cap_ld_[p]) {
        for (std::size_t l = 0; l < lv_.size(); ++l) {
            AMGLevelDev& v = *lv_[l];
            v.ld = l == 0 ? ld : v.n * v.bs;
            LevelVectors<T>& w = v.w<T>();
            const std::size_t tot = std::size_t(ncols) * std::size_t(v.ld);
            if (l > 0) {
                w.x.ensure(tot);
                w.b.ensure(tot);
            }
            if (l + 1 < lv_.size()) {
                w.r.ensure(tot);
                w.d.ensure(tot);
                w.z.ensure(tot);
                if (l == 0) w.t.ensure(tot);
            }
        }
        cap_cols_[p] = ncols;
        cap_ld_[p] = ld;
    }
    cycle<T>(0, r, ld, z, ld, ncols);
}

void AMG::vcycle(float const* r, float* z, int ncols, int64_t ld) { run(r, z, ncols, ld); }
void AMG::vcycle(double const* r, double* z, int ncols, int64_t ld) { run(r, z, ncols, ld); }

} // namespace tdg
