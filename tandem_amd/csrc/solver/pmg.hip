// p-multigrid V-cycle over DG operators of decreasing degree on one mesh, and the flexible CG it
// preconditions, for ncols right-hand sides at once (every operator application is one batched
// apply over all columns).
//
// Levels (tandem: PetscLinearSolver::setup_mg, app/common/PetscLinearSolver.cpp:53-114, levels from
// MGConfig::levels, app/common/MGConfig.cpp:11-40):
//   prolongation  P_l = the nodal interpolation from degree p_(l+1) to p_l per element and quantity
//                 (NodalInterpolation::assemble, app/localoperator/NodalInterpolation.cpp:8-38),
//   restriction   P_l^T,
//   coarse matrix the coarse operator itself, matrix-free.  tandem forms A_(l+1) = P^T A_l P; on
//                 nested spaces that is the fine bilinear form restricted to the coarse space, which
//                 the coarse operator reproduces with the fine facet penalty (galerkin_penalty) up to
//                 the quadrature of the material.
// Smoother (examples/options/mg_cheby.cfg: -mg_levels_ksp_type chebyshev, -mg_levels_pc_type
// jacobi): Chebyshev iteration on D^-1 A with D the element-block diagonal, on the interval
// [eig_lo, eig_hi] x lambda_max, lambda_max estimated by a few preconditioned CG (Lanczos) steps.
// Coarsest level: block-Jacobi CG to coarse_rtol (-mg_coarse_ksp_type cg, -mg_coarse_ksp_rtol).
// Because that coarse solve is not a fixed linear map, the outer Krylov method is flexible CG with
// one truncation direction (tandem's configuration uses FGMRES): beta = -(z . Ap) / (p . Ap),
// alpha = (p . r) / (p . Ap).
#include "pmg.h"

#include "../host/operator.h"
#include "amg.h"
#include "block_jacobi.h"
#include "multi_kernels.h"
#include "progress.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

namespace tdg {

struct PMGLevel {
    int64_t n = 0;              // local size (entries per column)
    int nbf = 0;                // basis functions per quantity
    DBuf<double> b, x;          // right-hand side and correction (levels below the finest)
    DBuf<double> r, d, z, t;    // smoother: residual, Chebyshev direction, D^-1 r, A d
    DBuf<double> I;             // interpolation from this level to the next finer (nbf_fine x nbf)
};

namespace {

constexpr int kT = 256;

unsigned grid_of(int64_t n) { return unsigned(std::min<int64_t>((n + kT - 1) / kT, 16384)); }

// yf(c, e, p, k) += sum_l I(k, l) xc(c, e, p, l): coarse-to-fine prolongation added to yf
template <class T>
__global__ __launch_bounds__(kT) void prolong_add(double const* __restrict__ I, int nf, int nc, int nq, int64_t E,
                                                  int ncols, T const* __restrict__ xc, int64_t ldc,
                                                  T* __restrict__ yf, int64_t ldf) {
    extern __shared__ double sI[];
    for (int i = threadIdx.x; i < nf * nc; i += kT) sI[i] = I[i];
    __syncthreads();
    const int64_t per = E * nq * nf, total = per * ncols;
    for (int64_t t = int64_t(blockIdx.x) * kT + threadIdx.x; t < total; t += int64_t(gridDim.x) * kT) {
        const int64_t c = t / per, r = t - c * per;
        const int64_t blk = r / nf; // (e, p)
        const int k = int(r - blk * nf);
        T const* xe = xc + c * ldc + blk * nc;
        double v = 0.0;
        for (int l = 0; l < nc; ++l) v += sI[k * nc + l] * double(xe[l]);
        yf[c * ldf + r] += T(v);
    }
}

// yc(c, e, p, l) = sum_k I(k, l) xf(c, e, p, k): fine-to-coarse restriction (transpose)
template <class T>
__global__ __launch_bounds__(kT) void restrict_t(double const* __restrict__ I, int nf, int nc, int nq, int64_t E,
                                                 int ncols, T const* __restrict__ xf, int64_t ldf,
                                                 T* __restrict__ yc, int64_t ldc) {
    extern __shared__ double sI[];
    for (int i = threadIdx.x; i < nf * nc; i += kT) sI[i] = I[i];
    __syncthreads();
    const int64_t per = E * nq * nc, total = per * ncols;
    for (int64_t t = int64_t(blockIdx.x) * kT + threadIdx.x; t < total; t += int64_t(gridDim.x) * kT) {
        const int64_t c = t / per, r = t - c * per;
        const int64_t blk = r / nc;
        const int l = int(r - blk * nc);
        T const* xe = xf + c * ldf + blk * nf;
        double v = 0.0;
        for (int k = 0; k < nf; ++k) v += sI[k * nc + l] * double(xe[k]);
        yc[c * ldc + r] = T(v);
    }
}

// r -= t
template <class T> __global__ __launch_bounds__(kB) void vsub(T* __restrict__ r, T const* __restrict__ t, Cols L) {
    const int64_t c = blockIdx.y, n = L.size(c), o = L.base(c);
    for (int64_t i = int64_t(blockIdx.x) * kB + threadIdx.x; i < n; i += int64_t(gridDim.x) * kB) r[o + i] -= t[o + i];
}
// r = b - t
template <class T>
__global__ __launch_bounds__(kB) void vresid(T* __restrict__ r, T const* __restrict__ b, T const* __restrict__ t, Cols L) {
    const int64_t c = blockIdx.y, n = L.size(c), o = L.base(c);
    for (int64_t i = int64_t(blockIdx.x) * kB + threadIdx.x; i < n; i += int64_t(gridDim.x) * kB)
        r[o + i] = b[o + i] - t[o + i];
}
// r += a t
__global__ __launch_bounds__(kB) void vaxpy(double* __restrict__ r, double const* __restrict__ t, double a, Cols L) {
    const int64_t c = blockIdx.y, n = L.size(c), o = L.base(c);
    for (int64_t i = int64_t(blockIdx.x) * kB + threadIdx.x; i < n; i += int64_t(gridDim.x) * kB) r[o + i] += a * t[o + i];
}
// p = z + beta p
__global__ __launch_bounds__(kB) void vxpay(double* __restrict__ p, double const* __restrict__ z, double beta, Cols L) {
    const int64_t c = blockIdx.y, n = L.size(c), o = L.base(c);
    for (int64_t i = int64_t(blockIdx.x) * kB + threadIdx.x; i < n; i += int64_t(gridDim.x) * kB)
        p[o + i] = z[o + i] + beta * p[o + i];
}
// first Chebyshev step: d = s z, x = d (zero initial guess) or x += d
template <class T>
__global__ __launch_bounds__(kB) void cheb_first(T* __restrict__ x, T* __restrict__ d, T const* __restrict__ z, double s,
                                                 int zero, Cols L) {
    const int64_t c = blockIdx.y, n = L.size(c), o = L.base(c);
    const T sc = T(s);
    for (int64_t i = int64_t(blockIdx.x) * kB + threadIdx.x; i < n; i += int64_t(gridDim.x) * kB) {
        const T v = sc * z[o + i];
        d[o + i] = v;
        x[o + i] = zero ? v : x[o + i] + v;
    }
}
// later steps: d = a d + b z, x += d
template <class T>
__global__ __launch_bounds__(kB) void cheb_next(T* __restrict__ x, T* __restrict__ d, T const* __restrict__ z, double a,
                                                double b, Cols L) {
    const int64_t c = blockIdx.y, n = L.size(c), o = L.base(c);
    const T ac = T(a), bc = T(b);
    for (int64_t i = int64_t(blockIdx.x) * kB + threadIdx.x; i < n; i += int64_t(gridDim.x) * kB) {
        const T v = ac * d[o + i] + bc * z[o + i];
        d[o + i] = v;
        x[o + i] += v;
    }
}
// d = s (precision conversion), columns at strides lds / ldd
template <class D, class S>
__global__ __launch_bounds__(kB) void vconvert(D* __restrict__ d, S const* __restrict__ s, int64_t n, int64_t ldd,
                                               int64_t lds) {
    const int64_t c = blockIdx.y;
    for (int64_t i = int64_t(blockIdx.x) * kB + threadIdx.x; i < n; i += int64_t(gridDim.x) * kB)
        d[c * ldd + i] = D(s[c * lds + i]);
}
// a fixed pseudo-random vector in [-1, 1) (splitmix64 of the index)
__global__ __launch_bounds__(kT) void fill_hash(double* __restrict__ v, int64_t n) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n; i += int64_t(gridDim.x) * kT) {
        uint64_t z = uint64_t(i) + 0x9e3779b97f4a7c15ull;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        z ^= z >> 31;
        v[i] = double(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
    }
}

// Largest eigenvalue of the symmetric tridiagonal matrix (diagonal a, off-diagonal b) by bisection
// on the Sturm count.
double tridiag_max_eig(std::vector<double> const& a, std::vector<double> const& b) {
    const int m = int(a.size());
    double hi = 0.0, lo = 0.0;
    for (int j = 0; j < m; ++j) {
        const double r = (j > 0 ? std::fabs(b[j - 1]) : 0.0) + (j + 1 < m ? std::fabs(b[j]) : 0.0);
        hi = std::max(hi, a[j] + r);
        lo = std::min(lo, a[j] - r);
    }
    auto below = [&](double x) { // eigenvalues < x
        int cnt = 0;
        double d = 1.0;
        for (int j = 0; j < m; ++j) {
            const double off = j > 0 ? b[j - 1] * b[j - 1] : 0.0;
            d = a[j] - x - (j > 0 ? off / d : 0.0);
            if (d == 0.0) d = -1e-300;
            if (d < 0.0) ++cnt;
        }
        return cnt;
    };
    for (int it = 0; it < 200 && hi - lo > 1e-12 * std::max(1.0, std::fabs(hi)); ++it) {
        const double mid = 0.5 * (lo + hi);
        if (below(mid) >= m) hi = mid;
        else lo = mid;
    }
    return hi;
}

} // namespace

PMG::PMG(std::vector<Operator*> levels, PMGOptions const& o) : ops_(std::move(levels)), o_(o) {
    if (ops_.size() < 2) throw std::invalid_argument("p-multigrid needs at least two levels");
    if (o_.smooth < 1 || !(o_.eig_hi > o_.eig_lo) || !(o_.eig_lo > 0.0) || o_.eig_iters < 2 || o_.coarse_max_iter < 1 ||
        !(o_.coarse_rtol > 0.0))
        throw std::invalid_argument("bad p-multigrid options");
    Operator const& f = *ops_[0];
    for (std::size_t l = 0; l < ops_.size(); ++l) {
        Operator const* a = ops_[l];
        if (!a) throw std::invalid_argument("null level operator");
        if (!a->ready()) throw std::logic_error("set_material() must be called on every level");
        if (a->kind() != f.kind() || a->dim() != f.dim() || a->num_elements() != f.num_elements() ||
            a->num_facets() != f.num_facets() || a->quantities() != f.quantities())
            throw std::invalid_argument("p-multigrid levels must discretise one problem on one mesh");
        if (l > 0 && a->degree() >= ops_[l - 1]->degree())
            throw std::invalid_argument("p-multigrid levels must have strictly decreasing degree");
    }
    for (std::size_t l = 1; l < ops_.size(); ++l) {
        if (o_.galerkin_penalty) ops_[l]->set_penalty(f.penalty());
        ops_[l]->set_stream(f.stream());
    }
    for (std::size_t l = 0; l < ops_.size(); ++l) {
        auto v = std::make_unique<PMGLevel>();
        v->n = int64_t(ops_[l]->local_size());
        v->nbf = int(ops_[l]->nbf());
        if (l > 0) {
            auto I = Operator::interpolation_matrix(f.dim(), ops_[l - 1]->degree(), ops_[l]->degree());
            v->I.alloc(I.size());
            MHIP(hipMemcpy(v->I.p, I.data(), I.size() * sizeof(double), hipMemcpyHostToDevice));
        }
        lv_.push_back(std::move(v));
    }
    lmax_.assign(ops_.size() - 1, 0.0);
    coarse_work_ = std::make_unique<MultiWork>();
}

PMG::~PMG() = default;

// lambda_max of D^-1 A on level l: Lanczos tridiagonal from eig_iters preconditioned CG steps on a
// fixed pseudo-random right-hand side.
double PMG::estimate(int l) {
    Operator& A = *ops_[l];
    PMGLevel& v = *lv_[l];
    BlockJacobi const& M = A.block_jacobi();
    hipStream_t st = A.stream();
    const int64_t n = v.n;
    v.r.ensure(n);
    v.d.ensure(n);
    v.z.ensure(n);
    v.t.ensure(n);
    DBuf<double> S, partial;
    S.alloc(kS);
    partial.alloc(kG);
    std::vector<double> hs(kS, 0.0);
    hs[5] = 1.0;
    MHIP(hipMemcpy(S.p, hs.data(), kS * sizeof(double), hipMemcpyHostToDevice));
    Cols C;
    C.ld = n;
    C.n = n;
    auto dot = [&](double const* a, double const* b) {
        hipLaunchKernelGGL((mdot<double, double>), dim3(kG, 1), dim3(kB), 0, st, a, b, C, partial.p);
        hipLaunchKernelGGL(mfinal, dim3(1), dim3(kB), 0, st, partial.p, S.p, 0, -1);
        MHIP(hipGetLastError());
        double h = 0.0;
        MHIP(hipMemcpyAsync(&h, S.p, sizeof(double), hipMemcpyDeviceToHost, st));
        MHIP(hipStreamSynchronize(st));
        return h;
    };
    hipLaunchKernelGGL(fill_hash, dim3(grid_of(n)), dim3(kT), 0, st, v.r.p, n);
    M.apply(v.r.p, v.z.p, st, 1, n);
    hipLaunchKernelGGL((mcopy<double>), dim3(kG, 1), dim3(kB), 0, st, v.d.p, v.z.p, C);
    double rho = dot(v.r.p, v.z.p);
    std::vector<double> al, be;
    for (int j = 0; j < o_.eig_iters && rho > 0.0; ++j) {
        A.apply_batched(v.d.p, v.t.p, 1, n);
        const double pq = dot(v.d.p, v.t.p);
        if (!(pq > 0.0)) break;
        const double alpha = rho / pq;
        hipLaunchKernelGGL(vaxpy, dim3(kG, 1), dim3(kB), 0, st, v.r.p, v.t.p, -alpha, C); // r -= alpha A p
        MHIP(hipGetLastError());
        al.push_back(alpha);
        M.apply(v.r.p, v.z.p, st, 1, n);
        const double rho1 = dot(v.r.p, v.z.p);
        const double beta = rho1 / rho;
        be.push_back(beta);
        hipLaunchKernelGGL(vxpay, dim3(kG, 1), dim3(kB), 0, st, v.d.p, v.z.p, beta, C); // p = z + beta p
        MHIP(hipGetLastError());
        rho = rho1;
    }
    const int m = int(al.size());
    if (m == 0) throw std::runtime_error("p-multigrid eigenvalue estimate failed");
    std::vector<double> a(m), b(m > 1 ? m - 1 : 0);
    for (int j = 0; j < m; ++j) {
        a[j] = 1.0 / al[j] + (j > 0 ? be[j - 1] / al[j - 1] : 0.0);
        if (j + 1 < m) b[j] = std::sqrt(std::max(be[j], 0.0)) / al[j];
    }
    MHIP(hipStreamSynchronize(st));
    return tridiag_max_eig(a, b);
}

void PMG::setup() {
    if (ready_) return;
    Operator& coarse = *ops_.back();
    if (o_.coarse_assembled && coarse.assembled_fast()) {
        coarse.assemble();
        coarse.set_variant(Variant_Assembled);
        // (a coarse level small enough for one dense solve keeps block Jacobi: an exact FP32
        // preconditioner leaves the FP32 recursion in rounding noise between two residual checks)
        const bool amg_shape = coarse.kind() == Kind_Elasticity &&
                               ((coarse.dim() == 2 && coarse.degree() <= 2) || (coarse.dim() == 3 && coarse.degree() == 1));
        if (o_.coarse_amg && amg_shape && int64_t(coarse.local_size()) > AMGOptions().coarse_dofs)
            amg_ = std::make_unique<AMG>(coarse);
    }
    for (auto* a : ops_) {
        a->upload_device();
        (void)a->block_jacobi();
    }
    for (std::size_t l = 0; l + 1 < ops_.size(); ++l) lmax_[l] = estimate(int(l));
    ready_ = true;
}

// Element-blocked level vector of precision T in a buffer sized in doubles (large enough for either).
template <class T> static T* as(DBuf<double>& b) { return reinterpret_cast<T*>(b.p); }

template <class T>
void PMG::smooth(int l, T const* b, T* x, int ncols, int64_t ld, bool zero, bool residual) {
    Operator& A = *ops_[l];
    PMGLevel& v = *lv_[l];
    BlockJacobi const& M = A.block_jacobi();
    hipStream_t st = A.stream();
    const double lo = o_.eig_lo * lmax_[l], hi = o_.eig_hi * lmax_[l];
    const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo), sigma = theta / delta;
    T *r = as<T>(v.r), *d = as<T>(v.d), *z = as<T>(v.z), *t = as<T>(v.t);
    Cols C;
    C.ld = ld;
    C.n = v.n;
    const dim3 g(kG, ncols), bl(kB);
    if (zero) {
        hipLaunchKernelGGL((mcopy<T>), g, bl, 0, st, r, b, C);
    } else {
        A.apply_batched(x, t, ncols, ld);
        hipLaunchKernelGGL((vresid<T>), g, bl, 0, st, r, b, t, C);
    }
    M.apply(r, z, st, ncols, ld);
    hipLaunchKernelGGL((cheb_first<T>), g, bl, 0, st, x, d, z, 1.0 / theta, zero ? 1 : 0, C);
    double rho0 = 1.0 / sigma;
    for (int i = 1; i < o_.smooth; ++i) {
        A.apply_batched(d, t, ncols, ld);
        hipLaunchKernelGGL((vsub<T>), g, bl, 0, st, r, t, C);
        M.apply(r, z, st, ncols, ld);
        const double rho1 = 1.0 / (2.0 * sigma - rho0);
        hipLaunchKernelGGL((cheb_next<T>), g, bl, 0, st, x, d, z, rho1 * rho0, 2.0 * rho1 / delta, C);
        rho0 = rho1;
    }
    if (residual) {
        A.apply_batched(d, t, ncols, ld);
        hipLaunchKernelGGL((vsub<T>), g, bl, 0, st, r, t, C);
    }
    MHIP(hipGetLastError());
}

template <class T> void PMG::cycle(int l, T const* b, T* x, int ncols, int64_t ld) {
    const int last = int(ops_.size()) - 1;
    if (l == last) {
        SolveOptions co;
        co.rtol = o_.coarse_rtol;
        co.max_iter = o_.coarse_max_iter;
        co.mode = o_.coarse_fp32 ? Solve_FP32_CG_FP64_RELIABLE : Solve_FP64_CG;
        co.reliable_delta = 1e-3; // no replacement above the loose coarse tolerance
        co.precond = amg_ ? Precond_None : Precond_BlockJacobi;
        co.check_every = amg_ ? 2 : 10; // AMG steps gain an order of magnitude each: check often
        co.graphs = 0;
        Cols C;
        C.ld = ld;
        C.n = lv_[l]->n;
        double const* b64 = reinterpret_cast<double const*>(b);
        double* x64 = reinterpret_cast<double*>(x);
        hipStream_t st = ops_[l]->stream();
        const dim3 g(kG, ncols), bl(kB);
        if constexpr (sizeof(T) == 4) { // the coarse solver takes FP64 right-hand sides
            PMGLevel& v = *lv_[l];
            hipLaunchKernelGGL((vconvert<double, T>), g, bl, 0, st, v.r.p, b, C.n, ld, ld);
            b64 = v.r.p;
            x64 = v.t.p;
        }
        CallerPrecond amg;
        amg.f32 = [&](float const* rr, float* zz) { amg_->vcycle(rr, zz, ncols, ld); };
        amg.f64 = [&](double const* rr, double* zz) { amg_->vcycle(rr, zz, ncols, ld); };
        auto r = lockstep_cg(*ops_[l], b64, x64, ncols, C, std::size_t(ncols) * std::size_t(ld), false, co,
                             coarse_work_.get(), amg_ ? &amg : nullptr);
        if constexpr (sizeof(T) == 4) hipLaunchKernelGGL((vconvert<T, double>), g, bl, 0, st, x, x64, C.n, ld, ld);
        MHIP(hipGetLastError());
        coarse_its_ += r.total_iterations;
        return;
    }
    Operator const& A = *ops_[l];
    PMGLevel& f = *lv_[l];
    PMGLevel& c = *lv_[l + 1];
    hipStream_t st = A.stream();
    const int nq = int(A.quantities());
    const int64_t E = int64_t(A.num_elements());
    smooth<T>(l, b, x, ncols, ld, true, true); // x = S b, r = b - A x
    const std::size_t shm = std::size_t(f.nbf) * std::size_t(c.nbf) * sizeof(double);
    hipLaunchKernelGGL((restrict_t<T>), dim3(grid_of(int64_t(ncols) * c.n)), dim3(kT), shm, st, c.I.p, f.nbf, c.nbf,
                       nq, E, ncols, as<T>(f.r), ld, as<T>(c.b), c.n);
    MHIP(hipGetLastError());
    cycle<T>(l + 1, as<T>(c.b), as<T>(c.x), ncols, c.n);
    hipLaunchKernelGGL((prolong_add<T>), dim3(grid_of(int64_t(ncols) * f.n)), dim3(kT), shm, st, c.I.p, f.nbf, c.nbf,
                       nq, E, ncols, as<T>(c.x), c.n, x, ld);
    MHIP(hipGetLastError());
    smooth<T>(l, b, x, ncols, ld, false, false);
}

void PMG::vcycle(double const* r, double* z, int ncols, int64_t ld) {
    if (ncols < 1 || ld < lv_[0]->n) throw std::invalid_argument("bad V-cycle column layout");
    setup();
    const std::size_t nf = std::size_t(ncols) * std::size_t(ld);
    const bool f32 = o_.vcycle_fp32 != 0;
    PMGLevel& f = *lv_[0];
    f.r.ensure(nf);
    f.d.ensure(nf);
    f.z.ensure(nf);
    f.t.ensure(nf);
    if (f32) {
        f.b.ensure(nf);
        f.x.ensure(nf);
    }
    const int last = int(lv_.size()) - 1;
    for (int l = 1; l <= last; ++l) {
        PMGLevel& v = *lv_[l];
        const std::size_t m = std::size_t(ncols) * std::size_t(v.n);
        v.b.ensure(m);
        v.x.ensure(m);
        if (l < last || f32) { // FP32: the coarsest level converts through r and t
            v.r.ensure(m);
            v.t.ensure(m);
        }
        if (l < last) {
            v.d.ensure(m);
            v.z.ensure(m);
        }
    }
    for (auto* a : ops_) a->reserve_columns(ncols);
    if (f32) {
        hipStream_t st = ops_[0]->stream();
        const dim3 g(kG, ncols), bl(kB);
        const int64_t n = lv_[0]->n;
        hipLaunchKernelGGL((vconvert<float, double>), g, bl, 0, st, as<float>(f.b), r, n, ld, ld);
        cycle<float>(0, as<float>(f.b), as<float>(f.x), ncols, ld);
        hipLaunchKernelGGL((vconvert<double, float>), g, bl, 0, st, z, as<float>(f.x), n, ld, ld);
        MHIP(hipGetLastError());
    } else {
        cycle<double>(0, r, z, ncols, ld);
    }
    ++cycles_;
}

MultiSolveResult solve_multi_pmg(PMG& M, double const* B, double* X, int ncols, int64_t ld, SolveOptions const& o) {
    Operator& op = M.fine();
    if (!op.ready()) throw std::logic_error("set_material() must be called before solve");
    if (ncols < 1 || ld < int64_t(op.local_size())) throw std::invalid_argument("bad column layout");
    M.setup();
    op.reserve_columns(ncols);
    hipStream_t st = op.stream();
    Cols L;
    L.ld = ld;
    L.n = int64_t(op.local_size());
    const std::size_t tot = std::size_t(ncols) * std::size_t(ld);
    const dim3 g(kG, ncols), bl(kB);
    MultiWork W;
    W.S.alloc(std::size_t(ncols) * kS);
    W.partial.alloc(std::size_t(ncols) * kG);
    W.r.alloc(tot);
    W.p.alloc(tot);
    W.q.alloc(tot);
    W.z.alloc(tot);
    double *S = W.S.p, *partial = W.partial.p, *r = W.r.p, *p = W.p.p, *q = W.q.p, *z = W.z.p;
    std::vector<double> hs(std::size_t(ncols) * kS, 0.0);
    auto pull = [&] {
        MHIP(hipMemcpyAsync(hs.data(), S, hs.size() * sizeof(double), hipMemcpyDeviceToHost, st));
        MHIP(hipStreamSynchronize(st));
    };
    auto push = [&] { MHIP(hipMemcpyAsync(S, hs.data(), hs.size() * sizeof(double), hipMemcpyHostToDevice, st)); };
    auto dot = [&](double const* a, double const* b, int slot) {
        hipLaunchKernelGGL((mdot<double, double>), g, bl, 0, st, a, b, L, partial);
        hipLaunchKernelGGL(mfinal, dim3(ncols), bl, 0, st, partial, S, slot, -1);
    };
    MultiSolveResult res;
    res.iterations.assign(ncols, 0);
    res.rel_residual.assign(ncols, 0.0);
    MHIP(hipStreamSynchronize(st));
    const auto t0 = std::chrono::steady_clock::now();
    Progress progress("p-multigrid flexible CG");
    for (int c = 0; c < ncols; ++c) hs[c * kS + 5] = 1.0;
    push();
    dot(B, B, 3);
    MHIP(hipGetLastError());
    pull();
    std::vector<double> target(ncols);
    for (int c = 0; c < ncols; ++c) {
        target[c] = o.rtol * o.rtol * hs[c * kS + 3];
        if (hs[c * kS + 3] == 0.0) hs[c * kS + 5] = 0.0;
    }
    push();
    hipLaunchKernelGGL((mfill<double>), g, bl, 0, st, X, 0.0, L);
    hipLaunchKernelGGL((mcopy<double>), g, bl, 0, st, r, B, L);
    M.vcycle(r, z, ncols, ld);
    hipLaunchKernelGGL((mcopy<double>), g, bl, 0, st, p, z, L);
    dot(p, r, 0);
    MHIP(hipGetLastError());
    int it = 0;
    auto worst = [&] {
        double w = 0.0;
        for (int c = 0; c < ncols; ++c)
            if (hs[c * kS + 5] != 0.0 && hs[c * kS + 3] > 0.0) w = std::max(w, std::sqrt(hs[c * kS + 4] / hs[c * kS + 3]));
        return w;
    };
    while (it < o.max_iter) {
        op.apply_batched(p, q, ncols, ld);
        dot(p, q, 2);
        hipLaunchKernelGGL((mxr<double, double>), g, bl, 0, st, X, r, p, q, L, S, 0, partial);
        hipLaunchKernelGGL(mfinal, dim3(ncols), bl, 0, st, partial, S, 4, -1);
        MHIP(hipGetLastError());
        ++it;
        pull();
        int active = 0;
        for (int c = 0; c < ncols; ++c) {
            if (hs[c * kS + 5] == 0.0) continue;
            const double rr = hs[c * kS + 4];
            if (!std::isfinite(rr) || !(hs[c * kS + 2] > 0.0))
                throw std::runtime_error("p-multigrid CG breakdown (non-positive p.Ap or non-finite residual)");
            res.iterations[c] = it;
            if (rr <= target[c]) hs[c * kS + 5] = 0.0;
            else ++active;
        }
        push();
        progress.tick(it, active, worst());
        if (!active) break;
        M.vcycle(r, z, ncols, ld);
        dot(z, q, 7);
        hipLaunchKernelGGL((mpflex<double>), g, bl, 0, st, p, z, L, S);
        dot(p, r, 0);
        MHIP(hipGetLastError());
    }
    res.total_iterations = it;
    // true residuals
    op.apply_batched(X, q, ncols, ld);
    hipLaunchKernelGGL(mresidual, g, bl, 0, st, B, q, r, (float*)nullptr, L, S, 1, partial);
    hipLaunchKernelGGL(mfinal, dim3(ncols), bl, 0, st, partial, S, 4, -1);
    MHIP(hipGetLastError());
    pull();
    res.converged = 1;
    for (int c = 0; c < ncols; ++c) {
        const double bb = hs[c * kS + 3];
        res.rel_residual[c] = bb > 0.0 ? std::sqrt(hs[c * kS + 4] / bb) : 0.0;
        if (!(res.rel_residual[c] <= o.rtol)) res.converged = 0;
    }
    res.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return res;
}

} // namespace tdg
