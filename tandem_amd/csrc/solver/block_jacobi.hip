// Element-block Jacobi preconditioner: probing of the diagonal blocks with the matrix-free apply,
// in-LDS Gauss-Jordan inversion (one workgroup per element), and the block application.
#include "block_jacobi.h"

#include "../host/operator.h"

#include <algorithm>
#include <stdexcept>
#include <string>

namespace tdg {

namespace {

#define BHIP(call)                                                                                 \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess) throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e_)); \
    } while (0)

constexpr int kThreads = 256;

// X[j][e bs + i] = 1 if element e has colour c and i = c0 + j, else 0 (j < nc)
__global__ __launch_bounds__(kThreads) void probe_columns(double* __restrict__ X, int32_t const* __restrict__ colour,
                                                          int c, int c0, int nc, int64_t n, int bs) {
    const int64_t total = int64_t(nc) * n;
    for (int64_t t = int64_t(blockIdx.x) * kThreads + threadIdx.x; t < total; t += int64_t(gridDim.x) * kThreads) {
        const int64_t j = t / n, k = t % n, e = k / bs;
        const int i = int(k % bs);
        X[t] = (colour[e] == c && i == c0 + int(j)) ? 1.0 : 0.0;
    }
}

// D[e][row + (c0 + j) bs] = Y[j][e bs + row] for the elements of colour c
__global__ __launch_bounds__(kThreads) void gather_columns(double const* __restrict__ Y, int32_t const* __restrict__ colour,
                                                           int c, int c0, int nc, int64_t n, int bs,
                                                           double* __restrict__ Dblk) {
    const int64_t total = int64_t(nc) * n;
    for (int64_t t = int64_t(blockIdx.x) * kThreads + threadIdx.x; t < total; t += int64_t(gridDim.x) * kThreads) {
        const int64_t j = t / n, k = t % n, e = k / bs;
        const int row = int(k % bs);
        if (colour[e] == c) Dblk[e * int64_t(bs) * bs + row + (c0 + j) * bs] = Y[t];
    }
}

// In-place inverse of each (symmetrised) SPD block by Gauss-Jordan elimination without pivoting,
// the block held in LDS; writes the FP64 inverse back and an FP32 copy.
template <int BSMAX>
__global__ __launch_bounds__(kThreads) void invert_blocks(double* __restrict__ Dblk, int bs) {
    __shared__ double A[BSMAX * BSMAX];
    __shared__ double colk[BSMAX];
    const int64_t e = blockIdx.x;
    double* g = Dblk + e * int64_t(bs) * bs;
    const int nn = bs * bs;
    for (int t = threadIdx.x; t < nn; t += kThreads) {
        const int i = t % bs, j = t / bs;
        A[t] = 0.5 * (g[i + j * bs] + g[j + i * bs]);
    }
    __syncthreads();
    for (int k = 0; k < bs; ++k) {
        for (int i = threadIdx.x; i < bs; i += kThreads) colk[i] = A[i + k * bs];
        __syncthreads();
        const double p = 1.0 / colk[k];
        for (int j = threadIdx.x; j < bs; j += kThreads) A[k + j * bs] = j == k ? p : A[k + j * bs] * p;
        __syncthreads();
        for (int t = threadIdx.x; t < nn; t += kThreads) {
            const int i = t % bs, j = t / bs;
            if (i == k) continue;
            A[t] = j == k ? -colk[i] * p : A[t] - colk[i] * A[k + j * bs];
        }
        __syncthreads();
    }
    for (int t = threadIdx.x; t < nn; t += kThreads) g[t] = A[t];
}

// Upper triangle (i <= j) of every symmetric inverse block, packed column by column:
// B(i, j) = B(j, i) at j (j + 1) / 2 + i.  Halves the bytes every application reads and makes the
// applied operator exactly symmetric.
__device__ __forceinline__ int packed_index(int i, int j) { return i <= j ? j * (j + 1) / 2 + i : i * (i + 1) / 2 + j; }

__global__ __launch_bounds__(kThreads) void pack_blocks(double const* __restrict__ full, double* __restrict__ p64,
                                                        float* __restrict__ p32, int bs) {
    const int64_t e = blockIdx.x;
    const int np = bs * (bs + 1) / 2;
    double const* g = full + e * int64_t(bs) * bs;
    for (int t = threadIdx.x; t < np; t += kThreads) {
        // column j of the packed upper triangle holds rows 0..j
        int j = int((sqrt(8.0 * t + 1.0) - 1.0) / 2.0);
        while (j * (j + 1) / 2 > t) --j;
        while ((j + 1) * (j + 2) / 2 <= t) ++j;
        const int i = t - j * (j + 1) / 2;
        const double v = g[i + j * bs];
        p64[e * np + t] = v;
        p32[e * np + t] = float(v);
    }
}

// z_e = inv_e r_e, one workgroup per element: the packed block and r_e staged in LDS, lane i
// computes row i
template <class T, int BSMAX>
__global__ __launch_bounds__(128) void apply_blocks(T const* __restrict__ inv, T const* __restrict__ rs,
                                                    T* __restrict__ zs, int bs) {
    __shared__ T P[BSMAX * (BSMAX + 1) / 2];
    __shared__ T r[BSMAX];
    const int64_t e = blockIdx.x;
    const int np = bs * (bs + 1) / 2;
    T const* Pg = inv + e * int64_t(np);
    for (int t = threadIdx.x; t < np; t += 128) P[t] = Pg[t];
    for (int t = threadIdx.x; t < bs; t += 128) r[t] = rs[e * bs + t];
    __syncthreads();
    const int i = threadIdx.x;
    if (i >= bs) return;
    T acc = T(0);
    for (int jj = 0; jj < bs; ++jj) acc += P[packed_index(i, jj)] * r[jj];
    zs[e * bs + i] = acc;
}

// Several columns: one workgroup per element holds the element's inverse block in LDS and applies
// it to every column (columns at stride ld), so a block is read from HBM once per application
// instead of once per column.  Each lane computes four columns of one row.
constexpr int kCols = 16; // columns staged per pass
template <class T, int BSMAX>
__global__ __launch_bounds__(256) void apply_blocks_cols(T const* __restrict__ inv, T const* __restrict__ rs,
                                                         T* __restrict__ zs, int bs, int ncols, int64_t ld) {
    __shared__ T B[BSMAX * (BSMAX + 1) / 2];
    __shared__ T R[kCols * BSMAX];
    const int64_t e = blockIdx.x;
    const int np = bs * (bs + 1) / 2;
    T const* Bg = inv + e * int64_t(np);
    for (int t = threadIdx.x; t < np; t += 256) B[t] = Bg[t];
    for (int c0 = 0; c0 < ncols; c0 += kCols) {
        const int nc = min(kCols, ncols - c0);
        __syncthreads(); // B is loaded / the previous pass's reads of R are done
        for (int t = threadIdx.x; t < nc * bs; t += 256) {
            const int c = t / bs, i = t % bs;
            R[c * bs + i] = rs[int64_t(c0 + c) * ld + e * bs + i];
        }
        __syncthreads();
        // outputs (row i, columns 4q..4q+3) of this pass
        const int nq4 = (nc + 3) / 4;
        for (int t = threadIdx.x; t < bs * nq4; t += 256) {
            const int i = t % bs, q = t / bs;
            T a0 = T(0), a1 = T(0), a2 = T(0), a3 = T(0);
            T const* r0 = R + (4 * q) * bs;
            for (int jj = 0; jj < bs; ++jj) {
                const T b = B[packed_index(i, jj)];
                a0 += b * r0[jj];
                a1 += b * r0[jj + bs];
                a2 += b * r0[jj + 2 * bs];
                a3 += b * r0[jj + 3 * bs];
            }
            const int c = c0 + 4 * q;
            T* z = zs + e * bs + i;
            z[int64_t(c) * ld] = a0;
            if (4 * q + 1 < nc) z[int64_t(c + 1) * ld] = a1;
            if (4 * q + 2 < nc) z[int64_t(c + 2) * ld] = a2;
            if (4 * q + 3 < nc) z[int64_t(c + 3) * ld] = a3;
        }
    }
}

unsigned grid_for(int64_t items) { return unsigned(std::min<int64_t>((items + kThreads - 1) / kThreads, 65536)); }

} // namespace

std::vector<int32_t> colour_elements(Operator const& op, int* ncolours) {
    const std::size_t E = op.num_elements();
    std::vector<std::vector<int64_t>> nb(E);
    for (auto const& f : op.topology().facets)
        if (f.elem[0] != f.elem[1]) {
            nb[f.elem[0]].push_back(f.elem[1]);
            nb[f.elem[1]].push_back(f.elem[0]);
        }
    std::vector<int32_t> col(E, -1);
    int nc = 0;
    for (std::size_t e = 0; e < E; ++e) {
        uint64_t used = 0;
        for (auto o : nb[e])
            if (col[o] >= 0 && col[o] < 64) used |= uint64_t(1) << col[o];
        int c = 0;
        while (used & (uint64_t(1) << c)) ++c;
        col[e] = c;
        nc = std::max(nc, c + 1);
    }
    *ncolours = nc;
    return col;
}

BlockJacobi::~BlockJacobi() { release(); }

void BlockJacobi::release() {
    if (inv64_) (void)hipFree(inv64_);
    if (inv32_) (void)hipFree(inv32_);
    inv64_ = nullptr;
    inv32_ = nullptr;
}

void BlockJacobi::build(Operator& op) {
    if (!op.ready()) throw std::logic_error("set_material() must be called before building the preconditioner");
    release();
    op.upload_device();
    hipStream_t st = op.stream();
    const int bs = int(op.block_size());
    const int64_t E = int64_t(op.num_elements()), n = int64_t(op.local_size());
    if (bs > 112) throw std::invalid_argument("block size above 112 is not supported by the block inversion");
    int nc = 0;
    std::vector<int32_t> col = colour_elements(op, &nc);
    int32_t* dcol = nullptr;
    double *X = nullptr, *Y = nullptr, *full = nullptr;
    const int64_t np = int64_t(bs) * (bs + 1) / 2;
    const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(bs, (int64_t(1) << 30) / (n * 8)));
    try {
        BHIP(hipMalloc(reinterpret_cast<void**>(&dcol), col.size() * sizeof(int32_t)));
        BHIP(hipMemcpy(dcol, col.data(), col.size() * sizeof(int32_t), hipMemcpyHostToDevice));
        BHIP(hipMalloc(reinterpret_cast<void**>(&full), std::size_t(E) * bs * bs * sizeof(double)));
        BHIP(hipMalloc(reinterpret_cast<void**>(&inv64_), std::size_t(E) * np * sizeof(double)));
        BHIP(hipMalloc(reinterpret_cast<void**>(&inv32_), std::size_t(E) * np * sizeof(float)));
        BHIP(hipMalloc(reinterpret_cast<void**>(&X), std::size_t(chunk) * n * sizeof(double)));
        BHIP(hipMalloc(reinterpret_cast<void**>(&Y), std::size_t(chunk) * n * sizeof(double)));
        for (int c = 0; c < nc; ++c)
            for (int c0 = 0; c0 < bs; c0 += int(chunk)) {
                const int m = int(std::min<int64_t>(chunk, bs - c0));
                hipLaunchKernelGGL(probe_columns, dim3(grid_for(m * n)), dim3(kThreads), 0, st, X, dcol, c, c0, m, n,
                                   bs);
                BHIP(hipGetLastError());
                op.apply_batched(X, Y, m, n);
                hipLaunchKernelGGL(gather_columns, dim3(grid_for(m * n)), dim3(kThreads), 0, st, Y, dcol, c, c0, m, n,
                                   bs, full);
                BHIP(hipGetLastError());
            }
        if (bs <= 32)
            hipLaunchKernelGGL((invert_blocks<32>), dim3(unsigned(E)), dim3(kThreads), 0, st, full, bs);
        else if (bs <= 64)
            hipLaunchKernelGGL((invert_blocks<64>), dim3(unsigned(E)), dim3(kThreads), 0, st, full, bs);
        else
            hipLaunchKernelGGL((invert_blocks<112>), dim3(unsigned(E)), dim3(kThreads), 0, st, full, bs);
        BHIP(hipGetLastError());
        hipLaunchKernelGGL(pack_blocks, dim3(unsigned(E)), dim3(kThreads), 0, st, full, inv64_, inv32_, bs);
        BHIP(hipGetLastError());
        BHIP(hipStreamSynchronize(st));
    } catch (...) {
        if (dcol) (void)hipFree(dcol);
        if (X) (void)hipFree(X);
        if (Y) (void)hipFree(Y);
        if (full) (void)hipFree(full);
        release();
        throw;
    }
    (void)hipFree(dcol);
    (void)hipFree(X);
    (void)hipFree(Y);
    (void)hipFree(full);
    bs_ = bs;
    E_ = E;
    colours_ = nc;
}

namespace {
template <class T> void launch_apply(T const* inv, T const* r, T* z, int64_t E, int bs, int ncols, int64_t ld,
                                     hipStream_t s) {
    if (ncols <= 1) {
        if (bs <= 64) hipLaunchKernelGGL((apply_blocks<T, 64>), dim3(unsigned(E)), dim3(128), 0, s, inv, r, z, bs);
        else hipLaunchKernelGGL((apply_blocks<T, 112>), dim3(unsigned(E)), dim3(128), 0, s, inv, r, z, bs);
    } else if (bs <= 64) {
        hipLaunchKernelGGL((apply_blocks_cols<T, 64>), dim3(unsigned(E)), dim3(256), 0, s, inv, r, z, bs, ncols, ld);
    } else {
        hipLaunchKernelGGL((apply_blocks_cols<T, 112>), dim3(unsigned(E)), dim3(256), 0, s, inv, r, z, bs, ncols, ld);
    }
    BHIP(hipGetLastError());
}
} // namespace

void BlockJacobi::apply(double const* r, double* z, hipStream_t s, int ncols, int64_t ld) const {
    if (!inv64_) throw std::logic_error("block Jacobi preconditioner not built");
    launch_apply(inv64_, r, z, E_, bs_, ncols, ld, s);
}

void BlockJacobi::apply(float const* r, float* z, hipStream_t s, int ncols, int64_t ld) const {
    if (!inv32_) throw std::logic_error("block Jacobi preconditioner not built");
    launch_apply(inv32_, r, z, E_, bs_, ncols, ld, s);
}

void BlockJacobi::inverse_block(int64_t e, double* out) const {
    if (!inv64_) throw std::logic_error("block Jacobi preconditioner not built");
    if (e < 0 || e >= E_) throw std::out_of_range("element index out of range");
    const int64_t np = int64_t(bs_) * (bs_ + 1) / 2;
    std::vector<double> p(static_cast<std::size_t>(np));
    BHIP(hipMemcpy(p.data(), inv64_ + e * np, std::size_t(np) * sizeof(double), hipMemcpyDeviceToHost));
    for (int j = 0; j < bs_; ++j)
        for (int i = 0; i <= j; ++i) out[i + j * bs_] = out[j + i * bs_] = p[std::size_t(j * (j + 1) / 2 + i)];
}

} // namespace tdg
