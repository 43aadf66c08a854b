// Element-block Jacobi preconditioner: probing of the diagonal blocks with the matrix-free apply,
// in-LDS Gauss-Jordan inversion (one workgroup per element), and the block application.
#include "block_jacobi.h"

#include "../host/operator.h"

#include <algorithm>
#include <stdexcept>
#include <string>
#include <type_traits>

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
__global__ __launch_bounds__(kThreads) void invert_blocks(double* __restrict__ Dblk, float* __restrict__ inv32, int bs) {
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
    // symmetrise the inverse (Gauss-Jordan leaves rounding-level asymmetry): CG sees an exactly
    // symmetric preconditioner
    float* g32 = inv32 ? inv32 + e * int64_t(bs) * bs : nullptr;
    for (int t = threadIdx.x; t < nn; t += kThreads) {
        const int i = t % bs, j = t / bs;
        const double v = 0.5 * (A[i + j * bs] + A[j + i * bs]);
        g[t] = v;
        if (g32) g32[t] = float(v);
    }
}

// Packed symmetric blocks: the lower triangle column by column, entry (i, j), i >= j, at
// start(j) + i - j with start(j) = j bs - j (j - 1) / 2, each element's block padded to tp entries
// (16-byte multiples).  The inverse blocks are exactly symmetric, so this halves the bytes of every
// application and changes no value.
__device__ __forceinline__ int packed_start(int j, int bs) { return j * bs - (j * (j - 1)) / 2; }
__device__ __forceinline__ int packed_index(int i, int j, int bs) {
    return i >= j ? packed_start(j, bs) + (i - j) : packed_start(i, bs) + (j - i);
}

template <class TB>
__global__ __launch_bounds__(kThreads) void pack_blocks(double const* __restrict__ full, TB* __restrict__ pk,
                                                        int bs, int64_t tp) {
    const int64_t e = blockIdx.x;
    double const* g = full + e * int64_t(bs) * bs;
    const int tri = bs * (bs + 1) / 2;
    for (int t = threadIdx.x; t < bs * bs; t += kThreads) {
        const int i = t % bs, j = t / bs;
        if (i < j) continue;
        pk[e * tp + packed_start(j, bs) + (i - j)] = TB(g[t]);
    }
    for (int64_t t = tri + threadIdx.x; t < tp; t += kThreads) pk[e * tp + t] = TB(0);
}

// z_e = inv_e r_e from packed blocks, one workgroup of 256 lanes per element: the packed block is
// copied to LDS with 16-byte loads (coalesced, half the bytes of the full block; four waves keep
// more loads in flight), then lane i < bs forms row i in the same order as the full-block kernel
// (bit-identical results).
template <class T, class TB>
__global__ __launch_bounds__(256) void apply_blocks_packed(TB const* __restrict__ pk, T const* __restrict__ rs_,
                                                           T* __restrict__ zs, int bs, int64_t tp, int64_t ld) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    TB* sB = reinterpret_cast<TB*>(lds_raw);
    T* sr = reinterpret_cast<T*>(lds_raw + tp * sizeof(TB));
    const int64_t e = blockIdx.x;
    T const* __restrict__ r = rs_ + int64_t(blockIdx.y) * ld;
    T* __restrict__ z = zs + int64_t(blockIdx.y) * ld;
    const int tid = threadIdx.x;
    if (tid < bs) sr[tid] = r[e * bs + tid];
    using V = typename std::conditional<sizeof(TB) == 8, double2, float4>::type;
    constexpr int per = 16 / sizeof(TB);
    V const* g = reinterpret_cast<V const*>(pk + e * tp);
    V* d = reinterpret_cast<V*>(sB);
#pragma unroll 4
    for (int64_t t = tid; t < tp / per; t += 256) d[t] = g[t];
    __syncthreads();
    if (tid >= bs) return;
    const int i = tid;
    T acc = T(0);
    // entry (i, j): (i, j) for j <= i, else (j, i); the packed index steps by bs - j - 1 down
    // column j while j < i and by 1 along column i afterwards
    int idx = i;
    for (int j = 0; j < bs; ++j) {
        acc += T(sB[idx]) * sr[j];
        idx += j < i ? bs - j - 1 : 1;
    }
    z[e * bs + i] = acc;
}

// z_e = inv_e r_e, one workgroup per element (column-major blocks: lanes read consecutive rows);
// blocks of precision TB, vectors and arithmetic of precision T
template <class T, class TB>
__global__ __launch_bounds__(128) void apply_blocks(TB const* __restrict__ inv, T const* __restrict__ rs_, T* __restrict__ zs,
                                                    int bs, int64_t ld) {
    __shared__ T rs[128];
    const int64_t e = blockIdx.x;
    const int i = threadIdx.x;
    T const* __restrict__ r = rs_ + int64_t(blockIdx.y) * ld;
    T* __restrict__ z = zs + int64_t(blockIdx.y) * ld;
    if (i < bs) rs[i] = r[e * bs + i];
    __syncthreads();
    if (i >= bs) return;
    TB const* B = inv + e * int64_t(bs) * bs;
    T acc = T(0);
    for (int j = 0; j < bs; ++j) acc += T(B[i + j * bs]) * rs[j];
    z[e * bs + i] = acc;
}

// Several columns: one workgroup per element holds the element's inverse block in LDS and applies
// it to every column (columns at stride ld), so a block is read from HBM once per application
// instead of once per column.  Each lane computes four columns of one row.
constexpr int kCols = 16; // columns staged per pass
template <class T, class TB, int BSMAX>
__global__ __launch_bounds__(256) void apply_blocks_cols(TB const* __restrict__ inv, T const* __restrict__ rs,
                                                         T* __restrict__ zs, int bs, int ncols, int64_t ld,
                                                         int64_t tp) {
    __shared__ TB B[BSMAX * BSMAX];
    __shared__ T R[kCols * BSMAX];
    const int64_t e = blockIdx.x;
    const int nn = bs * bs;
    if (tp > 0) { // packed symmetric blocks, unpacked into LDS
        TB const* Bg = inv + e * tp;
        for (int t = threadIdx.x; t < nn; t += 256) B[t] = Bg[packed_index(t % bs, t / bs, bs)];
    } else {
        TB const* Bg = inv + e * int64_t(bs) * bs;
        for (int t = threadIdx.x; t < nn; t += 256) B[t] = Bg[t];
    }
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
                const T b = T(B[i + jj * bs]);
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

// Small blocks (P1 / P2 levels): one lane per element keeps the element's inverse block in registers
// and applies it to a range of up to kSmallCols columns; waves stride over groups of 64 elements
// with kDotParts workgroups per column range.  Optionally the per-workgroup partial sums of r . z of
// every column (partial[c kDotParts + workgroup], the layout of the solvers' two-stage reductions).
constexpr int kSmallCols = 16;
template <class T, class TB, int BS>
__global__ __launch_bounds__(256) void apply_blocks_small(TB const* __restrict__ inv, T const* __restrict__ rs,
                                                          T* __restrict__ zs, int64_t E, int ncols, int64_t ld,
                                                          double* __restrict__ partial) {
    __shared__ double red[4][kSmallCols];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int c0 = blockIdx.y * kSmallCols;
    const int nc = min(kSmallCols, ncols - c0);
    double dacc[kSmallCols];
#pragma unroll
    for (int k = 0; k < kSmallCols; ++k) dacc[k] = 0.0;
    const int64_t groups = (E + 63) / 64;
    for (int64_t g = int64_t(blockIdx.x) * 4 + wave; g < groups; g += int64_t(gridDim.x) * 4) {
        const int64_t e = g * 64 + lane;
        const bool act = e < E;
        const int64_t ec = act ? e : 0;
        TB b[BS * BS];
#pragma unroll
        for (int t = 0; t < BS * BS; ++t) b[t] = inv[ec * (BS * BS) + t]; // column-major block
#pragma unroll
        for (int k = 0; k < kSmallCols; ++k) {
            if (k >= nc) break;
            T const* r = rs + int64_t(c0 + k) * ld + ec * BS;
            T rv[BS];
#pragma unroll
            for (int j = 0; j < BS; ++j) rv[j] = r[j];
            T z[BS];
#pragma unroll
            for (int i = 0; i < BS; ++i) z[i] = T(0);
#pragma unroll
            for (int j = 0; j < BS; ++j)
#pragma unroll
                for (int i = 0; i < BS; ++i) z[i] += T(b[i + j * BS]) * rv[j];
            if (act) {
                T* zo = zs + int64_t(c0 + k) * ld + ec * BS;
                double d = 0.0;
#pragma unroll
                for (int i = 0; i < BS; ++i) {
                    zo[i] = z[i];
                    d += double(rv[i]) * double(z[i]);
                }
                dacc[k] += d;
            }
        }
    }
    if (!partial) return;
#pragma unroll
    for (int k = 0; k < kSmallCols; ++k) {
        double v = dacc[k];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
        if (lane == 0) red[wave][k] = v;
    }
    __syncthreads();
    if (threadIdx.x < nc)
        partial[int64_t(c0 + threadIdx.x) * kDotParts + blockIdx.x] =
            ((red[0][threadIdx.x] + red[1][threadIdx.x]) + red[2][threadIdx.x]) + red[3][threadIdx.x];
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
    for (void* p : {static_cast<void*>(inv64_), static_cast<void*>(inv32_), static_cast<void*>(p32_),
                    static_cast<void*>(p64_)})
        if (p) (void)hipFree(p);
    inv64_ = nullptr;
    inv32_ = nullptr;
    p32_ = nullptr;
    p64_ = nullptr;
    bs_ = 0;
    tp_ = 0;
}

// Blocks above kSmallBlock rows are kept packed (lower triangle, FP64 and FP32); small blocks stay
// whole (the register kernel reads them).
constexpr int kSmallBlock = 12;

void BlockJacobi::build(Operator& op) {
    if (!op.ready()) throw std::logic_error("set_material() must be called before building the preconditioner");
    release();
    op.upload_device();
    hipStream_t st = op.stream();
    const int bs = int(op.block_size());
    const int64_t E = int64_t(op.num_elements()), n = int64_t(op.local_size());
    if (bs > 112) throw std::invalid_argument("block size above 112 is not supported by the block inversion");
    const bool packed = bs > kSmallBlock;
    const int64_t tp = packed ? (int64_t(bs) * (bs + 1) / 2 + 3) / 4 * 4 : 0;
    int nc = 0;
    std::vector<int32_t> col = colour_elements(op, &nc);
    int32_t* dcol = nullptr;
    double *X = nullptr, *Y = nullptr;
    const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(bs, (int64_t(1) << 30) / (n * 8)));
    try {
        BHIP(hipMalloc(reinterpret_cast<void**>(&dcol), col.size() * sizeof(int32_t)));
        BHIP(hipMemcpy(dcol, col.data(), col.size() * sizeof(int32_t), hipMemcpyHostToDevice));
        BHIP(hipMalloc(reinterpret_cast<void**>(&inv64_), std::size_t(E) * bs * bs * sizeof(double)));
        if (!packed) BHIP(hipMalloc(reinterpret_cast<void**>(&inv32_), std::size_t(E) * bs * bs * sizeof(float)));
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
                                   bs, inv64_);
                BHIP(hipGetLastError());
            }
        if (bs <= 32)
            hipLaunchKernelGGL((invert_blocks<32>), dim3(unsigned(E)), dim3(kThreads), 0, st, inv64_, inv32_, bs);
        else if (bs <= 64)
            hipLaunchKernelGGL((invert_blocks<64>), dim3(unsigned(E)), dim3(kThreads), 0, st, inv64_, inv32_, bs);
        else
            hipLaunchKernelGGL((invert_blocks<112>), dim3(unsigned(E)), dim3(kThreads), 0, st, inv64_, inv32_, bs);
        BHIP(hipGetLastError());
        if (packed) {
            BHIP(hipMalloc(reinterpret_cast<void**>(&p32_), std::size_t(E) * tp * sizeof(float)));
            BHIP(hipMalloc(reinterpret_cast<void**>(&p64_), std::size_t(E) * tp * sizeof(double)));
            hipLaunchKernelGGL((pack_blocks<float>), dim3(unsigned(E)), dim3(kThreads), 0, st, inv64_, p32_, bs, tp);
            hipLaunchKernelGGL((pack_blocks<double>), dim3(unsigned(E)), dim3(kThreads), 0, st, inv64_, p64_, bs, tp);
            BHIP(hipGetLastError());
        }
        BHIP(hipStreamSynchronize(st));
        if (packed) {
            (void)hipFree(inv64_);
            inv64_ = nullptr;
        }
    } catch (...) {
        if (dcol) (void)hipFree(dcol);
        if (X) (void)hipFree(X);
        if (Y) (void)hipFree(Y);
        release();
        throw;
    }
    (void)hipFree(dcol);
    (void)hipFree(X);
    (void)hipFree(Y);
    bs_ = bs;
    tp_ = tp;
    E_ = E;
    colours_ = nc;
}

namespace {
template <class T, class TB, int BS>
void launch_small(TB const* inv, T const* r, T* z, int64_t E, int ncols, int64_t ld, double* partial, hipStream_t s) {
    hipLaunchKernelGGL((apply_blocks_small<T, TB, BS>), dim3(kDotParts, unsigned((ncols + kSmallCols - 1) / kSmallCols)),
                       dim3(256), 0, s, inv, r, z, E, ncols, ld, partial);
    BHIP(hipGetLastError());
}

// The register kernel's block sizes: elasticity P1 (2D 6, 3D 12) and P2 (2D 12), Poisson P1 / P2
// (3, 4, 6, 10); 10 and 12 in FP32 only.
template <class T, class TB>
bool apply_small(TB const* inv, T const* r, T* z, int64_t E, int bs, int ncols, int64_t ld, double* partial,
                 hipStream_t s) {
    if (!inv || ncols < 1 || (ncols > 1 && ld < E * bs)) return false;
    switch (bs) {
    case 3: launch_small<T, TB, 3>(inv, r, z, E, ncols, ld, partial, s); return true;
    case 4: launch_small<T, TB, 4>(inv, r, z, E, ncols, ld, partial, s); return true;
    case 6: launch_small<T, TB, 6>(inv, r, z, E, ncols, ld, partial, s); return true;
    default: break;
    }
    if constexpr (sizeof(T) == 4 && sizeof(TB) == 4) { // FP64 registers would not hold these blocks
        if (bs == 10) {
            launch_small<T, TB, 10>(inv, r, z, E, ncols, ld, partial, s);
            return true;
        }
        if (bs == 12) {
            launch_small<T, TB, 12>(inv, r, z, E, ncols, ld, partial, s);
            return true;
        }
    }
    return false;
}

// full: blocks stored whole (bs x bs, column-major); tp > 0: packed lower triangles of tp entries
template <class T, class TB>
void launch_apply(TB const* inv, int64_t tp, T const* r, T* z, int64_t E, int bs, int ncols, int64_t ld,
                  hipStream_t s) {
    if (tp == 0 && apply_small(inv, r, z, E, bs, ncols, ld, static_cast<double*>(nullptr), s)) return;
    if (ncols <= 1) {
        if (tp > 0) {
            const std::size_t lds = std::size_t(tp) * sizeof(TB) + std::size_t(bs) * sizeof(T);
            hipLaunchKernelGGL((apply_blocks_packed<T, TB>), dim3(unsigned(E), 1u), dim3(256), lds, s, inv, r, z, bs,
                               tp, ld);
        } else {
            hipLaunchKernelGGL((apply_blocks<T, TB>), dim3(unsigned(E), 1u), dim3(128), 0, s, inv, r, z, bs, ld);
        }
    } else if (bs <= 64) {
        hipLaunchKernelGGL((apply_blocks_cols<T, TB, 64>), dim3(unsigned(E)), dim3(256), 0, s, inv, r, z, bs, ncols,
                           ld, tp);
    } else {
        hipLaunchKernelGGL((apply_blocks_cols<T, TB, 112>), dim3(unsigned(E)), dim3(256), 0, s, inv, r, z, bs, ncols,
                           ld, tp);
    }
    BHIP(hipGetLastError());
}
} // namespace

void BlockJacobi::apply(double const* r, double* z, hipStream_t s, int ncols, int64_t ld) const {
    if (!built()) throw std::logic_error("block Jacobi preconditioner not built");
    if (tp_) launch_apply(p64_, tp_, r, z, E_, bs_, ncols, ld, s);
    else launch_apply(inv64_, 0, r, z, E_, bs_, ncols, ld, s);
}

void BlockJacobi::apply_fp32_blocks(double const* r, double* z, hipStream_t s, int ncols, int64_t ld) const {
    if (!built()) throw std::logic_error("block Jacobi preconditioner not built");
    if (tp_) launch_apply(p32_, tp_, r, z, E_, bs_, ncols, ld, s);
    else launch_apply(inv32_, 0, r, z, E_, bs_, ncols, ld, s);
}

void BlockJacobi::apply(float const* r, float* z, hipStream_t s, int ncols, int64_t ld) const {
    if (!built()) throw std::logic_error("block Jacobi preconditioner not built");
    if (tp_) launch_apply(p32_, tp_, r, z, E_, bs_, ncols, ld, s);
    else launch_apply(inv32_, 0, r, z, E_, bs_, ncols, ld, s);
}

bool BlockJacobi::apply_dot(float const* r, float* z, hipStream_t s, int ncols, int64_t ld, double* partial) const {
    if (!built()) throw std::logic_error("block Jacobi preconditioner not built");
    return apply_small(inv32_, r, z, E_, bs_, ncols, ld, partial, s); // full (small) blocks only
}

bool BlockJacobi::apply_dot(double const* r, double* z, hipStream_t s, int ncols, int64_t ld, double* partial,
                            bool fp32_blocks) const {
    if (!built()) throw std::logic_error("block Jacobi preconditioner not built");
    return fp32_blocks ? apply_small(inv32_, r, z, E_, bs_, ncols, ld, partial, s)
                       : apply_small(inv64_, r, z, E_, bs_, ncols, ld, partial, s);
}

void BlockJacobi::inverse_block(int64_t e, double* out) const {
    if (!built()) throw std::logic_error("block Jacobi preconditioner not built");
    if (e < 0 || e >= E_) throw std::out_of_range("element index out of range");
    const int bs = bs_;
    if (!tp_) {
        BHIP(hipMemcpy(out, inv64_ + e * int64_t(bs) * bs, std::size_t(bs) * bs * sizeof(double),
                       hipMemcpyDeviceToHost));
        return;
    }
    std::vector<double> pk(static_cast<std::size_t>(tp_));
    BHIP(hipMemcpy(pk.data(), p64_ + e * tp_, std::size_t(tp_) * sizeof(double), hipMemcpyDeviceToHost));
    for (int j = 0; j < bs; ++j)
        for (int i = j; i < bs; ++i) {
            const double v = pk[std::size_t(j * bs - (j * (j - 1)) / 2 + (i - j))];
            out[i + j * bs] = v;
            out[j + i * bs] = v;
        }
}

} // namespace tdg
