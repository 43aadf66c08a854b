// Assembled form of the DG operator: one block row per element holding the element's own block and
// one coupling block per face (tandem's PetscDGMatrix, a BAIJ matrix preallocated with 1 + faces
// blocks per block row, app/common/PetscDGMatrix.cpp:14-77, filled by DGOperator::assemble,
// src/form/DGOperator.h:121-170).  The blocks are probed from the matrix-free apply (distance-2
// element colouring: with x = e_j on every element of one colour, y on an element and on each of its
// neighbours is one column of one block), then stored element-innermost:
//     A[slot][row][col][element],  slot 0 = own block, slot 1 + f = block of the face-f neighbour
// (zero where face f has no neighbour), nb[f][element] = that neighbour (the element itself where
// there is none).
//
// Apply kernels:
//   * register kernel (small blocks: P1 and P2 levels of p-multigrid): one lane per element; it keeps
//     R rows of its element's block row in registers (all BS rows where they fit, else BS / R row
//     groups one after another) and sweeps a range of columns, reading per column the element's and
//     its neighbours' x blocks (lanes of consecutive elements read consecutive addresses for the own
//     block) and writing R entries of y.  The matrix is read from HBM once per
//     column range instead of once per column, so a 64-column apply costs little more than the
//     64 x / y column reads and writes.
//   * generic kernel (any block size; the cross-check of the matrix-free apply at high order): one
//     lane per element row and column, the block row streamed from HBM per column.
#include "kernels.h"

#include <algorithm>

namespace tdg {

namespace {

constexpr int kThreads = 256;
constexpr int kWavesPerBlock = kThreads / 64;

// DOT: the workgroups stride over the element blocks (gridDim.x = kDotParts, column ranges of at
// most kDotCols) and write the per-workgroup partial sums of x . y of every column to
// partial[c kDotParts + workgroup] (the solvers' two-stage reduction layout).
constexpr int kDotCols = 16;
template <class T, int BS, int NB, int R, bool DOT>
__global__ __launch_bounds__(kThreads) void assembled_apply_kernel(AssembledArgs<T> a, T const* __restrict__ xs,
                                                                   T* __restrict__ ys, int64_t ld, int ncols,
                                                                   int cchunk, double* __restrict__ partial) {
    constexpr int PARTS = BS / R; // row groups, taken one after another by each lane
    constexpr int EPB = 64 * kWavesPerBlock;
    static_assert(!DOT || PARTS == 1, "the fused sums need the whole block row in one pass");
    __shared__ double red[kWavesPerBlock][kDotCols];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t E = a.num_elements;
    const int c0 = blockIdx.y * cchunk;
    const int c1 = min(ncols, c0 + cchunk);
    if (DOT && lane == 0)
        for (int k = 0; k < kDotCols; ++k) red[wave][k] = 0.0;
    auto block = [&](int64_t blk) {
        // one element per lane: 64 consecutive elements per wave, so every block-row load is one
        // contiguous 64-element stream
        const int64_t e = blk * EPB + int64_t(wave) * 64 + lane;
        const bool act = e < E;
        const int64_t ec = act ? e : 0; // idle lanes read element 0 and store nothing
        int64_t off[NB];
        off[0] = ec * BS;
#pragma unroll
        for (int f = 0; f + 1 < NB; ++f) off[1 + f] = int64_t(a.nb[int64_t(f) * E + ec]) * BS;
#pragma unroll 1
        for (int part = 0; part < PARTS; ++part) {
        T m[NB][R][BS];
#pragma unroll
        for (int b = 0; b < NB; ++b)
#pragma unroll
            for (int r = 0; r < R; ++r)
#pragma unroll
                for (int j = 0; j < BS; ++j) m[b][r][j] = a.A[((int64_t(b) * BS + part * R + r) * BS + j) * E + ec];
        // small block rows hold the next column's x blocks in flight during this column's products
        constexpr bool PRE = NB * BS * int(sizeof(T) / 4) <= 48;
        T v[NB][BS];
        if (PRE && c0 < c1) {
            T const* x = xs + int64_t(c0) * ld;
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int j = 0; j < BS; ++j) v[b][j] = x[off[b] + j];
        }
#pragma unroll 1
        for (int c = c0; c < c1; ++c) {
            T w[NB][BS];
            const int cn = PRE ? min(c + 1, c1 - 1) : c;
            T const* xn = xs + int64_t(cn) * ld;
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int j = 0; j < BS; ++j) w[b][j] = xn[off[b] + j];
            if constexpr (!PRE) {
#pragma unroll
                for (int b = 0; b < NB; ++b)
#pragma unroll
                    for (int j = 0; j < BS; ++j) v[b][j] = w[b][j];
            }
            T acc[R];
#pragma unroll
            for (int r = 0; r < R; ++r) acc[r] = T(0);
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int j = 0; j < BS; ++j)
#pragma unroll
                    for (int r = 0; r < R; ++r) acc[r] += m[b][r][j] * v[b][j];
            if (act) {
                T* y = ys + int64_t(c) * ld + ec * BS + part * R;
#pragma unroll
                for (int r = 0; r < R; ++r) y[r] = acc[r];
            }
            if constexpr (DOT) {
                double d = 0.0;
                if (act) {
#pragma unroll
                    for (int r = 0; r < R; ++r) d += double(v[0][r]) * double(acc[r]); // R == BS: part 0
                }
                for (int o = 32; o > 0; o >>= 1) d += __shfl_down(d, o, 64);
                if (lane == 0) red[wave][c - c0] += d;
            }
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int j = 0; j < BS; ++j) v[b][j] = w[b][j];
        }
        }
    };
    if constexpr (DOT) {
        const int64_t nblk = (E + EPB - 1) / EPB;
        for (int64_t blk = blockIdx.x; blk < nblk; blk += gridDim.x) block(blk);
        __syncthreads();
        if (threadIdx.x < c1 - c0)
            partial[int64_t(c0 + threadIdx.x) * kDotParts + blockIdx.x] =
                ((red[0][threadIdx.x] + red[1][threadIdx.x]) + red[2][threadIdx.x]) + red[3][threadIdx.x];
    } else {
        block(blockIdx.x);
    }
}

// One lane per (row, element), element fastest (coalesced block-row reads), one column per grid row.
template <class T>
__global__ __launch_bounds__(kThreads) void assembled_apply_generic(AssembledArgs<T> a, T const* __restrict__ xs,
                                                                    T* __restrict__ ys, int64_t ld) {
    const int64_t E = a.num_elements;
    const int bs = a.block_size;
    const int64_t t = int64_t(blockIdx.x) * kThreads + threadIdx.x;
    if (t >= E * bs) return;
    const int i = int(t / E);
    const int64_t e = t - int64_t(i) * E;
    T const* x = xs + int64_t(blockIdx.y) * ld;
    T acc = T(0);
    for (int b = 0; b < a.slots; ++b) {
        const int64_t o = (b == 0 ? e : int64_t(a.nb[int64_t(b - 1) * E + e])) * bs;
        T const* Ab = a.A + (int64_t(b) * bs + i) * bs * E + e;
        for (int j = 0; j < bs; ++j) acc += Ab[int64_t(j) * E] * x[o + j];
    }
    ys[int64_t(blockIdx.y) * ld + e * bs + i] = acc;
}

template <class T, int BS, int NB, int R>
hipError_t run_reg(AssembledArgs<T> const& a, T const* x, T* y, int ncols, int64_t ld, double* partial,
                   hipStream_t s) {
    constexpr int EPB = 64 * kWavesPerBlock;
    if (partial) {
        // one element per lane only: with several lanes per element the fused sums cost registers
        if constexpr (R == BS) {
            const int chunks = (ncols + kDotCols - 1) / kDotCols;
            hipLaunchKernelGGL((assembled_apply_kernel<T, BS, NB, R, true>), dim3(kDotParts, unsigned(chunks)),
                               dim3(kThreads), 0, s, a, x, y, ld, ncols, kDotCols, partial);
            return hipGetLastError();
        } else {
            return hipErrorNotSupported;
        }
    }
    const int64_t gx = (a.num_elements + EPB - 1) / EPB;
    // column ranges: enough waves to fill the chip (about 4 per SIMD), each range re-reads the matrix
    const int64_t waves = gx * kWavesPerBlock;
    int chunks = int(std::min<int64_t>(ncols, std::max<int64_t>(1, (4096 + waves - 1) / waves)));
    const int cchunk = (ncols + chunks - 1) / chunks;
    chunks = (ncols + cchunk - 1) / cchunk;
    hipLaunchKernelGGL((assembled_apply_kernel<T, BS, NB, R, false>), dim3(unsigned(gx), unsigned(chunks)),
                       dim3(kThreads), 0, s, a, x, y, ld, ncols, cchunk, static_cast<double*>(nullptr));
    return hipGetLastError();
}

// Rows per lane of the register kernel: the largest divisor of BS whose block rows fit in about 160
// (FP32) / 128 (FP64) 32-bit registers.
template <class T, int BS, int NB> constexpr int rows_per_lane() {
    constexpr int budget = sizeof(T) == 8 ? 128 : 160;
    int best = 1;
    for (int r = 1; r <= BS; ++r)
        if (BS % r == 0 && NB * r * BS * int(sizeof(T) / 4) <= budget) best = r;
    return best;
}

template <class T, int BS, int NB>
hipError_t run_bs(AssembledArgs<T> const& a, T const* x, T* y, int ncols, int64_t ld, double* partial,
                  hipStream_t s) {
    return run_reg<T, BS, NB, rows_per_lane<T, BS, NB>()>(a, x, y, ncols, ld, partial, s);
}

} // namespace

bool assembled_register_kernel(int block_size, int slots) {
    return (slots == 4 && (block_size == 3 || block_size == 6 || block_size == 12)) ||
           (slots == 5 && (block_size == 4 || block_size == 12));
}

template <class T>
hipError_t launch_assembled(AssembledArgs<T> const& a, T const* x, T* y, int ncols, int64_t ld, hipStream_t s,
                            double* partial) {
    if (a.num_elements <= 0 || ncols <= 0) return partial ? hipErrorInvalidValue : hipSuccess;
    if (ncols > 1 && ld < a.num_elements * a.block_size) return hipErrorInvalidValue;
    if (ncols > 65535 * kDotCols) return hipErrorInvalidValue;
    const int bs = a.block_size;
    if (a.slots == 4) {
        if (bs == 3) return run_bs<T, 3, 4>(a, x, y, ncols, ld, partial, s);
        if (bs == 6) return run_bs<T, 6, 4>(a, x, y, ncols, ld, partial, s);
        if (bs == 12) return run_bs<T, 12, 4>(a, x, y, ncols, ld, partial, s);
    } else if (a.slots == 5) {
        if (bs == 4) return run_bs<T, 4, 5>(a, x, y, ncols, ld, partial, s);
        if (bs == 12) return run_bs<T, 12, 5>(a, x, y, ncols, ld, partial, s);
    }
    if (partial) return hipErrorNotSupported; // the generic kernel computes no partial sums
    const int64_t rows = a.num_elements * bs;
    hipLaunchKernelGGL((assembled_apply_generic<T>), dim3(unsigned((rows + kThreads - 1) / kThreads), unsigned(ncols)),
                       dim3(kThreads), 0, s, a, x, y, ld);
    return hipGetLastError();
}

template hipError_t launch_assembled<double>(AssembledArgs<double> const&, double const*, double*, int, int64_t,
                                             hipStream_t, double*);
template hipError_t launch_assembled<float>(AssembledArgs<float> const&, float const*, float*, int, int64_t,
                                            hipStream_t, double*);

// ------------------------------------------------------------------------------ assembly by probing
namespace {

// X[j][e bs + i] = 1 if element e has colour c and i = c0 + j, else 0 (j < m)
__global__ __launch_bounds__(kThreads) void probe_kernel(double* __restrict__ X, int32_t const* __restrict__ colour,
                                                         int c, int c0, int m, int64_t n, int bs) {
    const int64_t t = int64_t(blockIdx.x) * kThreads + threadIdx.x;
    if (t >= int64_t(m) * n) return;
    const int j = int(t / n);
    const int64_t k = t - int64_t(j) * n;
    const int64_t e = k / bs;
    const int i = int(k - e * bs);
    X[t] = (colour[e] == c && i == c0 + j) ? 1.0 : 0.0;
}

// Column c0 + j of the blocks that colour c's probe isolates: y on element e is column c0 + j of e's
// own block when e has colour c, of the block of e's face-f neighbour when that neighbour has it.
__global__ __launch_bounds__(kThreads) void gather_kernel(double const* __restrict__ Y,
                                                          int32_t const* __restrict__ colour,
                                                          int32_t const* __restrict__ nb, double* __restrict__ A,
                                                          int c, int c0, int m, int64_t E, int bs, int faces) {
    const int64_t n = E * bs;
    const int64_t t = int64_t(blockIdx.x) * kThreads + threadIdx.x;
    if (t >= int64_t(m) * n) return;
    const int j = int(t / n);
    const int64_t k = t - int64_t(j) * n;
    const int64_t e = k / bs;
    const int i = int(k - e * bs);
    const int col = c0 + j;
    const double v = Y[t];
    if (colour[e] == c) A[((int64_t(0) * bs + i) * bs + col) * E + e] = v;
    for (int f = 0; f < faces; ++f) {
        const int64_t o = nb[int64_t(f) * E + e];
        if (o != e && colour[o] == c) A[((int64_t(1 + f) * bs + i) * bs + col) * E + e] = v;
    }
}

__global__ __launch_bounds__(kThreads) void narrow_kernel(double const* __restrict__ a, float* __restrict__ b,
                                                          int64_t n) {
    const int64_t t = int64_t(blockIdx.x) * kThreads + threadIdx.x;
    if (t < n) b[t] = float(a[t]);
}

unsigned grid_of(int64_t n) { return unsigned((n + kThreads - 1) / kThreads); }

} // namespace

hipError_t launch_assembly_probe(double* X, int32_t const* colour, int c, int c0, int m, int64_t E, int bs,
                                 hipStream_t s) {
    const int64_t n = E * bs;
    hipLaunchKernelGGL(probe_kernel, dim3(grid_of(int64_t(m) * n)), dim3(kThreads), 0, s, X, colour, c, c0, m, n, bs);
    return hipGetLastError();
}

hipError_t launch_assembly_gather(double const* Y, int32_t const* colour, int32_t const* nb, double* A, int c, int c0,
                                  int m, int64_t E, int bs, int faces, hipStream_t s) {
    hipLaunchKernelGGL(gather_kernel, dim3(grid_of(int64_t(m) * E * bs)), dim3(kThreads), 0, s, Y, colour, nb, A, c,
                       c0, m, E, bs, faces);
    return hipGetLastError();
}

hipError_t launch_narrow(double const* a, float* b, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(narrow_kernel, dim3(grid_of(n)), dim3(kThreads), 0, s, a, b, n);
    return hipGetLastError();
}

} // namespace tdg
