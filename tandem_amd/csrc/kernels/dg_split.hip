// Dispatch of the two-kernel matrix-core elasticity apply over (dimension, degree).
#include "dg_split_kernels.h"

namespace tdg {

#define TDG_SPLIT_DECL(D, N)                                                                                    \
    template <class T>                                                                                         \
    hipError_t launch_split_d##D##p##N(int shape, SplitArgs<T> const& a, T const* x, T* y, T* L, int ncols,  \
                                       int64_t ld, int64_t ldL, hipStream_t s);
TDG_SPLIT_DECL(3, 2)
TDG_SPLIT_DECL(3, 3)
TDG_SPLIT_DECL(3, 4)
TDG_SPLIT_DECL(2, 2)
TDG_SPLIT_DECL(2, 3)
TDG_SPLIT_DECL(2, 4)
TDG_SPLIT_DECL(2, 6)
#undef TDG_SPLIT_DECL

template <class T>
hipError_t launch_elasticity_split(int dim, int degree, int shape, SplitArgs<T> const& a, T const* x, T* y, T* L,
                                   int ncols, int64_t ld, int64_t ldL, hipStream_t s) {
    if (shape < 0 || shape >= kSplitShapes) return hipErrorInvalidValue;
    if (dim == 3) {
        switch (degree) {
        case 2: return launch_split_d3p2<T>(shape, a, x, y, L, ncols, ld, ldL, s);
        case 3: return launch_split_d3p3<T>(shape, a, x, y, L, ncols, ld, ldL, s);
        case 4: return launch_split_d3p4<T>(shape, a, x, y, L, ncols, ld, ldL, s);
        }
    } else if (dim == 2) {
        switch (degree) {
        case 2: return launch_split_d2p2<T>(shape, a, x, y, L, ncols, ld, ldL, s);
        case 3: return launch_split_d2p3<T>(shape, a, x, y, L, ncols, ld, ldL, s);
        case 4: return launch_split_d2p4<T>(shape, a, x, y, L, ncols, ld, ldL, s);
        case 6: return launch_split_d2p6<T>(shape, a, x, y, L, ncols, ld, ldL, s);
        }
    }
    return hipErrorInvalidValue;
}

template <class T> int64_t split_element_workgroups(int shape, int64_t num_groups) {
    return split::element_workgroups_of<T>(shape, num_groups);
}
template int64_t split_element_workgroups<double>(int, int64_t);
template int64_t split_element_workgroups<float>(int, int64_t);

namespace {
template <class T> __global__ void zero_slots_kernel(T* L, int32_t const* slots, int64_t n, int bs) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n * bs) L[int64_t(slots[i / bs]) * bs + i % bs] = T(0);
}
} // namespace

template <class T> hipError_t launch_zero_slots(T* L, int32_t const* slots, int64_t nslots, int bs, hipStream_t s) {
    if (nslots <= 0) return hipSuccess;
    const int64_t n = nslots * bs;
    const int64_t blocks = (n + 255) / 256;
    if (blocks > int64_t(0x7fffffff)) return hipErrorInvalidValue;
    hipLaunchKernelGGL((zero_slots_kernel<T>), dim3(unsigned(blocks)), dim3(256), 0, s, L, slots, nslots, bs);
    return hipGetLastError();
}
template hipError_t launch_zero_slots<double>(double*, int32_t const*, int64_t, int, hipStream_t);
template hipError_t launch_zero_slots<float>(float*, int32_t const*, int64_t, int, hipStream_t);

template hipError_t launch_elasticity_split<double>(int, int, int, SplitArgs<double> const&, double const*, double*,
                                                    double*, int, int64_t, int64_t, hipStream_t);
template hipError_t launch_elasticity_split<float>(int, int, int, SplitArgs<float> const&, float const*, float*, float*,
                                                   int, int64_t, int64_t, hipStream_t);

} // namespace tdg
