// Instantiation of the two-kernel matrix-core elasticity apply for dimension 2, degree 3.
#include "dg_split_kernels.h"

namespace tdg {

template <class T>
hipError_t launch_split_d2p3(int shape, SplitArgs<T> const& a, T const* x, T* y, T* L, int ncols, int64_t ld,
                             int64_t ldL, hipStream_t s) {
    return split::run_split<2, 3, T>(shape, a, x, y, L, ncols, ld, ldL, s);
}

template hipError_t launch_split_d2p3<double>(int, SplitArgs<double> const&, double const*, double*, double*, int,
                                               int64_t, int64_t, hipStream_t);
template hipError_t launch_split_d2p3<float>(int, SplitArgs<float> const&, float const*, float*, float*, int, int64_t,
                                              int64_t, hipStream_t);

} // namespace tdg
