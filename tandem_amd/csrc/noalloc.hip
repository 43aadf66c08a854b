// tdg_mult_noalloc: the FP64 apply with its facet slots in a fixed device array, for callers that
// cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
// forbid allocation).  No counterpart in the reference.
//
// The slots live in `scratch`, a file-scope device array of kScratchDoubles entries (64 MiB) that
// every GPU the library runs on holds until the process ends.  Host code never uses the array's
// name as an address: on each device's first call its address is looked up once with
// hipGetSymbolAddress and kept per device, and the kernels receive that pointer.  All calls on one
// GPU share the array: calls that overlap in time on one GPU (different streams, threads or
// operators) race on it; calls ordered on one stream do not.  Each call clears the slots of faces
// that carry no facet term, so the previous call's contents do no harm.
#include "abi.h"

#include <hip/hip_runtime.h>

#include <cstddef>
#include <mutex>
#include <string>

namespace {

constexpr std::size_t kScratchDoubles = std::size_t(1) << 23; // 64 MiB
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kScratchDoubles];

std::mutex g_mutex;
double* g_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device (looked up on its first use there).
double* scratch_on_current_device() {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device index out of range");
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!g_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw std::runtime_error(std::string("HIP error: scratch address lookup failed: ") + hipGetErrorString(e));
        g_ptr[dev] = static_cast<double*>(p);
    }
    return g_ptr[dev];
}

} // namespace

namespace tdg_abi {
// Set-up of the current device, run by tdg_upload: the address lookup happens there, outside any
// region the caller may be capturing.
void prepare_noalloc_scratch() { (void)scratch_on_current_device(); }
} // namespace tdg_abi

using namespace tdg_abi;

extern "C" int tdg_mult_noalloc(tdg_op* op, double const* x, double* y) {
    return guarded([&] {
        auto& o = ready(op);
        require(x && y && x != y, "x and y must be distinct device buffers");
        o.apply_noalloc(x, y, scratch_on_current_device(), kScratchDoubles);
    });
}
