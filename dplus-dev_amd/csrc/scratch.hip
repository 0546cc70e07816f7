// scratch.hip -- device memory of the allocation-free orientation average.
//
// dplus_orientation_average_noalloc takes every device buffer it needs (the
// docked-leaf tables, q and shell groups, Monte Carlo block partials, VEGAS bin
// sums and the intensity history) from one file-scope device array instead of
// allocating.  Each GPU the library runs on holds its own copy of the array
// until the process ends; its device address is looked up once per device, with
// that device current, and only that address is handed to kernels and copies
// (the array's name in host code is a host-side stand-in, not a device address).
#include <mutex>

#include "device.hpp"

namespace dplus {
constexpr size_t kScratchBytes = (size_t)64 << 20;
}

static __device__ __align__(256) unsigned char scratch[dplus::kScratchBytes];

namespace dplus {

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_use[kMaxDevices];

__global__ __launch_bounds__(256) void fill_zero_kernel(double *__restrict__ p, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}
}  // namespace

size_t scratch_capacity() { return kScratchBytes; }

// the device address of `device`'s scratch array; the caller has made `device`
// current (Device::require)
void *scratch_address(int device) {
    if (device < 0 || device >= kMaxDevices) throw Error(kErrInvalidArgs, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        int cur = -1;
        check_hip(hipGetDevice(&cur), "hipGetDevice");
        if (cur != device) throw Error(kErrGeneral, "scratch set-up with another device current");
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            (void)hipGetLastError();
            throw Error(kErrGeneral, std::string("hipGetSymbolAddress of the scratch array failed (") +
                                         hipGetErrorName(e) + ")");
        }
        g_scratch_ptr[device] = ptr;
    }
    return g_scratch_ptr[device];
}

// held for the whole of one allocation-free call on `device`: calls on one GPU
// share its array, so overlapping calls from several threads take turns
std::mutex &scratch_lock(int device) { return g_scratch_use[device]; }

// zero n doubles with a kernel (a fill node, unlike a memset node, is safe in
// captured graphs)
void launch_fill_zero(double *p, size_t n, hipStream_t s) {
    if (n == 0) return;
    const unsigned nb = (unsigned)((n + 255) / 256);
    fill_zero_kernel<<<nb, 256, 0, s>>>(p, (long long)n);
}

}  // namespace dplus
