// build_pca_space without allocation: gcam_build_pca_space_noalloc runs the
// kernels of gcam_build_pca_space (patch moments on the FP64 matrix cores,
// their fixed-order reduction, the leading-eigenpair solve) with the moment
// partials in a file-scope device array instead of a hipMallocAsync'd
// buffer, for callers that cannot allocate device memory.
//
// The array is one per GPU (each device's copy of the code object holds its
// own) and lives until the process ends.  Its device address is taken once
// per device with hipGetSymbolAddress, on the first call made with that
// device current; host code only ever passes that address, never the
// array's name (which on the host is a stand-in, not a device pointer).
// Every call on one GPU shares the array: calls that overlap on one GPU
// (different streams, or different host threads) race on it.  Nothing in it
// has to start at zero: the moment kernel writes every partial it uses and
// the reduction reads only those.
#include <mutex>
#include <string>

#include "gcam_internal.hpp"

namespace gcam {
namespace {

// Up to kMomentBlocks workgroup partials of the 350 moments and their sum:
// 1025 x 350 doubles = 2.87 MB per GPU.
constexpr size_t kScratchDoubles = (size_t)kMomentSlots * kMoments;
static_assert(kScratchDoubles * sizeof(double) == GCAM_PCA_NOALLOC_SCRATCH_BYTES,
              "the header states the scratch size");

__device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// The device address of this GPU's scratch array, looked up on the first
// call with `dev` current.
gcam_status scratch_for(int dev, double** out) {
    if (dev < 0 || dev >= kMaxDevices)
        return set_error(GCAM_ERR_UNSUPPORTED, "device index " + std::to_string(dev) + " beyond " +
                                                   std::to_string(kMaxDevices - 1));
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        double* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(reinterpret_cast<void**>(&p), HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_error(GCAM_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return GCAM_OK;
}

}  // namespace
}  // namespace gcam

using namespace gcam;

extern "C" gcam_status gcam_build_pca_space_noalloc(const float* patches, int64_t n, float* basis,
                                                   double* eigenvalues, void* stream) {
    if (!patches || !basis) return set_error(GCAM_ERR_INVALID_ARGUMENT, "NULL pointer");
    if (n < 2) return set_error(GCAM_ERR_INVALID_ARGUMENT, "build_pca_space needs at least 2 patches");
    if (n > GCAM_PCA_NOALLOC_MAX_PATCHES)
        return set_error(GCAM_ERR_INVALID_ARGUMENT, "gcam_build_pca_space_noalloc takes at most " +
                                                        std::to_string((long long)GCAM_PCA_NOALLOC_MAX_PATCHES) +
                                                        " patches, got " + std::to_string((long long)n));
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return set_error(GCAM_ERR_NO_DEVICE, "no HIP device");
    double* partials = nullptr;
    gcam_status st = scratch_for(dev, &partials);
    if (st != GCAM_OK) return st;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    int nblocks = 0;
    st = launch_buffer_moments(patches, n, partials, &nblocks, s);
    if (st != GCAM_OK) return st;
    return launch_pca_solve(partials, nblocks, n, basis, eigenvalues, nullptr, nullptr, s);
}
