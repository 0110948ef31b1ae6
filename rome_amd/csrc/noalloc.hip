// Allocation-free projection: rome_project_device_noalloc stages the caller's rotation matrices in a
// file-scope device array instead of a buffer from hipMalloc, for callers that cannot allocate (code
// under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
//
// `scratch` holds [kNoallocMaxRot][9] row-major matrices followed by [kNoallocMaxRot][6] projector rows
// (2 A^T restricted to the x, y columns, as the projector reads them).  Its host-side name is not a
// device address: host code uses only the per-device address hipGetSymbolAddress returns.  Every call
// on one device shares that device's array; two calls that overlap on one device race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "kernels.h"

namespace rome {

namespace {

constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[(size_t)kNoallocMaxRot * 15];

double* g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mu;

// Ai[r] = 2 (A00, A10, A01, A11, A02, A12): the same products (exact doublings) as the host-side table
__global__ void k_rot6(const double* __restrict__ A9, int nrot, double* __restrict__ Ai) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrot) return;
    const double* A = A9 + 9 * (size_t)r;
    double* o = Ai + 6 * (size_t)r;
    o[0] = 2.0 * A[0];
    o[1] = 2.0 * A[3];
    o[2] = 2.0 * A[1];
    o[3] = 2.0 * A[4];
    o[4] = 2.0 * A[2];
    o[5] = 2.0 * A[5];
}

}  // namespace

hipError_t noalloc_init(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch[device]) return hipSuccess;
    int cur = 0;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != device && (e = hipSetDevice(device)) != hipSuccess) return e;
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e == hipSuccess) g_scratch[device] = static_cast<double*>(p);
    if (cur != device) (void)hipSetDevice(cur);
    return e;
}

hipError_t project_plane_noalloc(hipStream_t s, int device, const double* A9_host, int nrot, const float2* grid,
                                 int pad, int padF, int c, int is3d, int size, int my_r_max, float2* out) {
    if (nrot < 0 || nrot > kNoallocMaxRot) return hipErrorInvalidValue;
    if (nrot == 0) return hipSuccess;
    if (device < 0 || device >= kMaxDevices || !g_scratch[device]) return hipErrorNotInitialized;
    double* A9 = g_scratch[device];
    double* Ai = A9 + 9 * (size_t)kNoallocMaxRot;
    hipError_t e = hipMemcpyAsync(A9, A9_host, 9 * (size_t)nrot * sizeof(double), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
    k_rot6<<<(nrot + 255) / 256, 256, 0, s>>>(A9, nrot, Ai);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    return launch_project_plane(s, grid, pad, padF, c, is3d, Ai, nrot, size, my_r_max, out);
}

}  // namespace rome
