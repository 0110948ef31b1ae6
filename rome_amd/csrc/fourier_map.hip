// Projector grid of a reference map on the GPU: MyProjector::computeFourierTransformMap
// (src/reconstruct.cpp:106-223) with griddingCorrect (src/reconstruct.cpp:61-103).
//
//   pad    : map / sinc^2(r / (ori * pf)) (trilinear interpolator, r > 0), zero-padded to pf*ori per axis
//            at the centred position and circularly shifted by pf*ori/2 (centerFFT3D/2D forward,
//            src/reconstruct.h:103-180): one thread per map voxel, FP64.
//   FFT    : hipFFT D2Z (FP64), then / (number of real samples) as normalizeComplex does
//            (src/fft_base.cpp:103-109).
//   window : coefficients with |k|^2 <= (pf r_max)^2, times pf^3 ori (3D) or pf^2 (2D), into the
//            [pad][pad][pad/2+1] grid (FP64 copy for the caller, FP32 copy for the projector), and the
//            shell power sum / count (block sums in LDS, one FP64 atomic per shell per block).
#include <hip/hip_runtime.h>

#include "device_util.h"
#include "kernels.h"

namespace rome {

namespace {

__global__ void k_fm_pad(const double* __restrict__ vol, int ori, int padori, int dim3, int do_gridding, int pf,
                         double* __restrict__ Mc, double* __restrict__ corrected) {
    const size_t n = dim3 ? (size_t)ori * ori * ori : (size_t)ori * ori;
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n) return;
    const int j = (int)(idx % ori);
    const int i = (int)((idx / ori) % ori);
    const int k = dim3 ? (int)(idx / ((size_t)ori * ori)) : 0;
    const int sh = -(int)((float)ori / 2.0f);
    double v = vol[idx];
    if (do_gridding) {
        const int x = j + sh, y = i + sh, z = dim3 ? k + sh : 0;
        const double r = sqrt((double)(x * x + y * y + z * z));
        if (r > 0.) {
            const double rval = r / (double)(ori * pf);
            const double sinc = sin(3.14159265358979323846 * rval) / (3.14159265358979323846 * rval);
            v /= sinc * sinc;
        }
    }
    if (corrected) corrected[idx] = v;
    // centred padding offset, then the forward centring shift by padori/2
    const int off = sh + (int)((float)padori / 2.0f);
    const int s = padori / 2;
    const int jp = (j + off + s) % padori, ip = (i + off + s) % padori;
    if (dim3) {
        const int kp = (k + off + s) % padori;
        Mc[((size_t)kp * padori + ip) * padori + jp] = v;
    } else {
        Mc[(size_t)ip * padori + jp] = v;
    }
}

constexpr int kFmShells = 520;  // ori/2 + 1 for ori up to 1038

__global__ __launch_bounds__(256) void k_fm_window(const double2* __restrict__ F, int padori, int dim3, int pf,
                                                   int r_max, int pad, int oriF, double size3, double normfft,
                                                   double2* __restrict__ grid64, float2* __restrict__ grid32,
                                                   double* __restrict__ pw, double* __restrict__ cnt) {
    __shared__ double s_pw[kFmShells], s_cnt[kFmShells];
    for (int q = threadIdx.x; q < oriF; q += blockDim.x) s_pw[q] = s_cnt[q] = 0.;
    __syncthreads();
    const int padoriF = padori / 2 + 1;
    const size_t n = (dim3 ? (size_t)padori : 1) * padori * padoriF;
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < n) {
        const int j = (int)(idx % padoriF);
        const int i = (int)((idx / padoriF) % padori);
        const int k = dim3 ? (int)(idx / ((size_t)padoriF * padori)) : 0;
        const int kp = dim3 ? (k < padoriF ? k : k - padori) : 0;
        const int ip = i < padoriF ? i : i - padori;
        const int jp = j;
        const int r2 = ip * ip + jp * jp + kp * kp;
        const int max_r2 = r_max * r_max * pf * pf;
        if (r2 <= max_r2) {
            const int c = (int)((float)pad / 2.0f);
            const int z = dim3 ? kp + c : 0, y = ip + c, x = jp;
            const double2 f = F[idx];
            const double re = (f.x / size3) * normfft, im = (f.y / size3) * normfft;
            const size_t g = dim3 ? ((size_t)z * pad + y) * (pad / 2 + 1) + x : (size_t)y * (pad / 2 + 1) + x;
            grid64[g] = make_double2(re, im);
            grid32[g] = make_float2((float)re, (float)im);
            const int ires = (int)round(sqrt((double)r2) / pf);
            atomicAdd(&s_pw[ires], (re * re + im * im) / 2.);
            atomicAdd(&s_cnt[ires], 1.);
        }
    }
    __syncthreads();
    for (int q = threadIdx.x; q < oriF; q += blockDim.x)
        if (s_cnt[q] != 0.) {
            atomicAdd(&pw[q], s_pw[q]);
            atomicAdd(&cnt[q], s_cnt[q]);
        }
}

}  // namespace

hipError_t launch_fm_pad(hipStream_t s, const double* vol, int ori, int padori, int dim3, int do_gridding, int pf,
                         double* Mc, double* corrected) {
    const size_t n = dim3 ? (size_t)ori * ori * ori : (size_t)ori * ori;
    k_fm_pad<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(vol, ori, padori, dim3, do_gridding, pf, Mc, corrected);
    return hipGetLastError();
}

hipError_t launch_fm_window(hipStream_t s, const double2* F, int padori, int dim3, int pf, int r_max, int pad,
                            int oriF, double normfft, double2* grid64, float2* grid32, double* pw, double* cnt) {
    if (oriF > kFmShells) return hipErrorInvalidValue;
    const size_t n = (dim3 ? (size_t)padori : 1) * padori * (padori / 2 + 1);
    const double size3 = dim3 ? (double)padori * padori * padori : (double)padori * padori;
    k_fm_window<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(F, padori, dim3, pf, r_max, pad, oriF, size3, normfft,
                                                           grid64, grid32, pw, cnt);
    return hipGetLastError();
}

}  // namespace rome
