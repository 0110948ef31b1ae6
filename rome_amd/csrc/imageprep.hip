// Image preparation of a pool on the device (ParticleModel::setFourierTransforms,
// src/exp_model.h:397-639), FP64 like the reference:
//   k_ip_shift  : norm correction (x avg_norm / NORM), integer translation by the rounded old offsets
//                 without wrapping (translate / applyGeometry, src/image.h:53-257), then CenterFFT
//                 (src/image.h:277-343) into the FFT input; the translated image is kept for the mask.
//   k_ip_mask   : softMaskOutsideMap (src/image.h:571-643) of the translated image, raised cosine over
//                 cosine_width pixels outside the radius, filled with the given noise image or the
//                 weighted background average; CenterFFT into the second FFT input.
//   hipFFT D2Z  : both transforms of the pool (batched); normalised by 1/ori^2 in the window kernel
//                 (FourierTransform with normalize = true).
//   k_ip_window : windowFourierTransform to current_size (src/image.h:496-566) of both, FP64 to the
//                 caller and FP32 for the engine.
//   k_ip_power  : power spectrum per rounded radius and highres_Xi2 of the masked transform
//                 (src/exp_model.h:569-604), when current_size < ori_size.
//   k_ip_ctf    : CTF::getFftwImage on the current_size grid (src/ctf.cpp:136-210, src/ctf.h:151-205).
#include <hip/hip_runtime.h>

#include "device_util.h"
#include "kernels.h"

namespace rome {

namespace {

constexpr double kPi = 3.14159265358979323846;

// C round(): half away from zero
DEV int c_round(double x) { return x >= 0. ? (int)floor(x + 0.5) : -(int)floor(-x + 0.5); }

__global__ void k_ip_shift(ImagePrepArgs a) {
    const int img = blockIdx.y;
    const int ori = a.ori, n2 = ori * ori;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n2) return;
    const int i = q / ori, j = q - i * ori;
    const int ox = c_round(a.xoff[img]), oy = c_round(a.yoff[img]);
    if (q == 0) {
        a.old_offx[img] = (double)ox;
        a.old_offy[img] = (double)oy;
    }
    const int si = i - oy, sj = j - ox;
    const double* in = a.images + (size_t)img * n2;
    // image * avg_norm / NORM as the reference multiplies it (one product, then the division)
    const double v = (si >= 0 && si < ori && sj >= 0 && sj < ori)
                         ? (a.do_norm ? in[si * ori + sj] * a.avg_norm / a.norm[img] : in[si * ori + sj])
                         : 0.0;
    a.timg[(size_t)img * n2 + q] = v;
    const int s = ori / 2;
    const int ci = (i + s) % ori, cj = (j + s) % ori;
    a.cbuf[(size_t)img * n2 + ci * ori + cj] = v;
}

// One block per image: background average (when no noise image), soft mask, CenterFFT.
__global__ __launch_bounds__(256) void k_ip_mask(ImagePrepArgs a) {
    __shared__ double red[16];
    const int img = blockIdx.x, tid = threadIdx.x;
    const int ori = a.ori, n2 = ori * ori;
    const double* t = a.timg + (size_t)img * n2;
    const double* noise = a.noise ? a.noise + (size_t)img * n2 : nullptr;
    const double radius = a.radius, width = a.cosine_width, radius_p = radius + width;
    const int o = -(int)((float)ori / 2.0f);
    double bg = 0.;
    if (!noise) {
        double sum = 0., sbg = 0.;
        for (int q = tid; q < n2; q += blockDim.x) {
            const int i = q / ori, j = q - i * ori;
            const double r = sqrt((double)((i + o) * (i + o) + (j + o) * (j + o)));
            if (r < radius) continue;
            if (r > radius_p) {
                sum += 1.;
                sbg += t[q];
            } else {
                const double rc = 0.5 + 0.5 * cos(kPi * (radius_p - r) / width);
                sum += rc;
                sbg += rc * t[q];
            }
        }
        sum = block_reduce(sum, [](double x, double y) { return x + y; }, red);
        sbg = block_reduce(sbg, [](double x, double y) { return x + y; }, red);
        bg = sbg / sum;
    }
    const int s = ori / 2;
    double* out = a.cbuf + (size_t)img * n2;
    for (int q = tid; q < n2; q += blockDim.x) {
        const int i = q / ori, j = q - i * ori;
        const double r = sqrt((double)((i + o) * (i + o) + (j + o) * (j + o)));
        double v = t[q];
        if (r >= radius) {
            const double add = noise ? noise[q] : bg;
            if (r <= radius_p) {
                const double rc = 0.5 + 0.5 * cos(kPi * (radius_p - r) / width);
                v = (1 - rc) * v + rc * add;
            } else {
                v = add;
            }
        }
        out[((i + s) % ori) * ori + (j + s) % ori] = v;
    }
}

// Window of both transforms to current_size: element (i, j) of the cur x (cur/2+1) output.
__global__ void k_ip_window(ImagePrepArgs a) {
    const int img = blockIdx.y;
    const int cur = a.cur, curF = cur / 2 + 1, ori = a.ori, oriF = ori / 2 + 1;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= cur * curF) return;
    const int i = q / curF, j = q - i * curF;
    double2 vn = make_double2(0., 0.), vm = make_double2(0., 0.);
    if (cur == ori) {
        vn = a.F1[(size_t)img * ori * oriF + q];
        vm = a.F2[(size_t)img * ori * oriF + q];
    } else {  // shrinking branch (current_size <= ori_size): the whole low-frequency rectangle
        const int ip = i < curF ? i : i - cur;
        const int in_i = ip < 0 ? ip + ori : ip;
        vn = a.F1[((size_t)img * ori + in_i) * oriF + j];
        vm = a.F2[((size_t)img * ori + in_i) * oriF + j];
    }
    const double inv = 1.0 / ((double)ori * (double)ori);
    vn.x *= inv;
    vn.y *= inv;
    vm.x *= inv;
    vm.y *= inv;
    const size_t o = (size_t)img * cur * curF + q;
    if (a.fnomask) a.fnomask[o] = vn;
    if (a.fmask) a.fmask[o] = vm;
    if (a.fnomask32) a.fnomask32[o] = make_float2((float)vn.x, (float)vn.y);
    if (a.fmask32) a.fmask32[o] = make_float2((float)vm.x, (float)vm.y);
}

// One block per image: shells of the masked transform (x = 0 column for ip >= 0 only).
__global__ __launch_bounds__(256) void k_ip_power(ImagePrepArgs a) {
    __shared__ double shell[kMaxShell];
    __shared__ double red[16];
    const int img = blockIdx.x, tid = threadIdx.x;
    const int ori = a.ori, oriF = ori / 2 + 1, curF = a.cur / 2 + 1;
    for (int r = tid; r < oriF; r += blockDim.x) shell[r] = 0.;
    __syncthreads();
    const double inv = 1.0 / ((double)ori * (double)ori);
    double hx = 0.;
    for (int q = tid; q < ori * oriF; q += blockDim.x) {
        const int i = q / oriF, jp = q - i * oriF;
        const int ip = i < oriF ? i : i - ori;
        const int ires = c_round(sqrt((double)(ip * ip + jp * jp)));
        if (ires > 0 && ires < oriF && !(jp == 0 && ip < 0)) {
            double2 f = a.F2[(size_t)img * ori * oriF + q];
            f.x *= inv;
            f.y *= inv;
            const double v = f.x * f.x + f.y * f.y;
            atomicAdd(&shell[ires], v);
            if (ires >= curF) hx += v;
        }
    }
    hx = block_reduce(hx, [](double x, double y) { return x + y; }, red);
    __syncthreads();
    for (int r = tid; r < oriF; r += blockDim.x) a.power[(size_t)img * oriF + r] = shell[r];
    if (tid == 0) a.hxi2[img] = hx;
}

__global__ void k_ip_ctf(ImagePrepArgs a) {
    const int img = blockIdx.y;
    const int cur = a.cur, h = cur / 2 + 1;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= cur * h) return;
    const int i = q / h, j = q - i * h;
    double v = 1.0;
    if (a.ctf) {
        const double* c = a.ctf + 7 * (size_t)img;  // defU, defV, angle, voltage, Cs, Q0, Bfac
        const double local_Cs = c[4] * 1e7, local_kV = c[3] * 1e3;
        const double rad_az = c[2] * kPi / 180.0;
        const double dfa = -(c[0] + c[1]) * 0.5, dfd = -(c[0] - c[1]) * 0.5;
        const double lambda = 12.2643247 / sqrt(local_kV * (1. + local_kV * 0.978466e-6));
        const double K1 = kPi / 2 * 2 * lambda;
        const double K2 = kPi / 2 * local_Cs * lambda * lambda * lambda;
        const double K3 = sqrt(1 - c[5] * c[5]);
        const double K4 = -c[6] / 4.;
        const double xs = (double)a.ori * a.angpix, ys = (double)a.ori * a.angpix;
        const int ip = i < h ? i : i - cur;
        const double X = (double)j / xs, Y = (double)ip / ys;
        const double u2 = X * X + Y * Y, u4 = u2 * u2;
        double deltaf = 0.;
        if (!(fabs(X) < 1e-6 && fabs(Y) < 1e-6)) deltaf = dfa + dfd * cos(2 * (atan2(Y, X) - rad_az));
        const double arg = K1 * deltaf * u2 + K2 * u4;
        v = (a.intact_first_peak && fabs(arg) < kPi / 2.) ? 1. : -(K3 * sin(arg) - c[5] * cos(arg));
        v *= exp(K4 * u2);
        if (a.ctf_abs) v = fabs(v);
        else if (a.ctf_flip) v = v < 0. ? -1. : 1.;
    }
    const size_t o = (size_t)img * cur * h + q;
    if (a.fctf) a.fctf[o] = v;
    if (a.fctf32) a.fctf32[o] = (float)v;
}

}  // namespace

hipError_t launch_imageprep_pre(hipStream_t s, int n, const ImagePrepArgs& a) {
    if (n <= 0) return hipSuccess;
    if (a.ori > 2 * (kMaxShell - 1)) return hipErrorInvalidValue;
    dim3 g((a.ori * a.ori + 255) / 256, n);
    k_ip_shift<<<g, 256, 0, s>>>(a);
    return hipGetLastError();
}

hipError_t launch_imageprep_mask(hipStream_t s, int n, const ImagePrepArgs& a) {
    if (n <= 0) return hipSuccess;
    k_ip_mask<<<n, 256, 0, s>>>(a);
    return hipGetLastError();
}

hipError_t launch_imageprep_post(hipStream_t s, int n, const ImagePrepArgs& a) {
    if (n <= 0) return hipSuccess;
    if (a.cur > a.ori) return hipErrorInvalidValue;
    const int cF = a.cur * (a.cur / 2 + 1);
    dim3 g((cF + 255) / 256, n);
    k_ip_window<<<g, 256, 0, s>>>(a);
    if (a.cur < a.ori) k_ip_power<<<n, 256, 0, s>>>(a);
    k_ip_ctf<<<g, 256, 0, s>>>(a);
    return hipGetLastError();
}

}  // namespace rome
