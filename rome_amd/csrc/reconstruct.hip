// M-step map of one class on the GPU: MyBackProjector::reconstruct (src/reconstruct.cpp:1929-2522),
// FP64 end to end (the reference keeps Fnewweight in double, :1946-1949), FFTs by hipFFT.
//
// Grids: data (complex) and weight (real) in the projector's centred layout [Dz][P][PF] (Dz = P in 3D,
// 1 in 2D; origin at (P/2, P/2, 0)).  "FFTW order" arrays use the same shape with the signed index
// k < PF ? k : k - P along z and y (decenter, src/reconstruct.h:281-372).
//
//   fold      enforceHermitianSymmetry (:1501-1537): each x = 0 Friedel pair summed, one thread per pair
//   decenter  weight -> Fweight (r^2 <= max_r2), later data -> Fconv
//   spectrum  sigma2 / counter per shell (r^2 < max_r2), block sums in LDS, FP64 atomics per shell
//   map       MAP term: data_vs_prior sums, Fweight += 1 / (osc tau2_fudge tau2) for shells >= minres_map
//   iterate   Fconv = Fnewweight Fweight -> c2r -> x blob transform (table) -> r2c / N -> Fnewweight /= |.|
//   window    FFTW-order Fconv cut to the 2 ori grid (windowFourierTransform) -> c2r -> centre, window to
//             ori (window, centerFFT) / normfft -> soft mask with the average background -> / sinc^2
#include <hip/hip_runtime.h>

#include "device_util.h"
#include "kernels.h"

namespace rome {

namespace {

constexpr double kPi = 3.14159265358979323846;

DEV int fftw_signed(int k, int half, int n) { return k < half ? k : k - n; }

__global__ void k_rc_fold(double2* __restrict__ data, double* __restrict__ weight, int Dz, int P, int PF) {
    const int oz = -(int)((float)Dz / 2.0f), oy = -(int)((float)P / 2.0f);
    const int a = blockIdx.x * blockDim.x + threadIdx.x;  // over Dz * P candidate (iz, iy)
    if (a >= Dz * P) return;
    const int iz = a / P + oz, iy = a % P + oy;
    if (iy < (iz < 0 ? 0 : 1)) return;
    const int zr = iz - oz, yr = iy - oy, zl = -iz - oz, yl = -iy - oy;
    const size_t r = ((size_t)zr * P + yr) * PF, l = ((size_t)zl * P + yl) * PF;
    const double fr = data[r].x + data[l].x, fi = data[r].y - data[l].y;
    data[r] = make_double2(fr, fi);
    data[l] = make_double2(fr, -fi);
    const double s = weight[r] + weight[l];
    weight[r] = s;
    weight[l] = s;
}

// FFTW-order element n -> signed (kp, ip, jp) and the centred source index
DEV void rc_coords(size_t n, int Dz, int P, int PF, int& kp, int& ip, int& jp, size_t& cen) {
    const int j = (int)(n % PF);
    const int i = (int)((n / PF) % P);
    const int k = (int)(n / ((size_t)PF * P));
    kp = Dz > 1 ? fftw_signed(k, PF, Dz) : 0;
    ip = fftw_signed(i, PF, P);
    jp = j;
    const int oz = -(int)((float)Dz / 2.0f), oy = -(int)((float)P / 2.0f);
    cen = ((size_t)(kp - oz) * P + (ip - oy)) * PF + jp;
}

__global__ void k_rc_decenter_real(const double* __restrict__ cen, double* __restrict__ out, int Dz, int P, int PF,
                                   int max_r2) {
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= (size_t)Dz * P * PF) return;
    int kp, ip, jp;
    size_t c;
    rc_coords(n, Dz, P, PF, kp, ip, jp, c);
    out[n] = (kp * kp + ip * ip + jp * jp <= max_r2) ? cen[c] : 0.;
}

constexpr int kRcShells = 520;

// per shell of r^2 < max_r2: s0 += osc * Fweight, s1 += 1 (sigma2 spectrum)
__global__ __launch_bounds__(256) void k_rc_spectrum(const double* __restrict__ Fw, int Dz, int P, int PF, int max_r2,
                                                     int pf, double osc, int oriF, double* __restrict__ s0,
                                                     double* __restrict__ s1) {
    __shared__ double a0[kRcShells], a1[kRcShells];
    for (int q = threadIdx.x; q < oriF; q += blockDim.x) a0[q] = a1[q] = 0.;
    __syncthreads();
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n < (size_t)Dz * P * PF) {
        int kp, ip, jp;
        size_t c;
        rc_coords(n, Dz, P, PF, kp, ip, jp, c);
        const int r2 = kp * kp + ip * ip + jp * jp;
        if (r2 < max_r2) {
            const int ires = (int)round(sqrt((double)r2) / pf);
            atomicAdd(&a0[ires], osc * Fw[n]);
            atomicAdd(&a1[ires], 1.);
        }
    }
    __syncthreads();
    for (int q = threadIdx.x; q < oriF; q += blockDim.x)
        if (a1[q] != 0.) {
            atomicAdd(&s0[q], a0[q]);
            atomicAdd(&s1[q], a1[q]);
        }
}

// MAP term (:2081-2139): data_vs_prior sums and the regularised weight; *bad = 1 on a negative tau2
__global__ __launch_bounds__(256) void k_rc_map(double* __restrict__ Fw, int Dz, int P, int PF, int max_r2, int pf,
                                                double osc, double tau2_fudge, const double* __restrict__ tau2,
                                                int minres, int oriF, double* __restrict__ dvp,
                                                double* __restrict__ cnt, int* __restrict__ bad) {
    __shared__ double a0[kRcShells], a1[kRcShells];
    for (int q = threadIdx.x; q < oriF; q += blockDim.x) a0[q] = a1[q] = 0.;
    __syncthreads();
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n < (size_t)Dz * P * PF) {
        int kp, ip, jp;
        size_t c;
        rc_coords(n, Dz, P, PF, kp, ip, jp, c);
        const int r2 = kp * kp + ip * ip + jp * jp;
        if (r2 < max_r2) {
            const int ires = (int)round(sqrt((double)r2) / pf);
            const double invw = Fw[n];
            const double t = tau2[ires];
            double invtau2;
            if (t > 0.) {
                invtau2 = 1. / (osc * tau2_fudge * t);
            } else if (t == 0.) {
                invtau2 = 1. / (0.001 * invw);
            } else {
                invtau2 = 0.;
                *bad = 1;
            }
            atomicAdd(&a0[ires], invw / invtau2);
            atomicAdd(&a1[ires], 1.);
            if (ires >= minres) Fw[n] = invw + invtau2;
        }
    }
    __syncthreads();
    for (int q = threadIdx.x; q < oriF; q += blockDim.x)
        if (a1[q] != 0.) {
            atomicAdd(&dvp[q], a0[q]);
            atomicAdd(&cnt[q], a1[q]);
        }
}

__global__ void k_rc_normalise(double* __restrict__ Fw, double2* __restrict__ data, size_t n, double norm) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Fw[i] /= norm;
    const double2 d = data[i];
    data[i] = make_double2(d.x / norm, d.y / norm);
}

__global__ void k_rc_newweight_init(double* __restrict__ Fnw, int Dz, int P, int PF, int max_r2) {
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= (size_t)Dz * P * PF) return;
    int kp, ip, jp;
    size_t c;
    rc_coords(n, Dz, P, PF, kp, ip, jp, c);
    Fnw[n] = (kp * kp + ip * ip + jp * jp < max_r2) ? 1. : 0.;
}

__global__ void k_rc_fconv(const double* __restrict__ Fnw, const double* __restrict__ Fw, double2* __restrict__ Fc,
                           size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) Fc[i] = make_double2(Fnw[i] * Fw[i], 0.);
}

// real-space array [Dz][P][P]: multiply by the normalised blob transform (convoluteBlobRealSpace, :1754-1777)
__global__ void k_rc_blob(double* __restrict__ M, int Dz, int P, int ori, int pf, const double* __restrict__ tab,
                          int ntab, double sampling, double norm) {
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= (size_t)Dz * P * P) return;
    const int padh = P / 2;
    const int j = (int)(n % P), i = (int)((n / P) % P), k = (int)(n / ((size_t)P * P));
    const int kp = Dz > 1 ? (k < padh ? k : k - P) : 0;
    const int ip = i < padh ? i : i - P, jp = j < padh ? j : j - P;
    const double rval = sqrt((double)(kp * kp + ip * ip + jp * jp)) / (double)(ori * pf);
    const int idx = (int)(fabs(rval) / sampling);
    const double t = idx < ntab ? tab[idx] : 0.;
    M[n] *= t / norm;
}

__global__ void k_rc_newweight(const double2* __restrict__ Fc, double* __restrict__ Fnw, int Dz, int P, int PF,
                               int max_r2, double size3) {
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= (size_t)Dz * P * PF) return;
    int kp, ip, jp;
    size_t c;
    rc_coords(n, Dz, P, PF, kp, ip, jp, c);
    if (kp * kp + ip * ip + jp * jp < max_r2) {
        const double re = Fc[n].x / size3, im = Fc[n].y / size3;
        const double w = fmax(1e-6, sqrt(re * re + im * im));
        Fnw[n] /= w;
    }
}

__global__ void k_rc_final_fconv(const double2* __restrict__ data, const double* __restrict__ Fnw,
                                 double2* __restrict__ Fc, int Dz, int P, int PF, int max_r2) {
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= (size_t)Dz * P * PF) return;
    int kp, ip, jp;
    size_t c;
    rc_coords(n, Dz, P, PF, kp, ip, jp, c);
    double2 v = make_double2(0., 0.);
    if (kp * kp + ip * ip + jp * jp <= max_r2) v = data[c];
    const double w = Fnw[n];
    Fc[n] = make_double2(v.x * w, v.y * w);
}

// windowFourierTransform, growing branch (src/reconstruct.h:498-517): [NZ][N][NF] (zeroed) <- [Dz][P][PF],
// entries with r^2 <= (PF - 1)^2
__global__ void k_rc_window_ft_grow(const double2* __restrict__ Fc, double2* __restrict__ Ft, int Dz, int P, int PF,
                                    int NZ, int N, int NF) {
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= (size_t)Dz * P * PF) return;
    const int j = (int)(n % PF), i = (int)((n / PF) % P), k = (int)(n / ((size_t)PF * P));
    const int kp = Dz > 1 ? (k < PF ? k : k - Dz) : 0;
    const int ip = i < PF ? i : i - P;
    if (kp * kp + ip * ip + j * j > (PF - 1) * (PF - 1)) return;
    const int k2 = kp < 0 ? kp + NZ : kp, i2 = ip < 0 ? ip + N : ip;
    Ft[((size_t)k2 * N + i2) * NF + j] = Fc[n];
}

// windowFourierTransform, shrinking branch (src/reconstruct.h:518-531): [NZ][N][NF] <- [Dz][P][PF]
__global__ void k_rc_window_ft(const double2* __restrict__ Fc, double2* __restrict__ Ft, int Dz, int P, int PF, int NZ,
                               int N, int NF) {
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= (size_t)NZ * N * NF) return;
    const int j = (int)(n % NF), i = (int)((n / NF) % N), k = (int)(n / ((size_t)NF * N));
    const int kp = NZ > 1 ? (k < NF ? k : k - NZ) : 0;
    const int ip = i < NF ? i : i - N;
    const int k2 = kp < 0 ? kp + Dz : kp, i2 = ip < 0 ? ip + P : ip;
    Ft[n] = Fc[((size_t)k2 * P + i2) * PF + j];
}

// centre (centerFFT forward) and window the [NZ][N][N] map to [OZ][ori][ori], divided by normfft
__global__ void k_rc_center_window(const double* __restrict__ Mout, double* __restrict__ vol, int NZ, int N, int OZ,
                                   int ori, double normfft) {
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= (size_t)OZ * ori * ori) return;
    const int j = (int)(n % ori), i = (int)((n / ori) % ori), k = (int)(n / ((size_t)ori * ori));
    const int off = -(int)((float)ori / 2.0f) + (int)((float)N / 2.0f);  // window offset
    auto src = [&](int a, int dim) { return ((a + off - dim / 2) % dim + dim) % dim; };
    const int ks = NZ > 1 ? src(k, N) : 0;
    vol[n] = Mout[((size_t)ks * N + src(i, N)) * N + src(j, N)] / normfft;
}

DEV double rc_radius(size_t n, int OZ, int ori) {
    const int j = (int)(n % ori), i = (int)((n / ori) % ori), k = (int)(n / ((size_t)ori * ori));
    const int o = -(int)((float)ori / 2.0f), oz = -(int)((float)OZ / 2.0f);
    const int x = j + o, y = i + o, z = k + oz;
    return sqrt((double)(z * z + y * y + x * x));
}

// background sums of softMaskOutsideMap (src/reconstruct.h:224-250): [0] weight, [1] weighted value
__global__ __launch_bounds__(256) void k_rc_bgsum(const double* __restrict__ vol, int OZ, int ori, double radius,
                                                  double width, double* __restrict__ sums) {
    __shared__ double red[16];
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    double s = 0., sv = 0.;
    if (n < (size_t)OZ * ori * ori) {
        const double r = rc_radius(n, OZ, ori), rp = radius + width;
        if (r > rp) {
            s = 1.;
            sv = vol[n];
        } else if (r >= radius) {
            const double rc = 0.5 + 0.5 * cos(kPi * (rp - r) / width);
            s = rc;
            sv = rc * vol[n];
        }
    }
    s = block_reduce(s, [](double a, double b) { return a + b; }, red);
    sv = block_reduce(sv, [](double a, double b) { return a + b; }, red);
    if (threadIdx.x == 0) {
        atomicAdd(&sums[0], s);
        atomicAdd(&sums[1], sv);
    }
}

// soft mask with the average background, then the trilinear gridding correction (griddingCorrect)
__global__ void k_rc_mask_grid(double* __restrict__ vol, int OZ, int ori, int pf, double radius, double width,
                               const double* __restrict__ sums) {
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= (size_t)OZ * ori * ori) return;
    const double bg = sums[1] / sums[0];
    const double r = rc_radius(n, OZ, ori), rp = radius + width;
    double v = vol[n];
    if (r > rp) {
        v = bg;
    } else if (r >= radius) {
        const double rc = 0.5 + 0.5 * cos(kPi * (rp - r) / width);
        v = (1. - rc) * v + rc * bg;
    }
    if (r > 0.) {
        const double rval = r / (double)(ori * pf);
        const double sinc = sin(kPi * rval) / (kPi * rval);
        v /= sinc * sinc;
    }
    vol[n] = v;
}

// power spectrum of the new map's transform (update_tau2_with_fsc, :2473-2500): [ori] shells by round(|k|)
__global__ __launch_bounds__(256) void k_rc_map_spectrum(const double2* __restrict__ F, int OZ, int ori, double size3,
                                                         double* __restrict__ spec, double* __restrict__ cnt) {
    __shared__ double a0[kRcShells * 2], a1[kRcShells * 2];
    const int nsh = ori;  // the reference's spectrum vector has vol_out.dimx entries
    for (int q = threadIdx.x; q < nsh; q += blockDim.x) a0[q] = a1[q] = 0.;
    __syncthreads();
    const int oF = ori / 2 + 1;
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n < (size_t)OZ * ori * oF) {
        const int j = (int)(n % oF), i = (int)((n / oF) % ori), k = (int)(n / ((size_t)oF * ori));
        const int kp = OZ > 1 ? fftw_signed(k, oF, OZ) : 0, ip = fftw_signed(i, oF, ori);
        const int idx = (int)round(sqrt((double)(kp * kp + ip * ip + j * j)));
        const double re = F[n].x / size3, im = F[n].y / size3;
        atomicAdd(&a0[idx], re * re + im * im);
        atomicAdd(&a1[idx], 1.);
    }
    __syncthreads();
    for (int q = threadIdx.x; q < nsh; q += blockDim.x)
        if (a1[q] != 0.) {
            atomicAdd(&spec[q], a0[q]);
            atomicAdd(&cnt[q], a1[q]);
        }
}

__global__ void k_rc_from_float(const float2* __restrict__ d32, const float* __restrict__ w32, double2* __restrict__ d64,
                                double* __restrict__ w64, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    d64[i] = make_double2((double)d32[i].x, (double)d32[i].y);
    w64[i] = (double)w32[i];
}

inline unsigned nblk(size_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

hipError_t rc_from_float(hipStream_t s, const float2* d32, const float* w32, double2* d64, double* w64, size_t n) {
    k_rc_from_float<<<nblk(n), 256, 0, s>>>(d32, w32, d64, w64, n);
    return hipGetLastError();
}

hipError_t rc_fold(hipStream_t s, double2* data, double* weight, int Dz, int P, int PF) {
    k_rc_fold<<<nblk((size_t)Dz * P), 256, 0, s>>>(data, weight, Dz, P, PF);
    return hipGetLastError();
}
hipError_t rc_decenter_real(hipStream_t s, const double* cen, double* out, int Dz, int P, int PF, int max_r2) {
    k_rc_decenter_real<<<nblk((size_t)Dz * P * PF), 256, 0, s>>>(cen, out, Dz, P, PF, max_r2);
    return hipGetLastError();
}
hipError_t rc_spectrum(hipStream_t s, const double* Fw, int Dz, int P, int PF, int max_r2, int pf, double osc, int oriF,
                       double* s0, double* s1) {
    if (oriF > kRcShells) return hipErrorInvalidValue;
    k_rc_spectrum<<<nblk((size_t)Dz * P * PF), 256, 0, s>>>(Fw, Dz, P, PF, max_r2, pf, osc, oriF, s0, s1);
    return hipGetLastError();
}
hipError_t rc_map(hipStream_t s, double* Fw, int Dz, int P, int PF, int max_r2, int pf, double osc, double tau2_fudge,
                  const double* tau2, int minres, int oriF, double* dvp, double* cnt, int* bad) {
    if (oriF > kRcShells) return hipErrorInvalidValue;
    k_rc_map<<<nblk((size_t)Dz * P * PF), 256, 0, s>>>(Fw, Dz, P, PF, max_r2, pf, osc, tau2_fudge, tau2, minres, oriF,
                                                       dvp, cnt, bad);
    return hipGetLastError();
}
hipError_t rc_normalise(hipStream_t s, double* Fw, double2* data, size_t n, double norm) {
    k_rc_normalise<<<nblk(n), 256, 0, s>>>(Fw, data, n, norm);
    return hipGetLastError();
}
hipError_t rc_newweight_init(hipStream_t s, double* Fnw, int Dz, int P, int PF, int max_r2) {
    k_rc_newweight_init<<<nblk((size_t)Dz * P * PF), 256, 0, s>>>(Fnw, Dz, P, PF, max_r2);
    return hipGetLastError();
}
hipError_t rc_fconv(hipStream_t s, const double* Fnw, const double* Fw, double2* Fc, size_t n) {
    k_rc_fconv<<<nblk(n), 256, 0, s>>>(Fnw, Fw, Fc, n);
    return hipGetLastError();
}
hipError_t rc_blob(hipStream_t s, double* M, int Dz, int P, int ori, int pf, const double* tab, int ntab,
                   double sampling, double norm) {
    k_rc_blob<<<nblk((size_t)Dz * P * P), 256, 0, s>>>(M, Dz, P, ori, pf, tab, ntab, sampling, norm);
    return hipGetLastError();
}
hipError_t rc_newweight(hipStream_t s, const double2* Fc, double* Fnw, int Dz, int P, int PF, int max_r2, double size3) {
    k_rc_newweight<<<nblk((size_t)Dz * P * PF), 256, 0, s>>>(Fc, Fnw, Dz, P, PF, max_r2, size3);
    return hipGetLastError();
}
hipError_t rc_final_fconv(hipStream_t s, const double2* data, const double* Fnw, double2* Fc, int Dz, int P, int PF,
                          int max_r2) {
    k_rc_final_fconv<<<nblk((size_t)Dz * P * PF), 256, 0, s>>>(data, Fnw, Fc, Dz, P, PF, max_r2);
    return hipGetLastError();
}
hipError_t rc_window_ft(hipStream_t s, const double2* Fc, double2* Ft, int Dz, int P, int PF, int NZ, int N, int NF) {
    if (NF < PF) {
        k_rc_window_ft<<<nblk((size_t)NZ * N * NF), 256, 0, s>>>(Fc, Ft, Dz, P, PF, NZ, N, NF);
    } else {  // N > P (current_size < ori_size): the grid grows; P is odd and N even, so never equal
        hipError_t e = hipMemsetAsync(Ft, 0, (size_t)NZ * N * NF * sizeof(double2), s);
        if (e != hipSuccess) return e;
        k_rc_window_ft_grow<<<nblk((size_t)Dz * P * PF), 256, 0, s>>>(Fc, Ft, Dz, P, PF, NZ, N, NF);
    }
    return hipGetLastError();
}
hipError_t rc_center_window(hipStream_t s, const double* Mout, double* vol, int NZ, int N, int OZ, int ori,
                            double normfft) {
    k_rc_center_window<<<nblk((size_t)OZ * ori * ori), 256, 0, s>>>(Mout, vol, NZ, N, OZ, ori, normfft);
    return hipGetLastError();
}
hipError_t rc_bgsum(hipStream_t s, const double* vol, int OZ, int ori, double radius, double width, double* sums) {
    k_rc_bgsum<<<nblk((size_t)OZ * ori * ori), 256, 0, s>>>(vol, OZ, ori, radius, width, sums);
    return hipGetLastError();
}
hipError_t rc_mask_grid(hipStream_t s, double* vol, int OZ, int ori, int pf, double radius, double width,
                        const double* sums) {
    k_rc_mask_grid<<<nblk((size_t)OZ * ori * ori), 256, 0, s>>>(vol, OZ, ori, pf, radius, width, sums);
    return hipGetLastError();
}
hipError_t rc_map_spectrum(hipStream_t s, const double2* F, int OZ, int ori, double size3, double* spec, double* cnt) {
    if (ori > 2 * kRcShells) return hipErrorInvalidValue;
    k_rc_map_spectrum<<<nblk((size_t)OZ * ori * (ori / 2 + 1)), 256, 0, s>>>(F, OZ, ori, size3, spec, cnt);
    return hipGetLastError();
}

}  // namespace rome
