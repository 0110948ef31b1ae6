/*
 * rome_oracle.cpp -- FP64 CPU restatement of ROME 1.1.2's MAP E-step and
 * Fourier reconstruction hot path (test infrastructure; see rome_oracle.h).
 *
 * Restated from the reference loops cited at each function.  The restatement
 * keeps the reference's iteration order wherever it decides a result
 * (argmax tie-breaking, significance threshold), its tabulated sin/cos shift,
 * its trilinear gather/scatter rules, and its FP64 arithmetic.
 */
#include "rome_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <utility>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

const double kPi = 3.14159265358979323846;          // src/macros.h:69, src/math.h:38
const double kHalfPi = 1.570796326794896619231321691639751442099; // Healpix lsconstants.h:43
const double kFltEpsilon = 1.19209e-07;              // src/math.h:33 fallback value

inline double rad2deg(double r) { return r * 180. / kPi; }  // src/math.h:40
inline double deg2rad(double d) { return d * kPi / 180.; }  // src/math.h:42
inline double acosd(double x) { return acos(x) * 180. / kPi; } // src/math.h:44
inline double sgn(double x) { return (x >= 0) ? 1 : -1; }   // src/math.h:46

int omp_threads(int want) {
#ifdef _OPENMP
    return want > 0 ? want : omp_get_max_threads();
#else
    (void)want;
    return 1;
#endif
}
int thread_id() {
#ifdef _OPENMP
    return omp_get_thread_num();
#else
    return 0;
#endif
}

/* ---------------- HEALPix NEST (Healpix_2.15a/healpix_base.cc) ---------------- */
const int kJrll[12] = {2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4};   // healpix_base.cc:60
const int kJpll[12] = {1, 3, 5, 7, 0, 2, 4, 6, 1, 3, 5, 7};   // healpix_base.cc:61

// NEST numbering interleaves the bits of ix (even positions) and iy (odd
// positions) inside a face (healpix_base.cc:127-143).
inline int compact_even_bits(int v) {
    int r = 0;
    for (int b = 0; b < 15; ++b) r |= ((v >> (2 * b)) & 1) << b;
    return r;
}
inline int spread_bits(int v) {
    int r = 0;
    for (int b = 0; b < 15; ++b) r |= ((v >> b) & 1) << (2 * b);
    return r;
}
void nest2xyf(int order, int pix, int& ix, int& iy, int& face) {
    int npface = 1 << (2 * order);
    face = pix >> (2 * order);
    pix &= (npface - 1);
    ix = compact_even_bits(pix);
    iy = compact_even_bits(pix >> 1);
}
int xyf2nest(int order, int ix, int iy, int face) {
    return (face << (2 * order)) + (spread_bits(ix) | (spread_bits(iy) << 1));
}
// Healpix_Base::pix2ang_z_phi, NEST branch (healpix_base.cc:460, 494-527)
void pix2ang_z_phi(int order, int pix, double& z, double& phi) {
    int nside = 1 << order;
    int npface = nside << order;
    int npix = 12 * npface;
    double fact2 = 4. / npix;
    double fact1 = (nside << 1) * fact2;
    int nl4 = nside * 4;
    int face, ix, iy;
    nest2xyf(order, pix, ix, iy, face);
    int jr = (kJrll[face] << order) - ix - iy - 1;
    int nr, kshift;
    if (jr < nside) {
        nr = jr;
        z = 1 - nr * nr * fact2;
        kshift = 0;
    } else if (jr > 3 * nside) {
        nr = nl4 - jr;
        z = nr * nr * fact2 - 1;
        kshift = 0;
    } else {
        nr = nside;
        z = (2 * nside - jr) * fact1;
        kshift = (jr - nside) & 1;
    }
    int jp = (kJpll[face] * nr + ix - iy + 1 + kshift) / 2;
    if (jp > nl4) jp -= nl4;
    if (jp < 1) jp += nl4;
    phi = (jp - (kshift + 1) * 0.5) * (kHalfPi / nr);
}

// HealpixSampler::checkDirection (src/sampling.h:59-75)
void check_direction(double& rot, double& tilt) {
    if (tilt < 0.) {
        tilt = -tilt;
        rot += 180.;
    }
    for (;;) {
        if (rot > 180.) rot -= 360.;
        else if (rot < -180.) rot += 360.;
        else break;
    }
}

/* ---------------- Euler (src/math.cpp:86-114, 149-195) ---------------- */
void euler_angles2matrix(double alpha, double beta, double gamma, double A[3][3]) {
    alpha = deg2rad(alpha);
    beta = deg2rad(beta);
    gamma = deg2rad(gamma);
    double ca = cos(alpha), cb = cos(beta), cg = cos(gamma);
    double sa = sin(alpha), sb = sin(beta), sg = sin(gamma);
    double cc = cb * ca, cs = cb * sa, sc = sb * ca, ss = sb * sa;
    A[0][0] = cg * cc - sg * sa;
    A[0][1] = cg * cs + sg * ca;
    A[0][2] = -cg * sb;
    A[1][0] = -sg * cc - cg * sa;
    A[1][1] = -sg * cs + cg * ca;
    A[1][2] = sg * sb;
    A[2][0] = sc;
    A[2][1] = ss;
    A[2][2] = cb;
}
void euler_matrix2angles(const double A[3][3], double& alpha, double& beta, double& gamma) {
    double abs_sb = sqrt(A[0][2] * A[0][2] + A[1][2] * A[1][2]);
    if (abs_sb > 16 * kFltEpsilon) {
        gamma = atan2(A[1][2], -A[0][2]);
        alpha = atan2(A[2][1], A[2][0]);
        double sign_sb;
        if (std::fabs(sin(gamma)) < kFltEpsilon)
            sign_sb = sgn(-A[0][2] / cos(gamma));
        else
            sign_sb = (sin(gamma) > 0) ? sgn(A[1][2]) : -sgn(A[1][2]);
        beta = atan2(sign_sb * abs_sb, A[2][2]);
    } else {
        if (sgn(A[2][2]) > 0) {
            alpha = 0;
            beta = 0;
            gamma = atan2(-A[1][0], A[0][0]);
        } else {
            alpha = 0;
            beta = kPi;
            gamma = atan2(A[1][0], -A[0][0]);
        }
    }
    gamma = rad2deg(gamma);
    beta = rad2deg(beta);
    alpha = rad2deg(alpha);
}

// HealpixSampler::addPerturbation (src/sampling.cpp:510-528): pose <- angles(A(pose) * R(p,p,p))
void add_perturbation(double p_deg, double& rot, double& tilt, double& psi) {
    if (!(std::fabs(p_deg) > 0.)) return;
    double A[3][3], R[3][3], AR[3][3];
    euler_angles2matrix(rot, tilt, psi, A);
    euler_angles2matrix(p_deg, p_deg, p_deg, R);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) AR[i][j] = A[i][0] * R[0][j] + A[i][1] * R[1][j] + A[i][2] * R[2][j];
    euler_matrix2angles(AR, rot, tilt, psi);
}

/* ---------------- tabulated Fourier shift (src/image.h:1478-1574) ---------------- */
struct SinCosTable {
    double s[5000], c[5000];
    SinCosTable() {
        const double step = 2 * kPi / 5000.0;
        for (int i = 0; i < 5000; ++i) {
            s[i] = sin((double)i * step);
            c[i] = cos((double)i * step);
        }
    }
};
const SinCosTable& table() {
    static SinCosTable t;
    return t;
}

void shift_image(const double* fin_re, const double* fin_im, double* fout_re, double* fout_im, int size,
                 double shift_x, double shift_y, int ori_size) {
    const SinCosTable& tb = table();
    const int nr_element = 5000;
    const int Fsize = size / 2 + 1;
    const int Fsize2 = size * Fsize;
    const double sampling = 2 * kPi / (double)nr_element;
    double xshift = shift_x / (-ori_size);
    double yshift = shift_y / (-ori_size);
    if (std::fabs(xshift) < 1e-6 && std::fabs(yshift) < 1e-6) {  // ROME_EQUAL_ACCURACY
        for (int i = 0; i < Fsize2; ++i) {
            fout_re[i] = fin_re[i];
            fout_im[i] = fin_im[i];
        }
        return;
    }
    for (int i = 0; i < size; ++i) {
        double y = (i < Fsize) ? i : i - size;
        for (int j = 0; j < Fsize; ++j) {
            double x = j;
            double dotp = 2 * kPi * (x * xshift + y * yshift);
            double a = tb.c[((int)(std::fabs(dotp) / sampling)) % nr_element];
            double b = tb.s[((int)(std::fabs(dotp) / sampling)) % nr_element];
            b = (dotp < 0) ? -b : b;
            double c = fin_re[i * Fsize + j], d = fin_im[i * Fsize + j];
            double ac = a * c, bd = b * d, ab_cd = (a + b) * (c + d);
            fout_re[i * Fsize + j] = ac - bd;
            fout_im[i * Fsize + j] = ab_cd - ac - bd;
        }
    }
}

// windowFourierTransform / windowTransform, down-sizing branch (src/image.h:420-560)
template <typename F>
void window_rows(int in_size, int out_size, F copy) {
    int in_Fsize = in_size / 2 + 1, out_Fsize = out_size / 2 + 1;
    if (in_size == out_size) {
        for (int i = 0; i < out_size; ++i)
            for (int j = 0; j < out_Fsize; ++j) copy(i * in_Fsize + j, i * out_Fsize + j);
        return;
    }
    if (out_Fsize > in_Fsize) {  // up-sizing: keep r <= in_Fsize-1 (src/image.h:517-540)
        int max_r2 = (in_Fsize - 1) * (in_Fsize - 1);
        for (int i = 0; i < in_size; ++i) {
            int ip = (i < in_Fsize) ? i : i - in_size;
            int in_i = (ip < 0) ? ip + in_size : ip;
            int out_i = (ip < 0) ? ip + out_size : ip;
            int jp2 = max_r2 - ip * ip;
            if (jp2 < 0) continue;
            int j_max = (int)sqrt((double)jp2) + 1;
            j_max = j_max < in_Fsize ? j_max : in_Fsize;
            for (int j = 0; j < j_max; ++j) copy(in_i * in_Fsize + j, out_i * out_Fsize + j);
        }
        return;
    }
    for (int i = 0; i < out_size; ++i) {
        int ip = (i < out_Fsize) ? i : i - out_size;
        int in_i = (ip < 0) ? ip + in_size : ip;
        int out_i = (ip < 0) ? ip + out_size : ip;
        for (int j = 0; j < out_Fsize; ++j) copy(in_i * in_Fsize + j, out_i * out_Fsize + j);
    }
}
void window_fourier(const double* in_re, const double* in_im, int in_size, double* out_re, double* out_im,
                    int out_size) {
    int out_Fsize2 = out_size * (out_size / 2 + 1);
    for (int i = 0; i < out_Fsize2; ++i) out_re[i] = out_im[i] = 0;
    window_rows(in_size, out_size, [&](int a, int b) {
        out_re[b] = in_re[a];
        out_im[b] = in_im[a];
    });
}
void window_real(const double* in, int in_size, double* out, int out_size) {
    int out_Fsize2 = out_size * (out_size / 2 + 1);
    for (int i = 0; i < out_Fsize2; ++i) out[i] = 0;
    window_rows(in_size, out_size, [&](int a, int b) { out[b] = in[a]; });
}

// MAPModel::updateImageSizeAndResolutionPointers, Mresol part (src/map_model.cpp:545-575)
void mresol(int size, int* out) {
    int Fsize = size / 2 + 1;
    for (int i = 0; i < size; ++i)
        for (int j = 0; j < Fsize; ++j) {
            int ip = (i < Fsize) ? i : i - size;
            int jp = j;
            int ires = (int)round(sqrt((double)(ip * ip + jp * jp)));
            out[i * Fsize + j] = (ires < Fsize && !(jp == 0 && ip < 0)) ? ires : -1;
        }
}

/* ---------------- projector (src/reconstruct.cpp:412-623, 823-958) ---------------- */
inline double lin(double a, double l, double h) { return l + (h - l) * a; }  // src/reconstruct.h:544

void project3d(const double* grid, int pad, int r_max, const double A[3][3], int f2d_size, double* out_re,
               double* out_im) {
    const int padF = pad / 2 + 1;
    const int shift = -(int)((float)pad / 2.0);
    const int Fsize = f2d_size / 2 + 1;
    int my_r_max = std::min(r_max, f2d_size / 2);
    int max_r2 = my_r_max * my_r_max;
    double Ai[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Ai[i][j] = A[j][i] * 2.0;  // padding_factor 2, transpose
    for (int n = 0; n < f2d_size * Fsize; ++n) out_re[n] = out_im[n] = 0.;
    auto at = [&](int z, int y, int x) { return grid + 2 * (((size_t)z * pad + y) * padF + x); };
    for (int i = 0; i < f2d_size; ++i) {
        int y;
        if (i <= my_r_max) y = i;
        else if (i >= f2d_size - my_r_max) y = i - f2d_size;
        else continue;
        int y2 = y * y;
        for (int x = 0; x <= my_r_max; ++x) {
            int r2 = x * x + y2;
            if (r2 > max_r2) continue;
            double xp = Ai[0][0] * x + Ai[0][1] * y;
            double yp = Ai[1][0] * x + Ai[1][1] * y;
            double zp = Ai[2][0] * x + Ai[2][1] * y;
            bool neg = false;
            if (xp < 0) {
                xp = -xp;
                yp = -yp;
                zp = -zp;
                neg = true;
            }
            int x0 = (int)floor(xp);
            double fx = xp - x0;
            int y0 = (int)floor(yp);
            double fy = yp - y0;
            y0 -= shift;
            int z0 = (int)floor(zp);
            double fz = zp - z0;
            z0 -= shift;
            const double* d00 = at(z0, y0, x0);
            const double* d01 = at(z0, y0 + 1, x0);
            const double* d10 = at(z0 + 1, y0, x0);
            const double* d11 = at(z0 + 1, y0 + 1, x0);
            double dx00r = lin(fx, d00[0], d00[2]), dx00i = lin(fx, d00[1], d00[3]);
            double dx01r = lin(fx, d10[0], d10[2]), dx01i = lin(fx, d10[1], d10[3]);
            double dx10r = lin(fx, d01[0], d01[2]), dx10i = lin(fx, d01[1], d01[3]);
            double dx11r = lin(fx, d11[0], d11[2]), dx11i = lin(fx, d11[1], d11[3]);
            double dxy0r = lin(fy, dx00r, dx10r), dxy0i = lin(fy, dx00i, dx10i);
            double dxy1r = lin(fy, dx01r, dx11r), dxy1i = lin(fy, dx01i, dx11i);
            out_re[i * Fsize + x] = lin(fz, dxy0r, dxy1r);
            double im = lin(fz, dxy0i, dxy1i);
            out_im[i * Fsize + x] = neg ? -1.0 * im : im;
        }
    }
}

void rotate2d(const double* grid, int pad, int r_max, const double A[3][3], int f2d_size, double* out_re,
              double* out_im) {
    const int padF = pad / 2 + 1;
    const int shift = -(int)((float)pad / 2.0);
    const int Fsize = f2d_size / 2 + 1;
    int my_r_max = std::min(r_max, f2d_size / 2);
    int max_r2 = my_r_max * my_r_max;
    double Ai[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Ai[i][j] = A[j][i] * 2.0;
    for (int n = 0; n < f2d_size * Fsize; ++n) out_re[n] = out_im[n] = 0.;
    for (int i = 0; i < f2d_size; ++i) {
        int y;
        if (i <= my_r_max) y = i;
        else if (i >= f2d_size - my_r_max) y = i - f2d_size;
        else continue;
        int y2 = y * y;
        for (int x = 0; x <= my_r_max; ++x) {
            int r2 = x * x + y2;
            if (r2 > max_r2) continue;
            double xp = Ai[0][0] * x + Ai[0][1] * y;
            double yp = Ai[1][0] * x + Ai[1][1] * y;
            bool neg = false;
            if (xp < 0) {
                xp = -xp;
                yp = -yp;
                neg = true;
            }
            int x0 = (int)floor(xp);
            double fx = xp - x0;
            int y0 = (int)floor(yp);
            double fy = yp - y0;
            y0 -= shift;
            const double* d0 = grid + 2 * ((size_t)y0 * padF + x0);
            const double* d1 = grid + 2 * ((size_t)(y0 + 1) * padF + x0);
            double dx0r = lin(fx, d0[0], d0[2]), dx0i = lin(fx, d0[1], d0[3]);
            double dx1r = lin(fx, d1[0], d1[2]), dx1i = lin(fx, d1[1], d1[3]);
            out_re[i * Fsize + x] = lin(fy, dx0r, dx1r);
            double im = lin(fy, dx0i, dx1i);
            out_im[i * Fsize + x] = neg ? -1.0 * im : im;
        }
    }
}

/* ---------------- backprojector (src/reconstruct.cpp:977-1164, 1353-1499) ---------------- */
void backproject3d(double* data, double* weight, int pad, int r_max, const double A[3][3], int f2d_size,
                   const double* in_re, const double* in_im, const double* mweight) {
    const int padF = pad / 2 + 1;
    const int shift = -(int)((float)pad / 2.0);
    const int Fsize = f2d_size / 2 + 1;
    int max_r2 = r_max * r_max;
    double Ai[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Ai[i][j] = A[j][i] * 2.0;
    for (int i = 0; i < f2d_size; ++i) {
        int y, first_x;
        if (i <= r_max) {
            y = i;
            first_x = 0;
        } else if (i >= f2d_size - r_max) {
            y = i - f2d_size;
            first_x = 1;
        } else
            continue;
        int y2 = y * y;
        for (int x = first_x; x <= r_max; ++x) {
            int r2 = x * x + y2;
            if (r2 > max_r2) continue;
            double vr = in_re[i * Fsize + x], vi = in_im[i * Fsize + x];
            double w = mweight ? mweight[i * Fsize + x] : 1.;
            if (!(w > 0.)) continue;
            double xp = Ai[0][0] * x + Ai[0][1] * y;
            double yp = Ai[1][0] * x + Ai[1][1] * y;
            double zp = Ai[2][0] * x + Ai[2][1] * y;
            bool neg = false;
            if (xp < 0) {
                xp = -xp;
                yp = -yp;
                zp = -zp;
                neg = true;
            }
            int x0 = (int)floor(xp);
            double fx = xp - x0;
            int y0 = (int)floor(yp);
            double fy = yp - y0;
            y0 -= shift;
            int z0 = (int)floor(zp);
            double fz = zp - z0;
            z0 -= shift;
            double mfx = 1. - fx, mfy = 1. - fy, mfz = 1. - fz;
            double dd[8] = {mfz * mfy * mfx, mfz * mfy * fx, mfz * fy * mfx, mfz * fy * fx,
                            fz * mfy * mfx,  fz * mfy * fx,  fz * fy * mfx,  fz * fy * fx};
            if (neg) vi = -vi;
            for (int c = 0; c < 8; ++c) {
                int dz = c >> 2, dy = (c >> 1) & 1, dx = c & 1;
                size_t idx = ((size_t)(z0 + dz) * pad + (y0 + dy)) * padF + (x0 + dx);
                data[2 * idx] += dd[c] * vr;
                data[2 * idx + 1] += dd[c] * vi;
                weight[idx] += dd[c] * w;
            }
        }
    }
}

void backrotate2d(double* data, double* weight, int pad, int r_max, const double A[3][3], int f2d_size,
                  const double* in_re, const double* in_im, const double* mweight) {
    const int padF = pad / 2 + 1;
    const int shift = -(int)((float)pad / 2.0);
    const int Fsize = f2d_size / 2 + 1;
    int max_r2 = r_max * r_max;
    double Ai[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Ai[i][j] = A[j][i] * 2.0;
    for (int i = 0; i < f2d_size; ++i) {
        int y, first_x;
        if (i <= r_max) {
            y = i;
            first_x = 0;
        } else if (i >= f2d_size - r_max) {
            y = i - f2d_size;
            first_x = 1;
        } else
            continue;
        int y2 = y * y;
        for (int x = first_x; x <= r_max; ++x) {
            int r2 = x * x + y2;
            if (r2 > max_r2) continue;
            double vr = in_re[i * Fsize + x], vi = in_im[i * Fsize + x];
            double w = mweight ? mweight[i * Fsize + x] : 1.;
            if (!(w > 0.)) continue;
            double xp = Ai[0][0] * x + Ai[0][1] * y;
            double yp = Ai[1][0] * x + Ai[1][1] * y;
            bool neg = false;
            if (xp < 0) {
                xp = -xp;
                yp = -yp;
                neg = true;
            }
            int x0 = (int)floor(xp);
            double fx = xp - x0;
            int y0 = (int)floor(yp);
            double fy = yp - y0;
            y0 -= shift;
            double mfx = 1. - fx, mfy = 1. - fy;
            double dd[4] = {mfy * mfx, mfy * fx, fy * mfx, fy * fx};
            if (neg) vi = -vi;
            for (int c = 0; c < 4; ++c) {
                int dy = c >> 1, dx = c & 1;
                size_t idx = (size_t)(y0 + dy) * padF + (x0 + dx);
                data[2 * idx] += dd[c] * vr;
                data[2 * idx + 1] += dd[c] * vi;
                weight[idx] += dd[c] * w;
            }
        }
    }
}

/* ---------------- E-step (src/map3d_optimizer.cpp:883-2504, src/map2d_optimizer.cpp:802-1873) ---------------- */
typedef std::vector<std::pair<int, double>> Sub;

struct Pass {
    int size, Fsize2, os, nover_rot, nover_trans;
    const double* sx;  // [T*nover_trans]
    const double* sy;
};

}  // namespace

extern "C" {

int or_healpix_pix2ang_nest(int order, double* z, double* phi) {
    int npix = 12 << (2 * order);
    for (int p = 0; p < npix; ++p) pix2ang_z_phi(order, p, z[p], phi[p]);
    return npix;
}
int or_healpix_nest2xyf(int order, int pix, int* ix, int* iy, int* face) {
    nest2xyf(order, pix, *ix, *iy, *face);
    return 0;
}
int or_healpix_xyf2nest(int order, int ix, int iy, int face) { return xyf2nest(order, ix, iy, face); }

// HealpixSampler::setTranslations (src/sampling.cpp:121-172): disc of radius offset_range
int or_translations(double offset_range, double offset_step, double* tx, double* ty) {
    int maxr = (int)ceil(offset_range / offset_step);
    int n = 0;
    for (int ix = -maxr; ix <= maxr; ++ix) {
        double xoff = ix * offset_step;
        for (int iy = -maxr; iy <= maxr; ++iy) {
            double yoff = iy * offset_step;
            if (xoff * xoff + yoff * yoff <= offset_range * offset_range) {
                if (tx) tx[n] = xoff;
                if (ty) ty[n] = yoff;
                ++n;
            }
        }
    }
    return n;
}

void or_euler_angles2matrix(double rot, double tilt, double psi, double* A9) {
    double A[3][3];
    euler_angles2matrix(rot, tilt, psi, A);
    for (int i = 0; i < 9; ++i) A9[i] = A[i / 3][i % 3];
}
void or_euler_matrix2angles(const double* A9, double* rot, double* tilt, double* psi) {
    double A[3][3];
    for (int i = 0; i < 9; ++i) A[i / 3][i % 3] = A9[i];
    euler_matrix2angles(A, *rot, *tilt, *psi);
}

// HealpixSampler::setOrientations + getOrientations3D + addPerturbation
// (src/sampling.cpp:174-240, 530-585, 510-528) for C1 symmetry.
int or_pose_table3d(int order, int oversampling, double random_perturbation, double* out) {
    int npix = 12 << (2 * order);
    double psi_step = 360. / (6 * round(std::pow(2., order)));
    int nr_psi = (int)ceil(360. / psi_step);
    psi_step = 360. / (double)nr_psi;
    double ang_sampling = 360. / (6 * round(std::pow(2., order)));  // getAngularSampling(0)
    double myperturb = random_perturbation * ang_sampling;
    std::vector<double> rot(npix), tilt(npix);
    for (int p = 0; p < npix; ++p) {
        double z, phi;
        pix2ang_z_phi(order, p, z, phi);
        rot[p] = rad2deg(phi);
        tilt[p] = acosd(z);
        check_direction(rot[p], tilt[p]);
    }
    if (oversampling == 0) {
        for (int d = 0; d < npix; ++d)
            for (int s = 0; s < nr_psi; ++s) {
                double r = rot[d], t = tilt[d], ps = s * psi_step;
                add_perturbation(myperturb, r, t, ps);
                double* o = out + 3 * ((size_t)d * nr_psi + s);
                o[0] = r;
                o[1] = t;
                o[2] = ps;
            }
        return npix * nr_psi;
    }
    int oorder = order + oversampling;
    int fact = 1 << oversampling;
    int nr_ipsi_over = 1 << oversampling;
    int nr_over = (1 << (2 * oversampling)) * nr_ipsi_over;
    for (int d = 0; d < npix; ++d) {
        int x, y, face;
        nest2xyf(order, d, x, y, face);
        double crot[64], ctilt[64];
        int idir_over = 0;
        for (int j = fact * y; j < fact * (y + 1); ++j)
            for (int i = fact * x; i < fact * (x + 1); ++i) {
                int overpix = xyf2nest(oorder, i, j, face);
                double z, phi;
                pix2ang_z_phi(oorder, overpix, z, phi);
                double r = rad2deg(phi), t = acosd(z);
                check_direction(r, t);
                crot[idir_over] = r;
                ctilt[idir_over] = t;
                ++idir_over;
            }
        for (int s = 0; s < nr_psi; ++s) {
            double* o = out + 3 * (((size_t)d * nr_psi + s) * nr_over);
            for (int io = 0; io < idir_over; ++io)
                for (int ip = 0; ip < nr_ipsi_over; ++ip) {
                    double overpsi = s * psi_step - 0.5 * psi_step + (0.5 + ip) * psi_step / nr_ipsi_over;
                    double r = crot[io], t = ctilt[io], ps = overpsi;
                    add_perturbation(myperturb, r, t, ps);
                    double* q = o + 3 * (io * nr_ipsi_over + ip);
                    q[0] = r;
                    q[1] = t;
                    q[2] = ps;
                }
        }
    }
    return npix * nr_psi * nr_over;
}

// HealpixSampler::getOrientations2D (src/sampling.cpp:484-508), 2D psi_step given.
int or_pose_table2d(double psi_step, int oversampling, double random_perturbation, double* out) {
    int nr_psi = (int)ceil(360. / psi_step);
    psi_step = 360. / (double)nr_psi;
    int nover = 1 << oversampling;
    double myperturb = random_perturbation * psi_step;  // getAngularSampling(0) in 2D = psi_step
    for (int s = 0; s < nr_psi; ++s) {
        double base = s * psi_step;
        for (int io = 0; io < nover; ++io) {
            double v = (oversampling == 0) ? base : base - 0.5 * psi_step + (0.5 + io) * psi_step / nover;
            if (std::fabs(random_perturbation) > 0.) v += myperturb;
            out[s * nover + io] = v;
        }
    }
    return nr_psi * nover;
}

// HealpixSampler::getTranslations (src/sampling.cpp:312-349)
int or_shift_table(double offset_range, double offset_step, int oversampling, double random_perturbation,
                   double* out_x, double* out_y) {
    int T = or_translations(offset_range, offset_step, nullptr, nullptr);
    std::vector<double> tx(T), ty(T);
    or_translations(offset_range, offset_step, tx.data(), ty.data());
    int nos = 1 << oversampling;
    int nover = nos * nos;
    double myperturb = random_perturbation * offset_step;
    for (int t = 0; t < T; ++t) {
        for (int iy = 0; iy < nos; ++iy)
            for (int ix = 0; ix < nos; ++ix) {
                double ox, oy;
                if (oversampling == 0) {
                    ox = tx[t];
                    oy = ty[t];
                } else {
                    oy = ty[t] - 0.5 * offset_step + (0.5 + iy) * offset_step / nos;
                    ox = tx[t] - 0.5 * offset_step + (0.5 + ix) * offset_step / nos;
                }
                if (std::fabs(random_perturbation) > 0.) {
                    ox += myperturb;
                    oy += myperturb;
                }
                out_x[t * nover + iy * nos + ix] = ox;
                out_y[t * nover + iy * nos + ix] = oy;
            }
    }
    return T * nover;
}

void or_shift_image(const double* in_re, const double* in_im, double* out_re, double* out_im, int size,
                    double shift_x, double shift_y, int ori_size) {
    shift_image(in_re, in_im, out_re, out_im, size, shift_x, shift_y, ori_size);
}
void or_window_fourier(const double* in_re, const double* in_im, int in_size, double* out_re, double* out_im,
                       int out_size) {
    window_fourier(in_re, in_im, in_size, out_re, out_im, out_size);
}
void or_mresol(int size, int* out) { mresol(size, out); }

void or_project3d(const double* grid, int pad, int r_max, const double* A9, int f2d_size, double* out_re,
                  double* out_im) {
    double A[3][3];
    for (int i = 0; i < 9; ++i) A[i / 3][i % 3] = A9[i];
    project3d(grid, pad, r_max, A, f2d_size, out_re, out_im);
}
void or_rotate2d(const double* grid, int pad, int r_max, const double* A9, int f2d_size, double* out_re,
                 double* out_im) {
    double A[3][3];
    for (int i = 0; i < 9; ++i) A[i / 3][i % 3] = A9[i];
    rotate2d(grid, pad, r_max, A, f2d_size, out_re, out_im);
}
void or_backproject3d(double* grid, double* weight, int pad, int r_max, const double* A9, int f2d_size,
                      const double* in_re, const double* in_im, const double* mweight) {
    double A[3][3];
    for (int i = 0; i < 9; ++i) A[i / 3][i % 3] = A9[i];
    backproject3d(grid, weight, pad, r_max, A, f2d_size, in_re, in_im, mweight);
}
void or_backrotate2d(double* grid, double* weight, int pad, int r_max, const double* A9, int f2d_size,
                     const double* in_re, const double* in_im, const double* mweight) {
    double A[3][3];
    for (int i = 0; i < 9; ++i) A[i / 3][i % 3] = A9[i];
    backrotate2d(grid, weight, pad, r_max, A, f2d_size, in_re, in_im, mweight);
}

int or_estep(const OrConfig* cfg, const OrImages* im, OrOutputs* out, char* errbuf, int errlen) {
    auto fail = [&](const char* msg) {
        if (errbuf && errlen > 0) snprintf(errbuf, errlen, "%s", msg);
        return -1;
    };
    const OrConfig& c = *cfg;
    const int N = im->n, K = c.nr_classes, D = c.nr_dir, PSI = c.nr_psi, T = c.nr_trans;
    const int ori_Fsize = c.ori_size / 2 + 1;
    const int cur = c.current_size, curF = cur / 2 + 1, curF2 = cur * curF;
    const bool is3d = c.is_3d != 0;
    const int os = c.adaptive_oversampling;
    const int npasses = os > 0 ? 2 : 1;
    const int pad = c.pad, padF = pad / 2 + 1;
    const size_t gridN = is3d ? (size_t)pad * pad * padF : (size_t)pad * padF;
    const int r_max = cur / 2;
    const int nthreads = omp_threads(c.nr_threads);
    if (os > 1) return fail("adaptive_oversampling > 1 not supported");
    if (os > 0 && !c.pose_fine) return fail("fine pose table missing");
    const int nover_rot_fine = os ? (is3d ? 8 : 2) : 1;
    const int nover_trans_fine = os ? 4 : 1;

    std::vector<int> mres_fine(curF2);
    mresol(cur, mres_fine.data());

    // Weight store: 3D -> per (img,k,dir,psi) sparse list (src/exp_mweight.h:74-113);
    // 2D -> same layout with dir = 0 (the dense 2D store holds the same entries, in the same order).
    const size_t nsub = (size_t)N * K * D * PSI;
    std::vector<Sub> store(nsub);
    auto sub = [&](int i, int k, int d, int s) -> Sub& { return store[(((size_t)i * K + k) * D + d) * PSI + s]; };
    std::vector<unsigned char> coarse_sig((size_t)N * K * D * PSI * T, 0);
    std::vector<double> min_diff2(N), sum_w(N), sig_w(N), max_w(N);
    std::vector<int> best(N * 6, 0);
    std::vector<double> meta(N * 9, 0.);
    std::vector<int> nr_sign(N, 0);

    Pass last{};
    std::vector<double> local_ctf_fine, local_minv_fine;  // last pass (fine) per-image CTF and Minv

    // coarse significance given (the fine pass of getAllSquaredDifferencesFine over exp_Mcoarse_significant,
    // src/map3d_optimizer.cpp:1261-1372, without the coarse pass that would have set it)
    const bool given = im->coarse_sig_in != nullptr;
    if (given) {
        if (npasses != 2) return fail("given coarse significance needs adaptive_oversampling = 1");
        std::memcpy(coarse_sig.data(), im->coarse_sig_in, coarse_sig.size());
        for (int i = 0; i < N; ++i) {
            const unsigned char* cs = &coarse_sig[(size_t)i * K * D * PSI * T];
            int cnt = 0;
            for (size_t q = 0; q < (size_t)K * D * PSI * T; ++q) cnt += cs[q] != 0;
            if (cnt == 0) return fail("my_nr_significant_coarse_samples == 0");
            nr_sign[i] = cnt;
            const double nan = std::numeric_limits<double>::quiet_NaN();
            if (out->min_diff2_coarse) out->min_diff2_coarse[i] = nan;
            if (out->sum_weight_coarse) out->sum_weight_coarse[i] = nan;
            if (out->sig_weight_coarse) out->sig_weight_coarse[i] = nan;
        }
    }

    for (int ipass = given ? 1 : 0; ipass < npasses; ++ipass) {
        const bool coarse = (ipass == 0);
        Pass P;
        P.size = (coarse && os > 0) ? c.coarse_size : cur;
        P.Fsize2 = P.size * (P.size / 2 + 1);
        P.os = coarse ? 0 : os;
        P.nover_rot = coarse ? 1 : nover_rot_fine;
        P.nover_trans = coarse ? 1 : nover_trans_fine;
        P.sx = coarse ? c.shift_coarse_x : c.shift_fine_x;
        P.sy = coarse ? c.shift_coarse_y : c.shift_fine_y;
        const int F2 = P.Fsize2;
        const int NOT = P.nover_trans, NOR = P.nover_rot;
        std::vector<int> mres(F2);
        mresol(P.size, mres.data());
        std::vector<double> ctf_p((size_t)N * F2), minv_p((size_t)N * F2);

        for (auto& s : store) s.clear();
        // ---- getAllSquaredDifferences{Coarse,Fine} (src/map3d_optimizer.cpp:1195-1372),
        //      preShiftedImagesCtfsAndInvSigma2s (src/exp_model.cpp:935-1103),
        //      2D: src/map2d_optimizer.cpp:907-1074
#pragma omp parallel for schedule(dynamic) num_threads(nthreads)
        for (int i = 0; i < N; ++i) {
            const int g = im->group[i];
            std::vector<double> fr(F2), fi(F2);
            window_fourier(im->fmask_re + (size_t)i * curF2, im->fmask_im + (size_t)i * curF2, cur, fr.data(),
                           fi.data(), P.size);
            double* ctf = ctf_p.data() + (size_t)i * F2;
            double* minv = minv_p.data() + (size_t)i * F2;
            window_real(im->fctf + (size_t)i * curF2, cur, ctf, P.size);
            for (int n = 0; n < F2; ++n) {
                int ires = mres[n];
                minv[n] = (ires > 0) ? 1. / (c.sigma2_fudge * c.sigma2_noise[(size_t)g * ori_Fsize + ires]) : 0;
            }
            std::vector<double> sh_re((size_t)T * NOT * F2), sh_im((size_t)T * NOT * F2);
            for (int t = 0; t < T; ++t)
                for (int ot = 0; ot < NOT; ++ot)
                    shift_image(fr.data(), fi.data(), &sh_re[(size_t)(t * NOT + ot) * F2],
                                &sh_im[(size_t)(t * NOT + ot) * F2], P.size, P.sx[t * NOT + ot], P.sy[t * NOT + ot],
                                c.ori_size);
            const double hxi2 = im->highres_Xi2[i];
            std::vector<double> ref_re((size_t)NOR * F2), ref_im((size_t)NOR * F2);
            double mymin = std::numeric_limits<double>::max();
            for (int k = 0; k < K; ++k) {
                if (c.pdf_class[k] <= 0.) continue;
                const double* grid = c.grids + 2 * gridN * k;
                for (int d = 0; d < D; ++d) {
                    if (c.pdf_direction[(size_t)k * D + d] <= 0.) continue;
                    for (int s = 0; s < PSI; ++s) {
                        Sub& L = sub(i, k, d, s);
                        bool projected = false;
                        for (int t = 0; t < T; ++t) {
                            size_t crt = (((size_t)k * D + d) * PSI + s) * T + t;
                            if (!coarse && !coarse_sig[(size_t)i * K * D * PSI * T + crt]) continue;
                            if (!projected) {
                                for (int orot = 0; orot < NOR; ++orot) {
                                    const double* pose;
                                    double A[3][3];
                                    if (is3d) {
                                        pose = coarse ? c.pose_coarse + 3 * ((size_t)d * PSI + s)
                                                      : c.pose_fine + 3 * (((size_t)d * PSI + s) * NOR + orot);
                                        euler_angles2matrix(pose[0], pose[1], pose[2], A);
                                        project3d(grid, pad, r_max, A, P.size, &ref_re[(size_t)orot * F2],
                                                  &ref_im[(size_t)orot * F2]);
                                    } else {
                                        double psi = coarse ? c.pose_coarse[s] : c.pose_fine[(size_t)s * NOR + orot];
                                        euler_angles2matrix(0, 0, psi, A);
                                        rotate2d(grid, pad, r_max, A, P.size, &ref_re[(size_t)orot * F2],
                                                 &ref_im[(size_t)orot * F2]);
                                    }
                                    if (c.do_ctf_correction && c.refs_are_ctf_corrected)
                                        for (int n = 0; n < F2; ++n) {
                                            ref_re[(size_t)orot * F2 + n] *= ctf[n];
                                            ref_im[(size_t)orot * F2 + n] *= ctf[n];
                                        }
                                }
                                projected = true;
                            }
                            for (int orot = 0; orot < NOR; ++orot)
                                for (int ot = 0; ot < NOT; ++ot) {
                                    const double* rr = &ref_re[(size_t)orot * F2];
                                    const double* ri = &ref_im[(size_t)orot * F2];
                                    const double* xr = &sh_re[(size_t)(t * NOT + ot) * F2];
                                    const double* xi = &sh_im[(size_t)(t * NOT + ot) * F2];
                                    double diff2 = 0;
                                    for (int n = 0; n < F2; ++n) {
                                        double dr = rr[n] - xr[n], di = ri[n] - xi[n];
                                        diff2 += (dr * dr + di * di) * minv[n];
                                    }
                                    diff2 = (diff2 + hxi2) / 2;
                                    int ihidden = (t * NOR + orot) * NOT + ot;
                                    L.push_back(std::make_pair(ihidden, diff2));
                                    if (diff2 < mymin) mymin = diff2;
                                }
                        }
                    }
                }
            }
            min_diff2[i] = mymin;
            // single pass: the one pass is both the coarse and the last ("fine") pass
            double* dumps[2] = {coarse ? out->coarse_diff2 : out->fine_diff2,
                                (coarse && npasses == 1) ? out->fine_diff2 : nullptr};
            for (double* dump : dumps) {
                if (!dump) continue;
                const size_t per = (size_t)K * D * PSI * T * NOR * NOT;
                double* di = dump + (size_t)i * per;
                for (size_t q = 0; q < per; ++q) di[q] = std::numeric_limits<double>::quiet_NaN();
                for (int k = 0; k < K; ++k)
                    for (int d = 0; d < D; ++d)
                        for (int s = 0; s < PSI; ++s)
                            for (auto& e : sub(i, k, d, s))
                                di[(((size_t)k * D + d) * PSI + s) * T * NOR * NOT + e.first] = e.second;
            }
        }

        // ---- convertSquaredDifferencesToWeights (src/map3d_optimizer.cpp:1610-1810; 2D :1077-1224)
        //      + MLModel::calculatePdfOffset (src/ml_model.h:231-242)
#pragma omp parallel for schedule(dynamic) num_threads(nthreads)
        for (int i = 0; i < N; ++i) {
            double sumw = 0., maxw = std::numeric_limits<double>::min();
            int bk = 0, bd = 0, bs = 0, bor = 0, bt = 0, bot = 0;
            for (int k = 0; k < K; ++k)
                for (int d = 0; d < D; ++d)
                    for (int s = 0; s < PSI; ++s) {
                        double pdf_orient = c.pdf_direction[(size_t)k * D + d];
                        Sub& L = sub(i, k, d, s);
                        int pre_t = std::numeric_limits<int>::max();
                        double pdf_offset = std::numeric_limits<double>::max();
                        // 2D visits (t, orot, ot) in the same order as the 3D sub-list
                        for (auto& e : L) {
                            int ot = e.first % NOT, rest = e.first / NOT;
                            int orot = rest % NOR, t = rest / NOR;
                            if (t != pre_t) {
                                pre_t = t;
                                double ox = c.trans_x[t] + im->old_offx[i];
                                double oy = c.trans_y[t] + im->old_offy[i];
                                double tdiff2 = (ox - c.prior_offx[k]) * (ox - c.prior_offx[k]) +
                                                (oy - c.prior_offy[k]) * (oy - c.prior_offy[k]);
                                if (c.sigma2_offset < 0.0001)
                                    pdf_offset = (tdiff2 > 0.) ? 0. : 1.;
                                else
                                    pdf_offset = exp(tdiff2 / (-2. * c.sigma2_offset)) / (2. * kPi * c.sigma2_offset);
                            }
                            double w = pdf_orient * pdf_offset;
                            double dd = e.second - min_diff2[i];
                            if (dd > 700.) w = 0.;
                            else w *= exp(-dd);
                            e.second = w;
                            sumw += w;
                            if (w > maxw) {
                                maxw = w;
                                bk = k; bd = d; bs = s; bor = orot; bt = t; bot = ot;
                            }
                        }
                    }
            sum_w[i] = sumw;
            max_w[i] = maxw;
            int* b = &best[i * 6];
            b[0] = bk; b[1] = bd; b[2] = bs; b[3] = bor; b[4] = bt; b[5] = bot;
            double* m = &meta[i * 9];
            m[0] = maxw / sumw;
            m[1] = bk + 1;
            if (is3d) {
                const double* pose = coarse ? c.pose_coarse + 3 * ((size_t)bd * PSI + bs)
                                            : c.pose_fine + 3 * (((size_t)bd * PSI + bs) * NOR + bor);
                m[2] = pose[0];
                m[3] = pose[1];
                m[4] = pose[2];
            } else {
                double psi = coarse ? c.pose_coarse[bs] : c.pose_fine[(size_t)bs * NOR + bor];
                if (psi > 180.) psi -= 360.;
                if (psi < -180.) psi += 360.;
                m[2] = 0;
                m[3] = 0;
                m[4] = psi;
            }
            m[5] = im->old_offx[i] + P.sx[bt * NOT + bot];
            m[6] = im->old_offy[i] + P.sy[bt * NOT + bot];
        }

        // ---- findAllSignificantPoints (src/map3d_optimizer.cpp:1812-1931; 2D :1226-1311)
        bool zero_sig = false;
#pragma omp parallel for schedule(dynamic) num_threads(nthreads)
        for (int i = 0; i < N; ++i) {
            std::vector<double> sorted;
            for (int k = 0; k < K; ++k)
                for (int d = 0; d < D; ++d)
                    for (int s = 0; s < PSI; ++s)
                        for (auto& e : sub(i, k, d, s))
                            if (is3d || e.second > 0.) sorted.push_back(e.second);
            std::sort(sorted.begin(), sorted.end());
            double frac = 0., my_sig = 0.;
            int count = 0;
            for (int j = (int)sorted.size() - 1; j >= 0; --j) {
                if (coarse) ++count;
                my_sig = sorted[j];
                frac += my_sig;
                if (frac > c.adaptive_fraction * sum_w[i]) break;
            }
            if (coarse && count == 0) zero_sig = true;
            if (coarse) {
                nr_sign[i] = count;
                unsigned char* cs = &coarse_sig[(size_t)i * K * D * PSI * T];
                for (int k = 0; k < K; ++k)
                    for (int d = 0; d < D; ++d)
                        for (int s = 0; s < PSI; ++s)
                            for (auto& e : sub(i, k, d, s))
                                if (e.second >= my_sig) cs[(((size_t)k * D + d) * PSI + s) * T + e.first / (NOR * NOT)] = 1;
            }
            sig_w[i] = my_sig;
        }
        if (zero_sig) return fail("my_nr_significant_coarse_samples == 0");

        // ---- optional dumps for parity tests
        if (coarse) {
            for (int i = 0; i < N; ++i) {
                if (out->min_diff2_coarse) out->min_diff2_coarse[i] = min_diff2[i];
                if (out->sum_weight_coarse) out->sum_weight_coarse[i] = sum_w[i];
                if (out->sig_weight_coarse) out->sig_weight_coarse[i] = sig_w[i];
            }
            if (out->coarse_sig) std::memcpy(out->coarse_sig, coarse_sig.data(), coarse_sig.size());
        }
        last = P;
        local_ctf_fine.swap(ctf_p);
        local_minv_fine.swap(minv_p);
        if (!coarse || npasses == 1) {
            if (out->fine_weight) {
                size_t per = (size_t)K * D * PSI * T * NOR * NOT;
                for (size_t q = 0; q < (size_t)N * per; ++q) out->fine_weight[q] = std::numeric_limits<double>::quiet_NaN();
                for (int i = 0; i < N; ++i)
                    for (int k = 0; k < K; ++k)
                        for (int d = 0; d < D; ++d)
                            for (int s = 0; s < PSI; ++s)
                                for (auto& e : sub(i, k, d, s))
                                    out->fine_weight[(size_t)i * per + (((size_t)k * D + d) * PSI + s) * T * NOR * NOT +
                                                     e.first] = e.second;
            }
        }
    }

    // ---- updateOtherParams + CoarseAndFineSearch::updateModel (src/map3d_optimizer.cpp:2059-2236, 1484-1602;
    //      2D src/map2d_optimizer.cpp:1417-1560).  Only the last-pass store contributes.
    //      In 3D without oversampling the fine store is empty (src/map3d_optimizer.cpp:1515,2156,2389).
    const bool have_store = !(is3d && os == 0);
    const Pass& P = last;
    const int F2 = P.Fsize2, NOT = P.nover_trans, NOR = P.nover_rot;
    std::vector<double> wsum_norm(N, 0.);
    std::vector<double> wsig_img((size_t)N * F2, 0.);  // per image per pixel, later mapped to shells
    std::vector<double> sumw_img(N, 0.);
    std::vector<std::vector<double>> th_pdf_class(nthreads, std::vector<double>(K, 0.));
    std::vector<std::vector<double>> th_pdf_dir(nthreads, std::vector<double>((size_t)K * D, 0.));
    std::vector<double> th_sig_off(nthreads, 0.);
    std::vector<std::vector<double>> th_prior_x(nthreads, std::vector<double>(K, 0.));
    std::vector<std::vector<double>> th_prior_y(nthreads, std::vector<double>(K, 0.));
    if (have_store) {
#pragma omp parallel for schedule(dynamic) num_threads(nthreads)
        for (int i = 0; i < N; ++i) {
            int tid = thread_id();
            const double* ctf = local_ctf_fine.data() + (size_t)i * F2;
            std::vector<double> fr(F2), fi(F2), xr(F2), xi(F2);
            window_fourier(im->fmask_re + (size_t)i * curF2, im->fmask_im + (size_t)i * curF2, cur, fr.data(),
                           fi.data(), P.size);
            std::vector<double> ref_re(F2), ref_im(F2);
            for (int k = 0; k < K; ++k) {
                const double* grid = c.grids + 2 * gridN * k;
                for (int d = 0; d < D; ++d)
                    for (int s = 0; s < PSI; ++s) {
                        int cur_orot = -1;
                        for (auto& e : sub(i, k, d, s)) {
                            if (e.second < sig_w[i]) continue;
                            double w = e.second / sum_w[i];
                            int ot = e.first % NOT, rest = e.first / NOT;
                            int orot = rest % NOR, t = rest / NOR;
                            // sigma2 / norm sums
                            if (orot != cur_orot) {
                                double A[3][3];
                                if (is3d) {
                                    const double* pose = P.os ? c.pose_fine + 3 * (((size_t)d * PSI + s) * NOR + orot)
                                                              : c.pose_coarse + 3 * ((size_t)d * PSI + s);
                                    euler_angles2matrix(pose[0], pose[1], pose[2], A);
                                    project3d(grid, pad, r_max, A, P.size, ref_re.data(), ref_im.data());
                                } else {
                                    double psi = P.os ? c.pose_fine[(size_t)s * NOR + orot] : c.pose_coarse[s];
                                    euler_angles2matrix(0, 0, psi, A);
                                    rotate2d(grid, pad, r_max, A, P.size, ref_re.data(), ref_im.data());
                                }
                                if (c.do_ctf_correction && c.refs_are_ctf_corrected)
                                    for (int n = 0; n < F2; ++n) {
                                        ref_re[n] *= ctf[n];
                                        ref_im[n] *= ctf[n];
                                    }
                                cur_orot = orot;
                            }
                            shift_image(fr.data(), fi.data(), xr.data(), xi.data(), P.size, P.sx[t * NOT + ot],
                                        P.sy[t * NOT + ot], c.ori_size);
                            double sum_wdiff2 = 0.;
                            double* ws = &wsig_img[(size_t)i * F2];
                            for (int n = 0; n < F2; ++n) {
                                if (mres_fine[n] > -1) {
                                    double dr = ref_re[n] - xr[n], di = ref_im[n] - xi[n];
                                    double wd = w * (dr * dr + di * di);
                                    ws[n] += wd;
                                    sum_wdiff2 += wd;
                                }
                            }
                            wsum_norm[i] += sum_wdiff2;
                            // pdf / offset sums (src/map3d_optimizer.cpp:2141-2177)
                            sumw_img[i] += w;
                            th_pdf_class[tid][k] += w;
                            th_pdf_dir[tid][(size_t)k * D + d] += w;
                            double offx = c.prior_offx[k] - im->old_offx[i] - P.sx[t * NOT + ot];
                            double offy = c.prior_offy[k] - im->old_offy[i] - P.sy[t * NOT + ot];
                            th_sig_off[tid] += w * (offx * offx + offy * offy);
                            if (!is3d) {
                                th_prior_x[tid][k] += w * (im->old_offx[i] + P.sx[t * NOT + ot]);
                                th_prior_y[tid][k] += w * (im->old_offy[i] + P.sy[t * NOT + ot]);
                            }
                        }
                    }
            }
        }
    }
    // reduce to groups / shells
    for (int i = 0; i < N; ++i) {
        int g = im->group[i];
        for (int n = 0; n < F2; ++n) {
            int ires = mres_fine[n];
            if (ires > -1) out->wsum_sigma2_noise[(size_t)g * ori_Fsize + ires] += wsig_img[(size_t)i * F2 + n];
        }
        out->wsum_sumw_group[g] += sumw_img[i];
    }
    for (int tt = 0; tt < nthreads; ++tt) {
        out->wsum_scalars[0] += th_sig_off[tt];
        for (int k = 0; k < K; ++k) out->wsum_pdf_class[k] += th_pdf_class[tt][k];
        for (int k = 0; k < K; ++k) {
            if (out->wsum_prior_offx) out->wsum_prior_offx[k] += th_prior_x[tt][k];
            if (out->wsum_prior_offy) out->wsum_prior_offy[k] += th_prior_y[tt][k];
        }
        for (size_t q = 0; q < (size_t)K * D; ++q) out->wsum_pdf_direction[q] += th_pdf_dir[tt][q];
    }

    // ---- backProjection (src/map3d_optimizer.cpp:2246-2481; 2D src/map2d_optimizer.cpp:1666-1873)
    if (have_store) {
        std::vector<std::vector<double>> th_data(nthreads), th_weight(nthreads);
        for (int k = 0; k < K; ++k) {
#pragma omp parallel for schedule(dynamic) num_threads(nthreads)
            for (int i = 0; i < N; ++i) {
                int tid = thread_id();
                if (th_data[tid].empty()) {
                    th_data[tid].assign(2 * gridN, 0.);
                    th_weight[tid].assign(gridN, 0.);
                }
                const int g = im->group[i];
                std::vector<double> minv(local_minv_fine.begin() + (size_t)i * F2,
                                         local_minv_fine.begin() + (size_t)(i + 1) * F2);
                std::vector<double> ctf(local_ctf_fine.begin() + (size_t)i * F2,
                                        local_ctf_fine.begin() + (size_t)(i + 1) * F2);
                minv[0] = 1. / (c.sigma2_fudge * c.sigma2_noise[(size_t)g * ori_Fsize + 0]);
                if (!c.do_map)
                    for (int n = 0; n < F2; ++n) minv[n] = 1.0;
                if (!c.do_ctf_correction)
                    for (int n = 0; n < F2; ++n) ctf[n] = 1.0;
                std::vector<double> mc(F2), mc2(F2), nr(F2), ni(F2);
                for (int n = 0; n < F2; ++n) mc[n] = minv[n] * ctf[n];
                window_fourier(im->fnomask_re + (size_t)i * curF2, im->fnomask_im + (size_t)i * curF2, cur, nr.data(),
                               ni.data(), P.size);
                for (int n = 0; n < F2; ++n) {
                    nr[n] *= mc[n];
                    ni[n] *= mc[n];
                    mc2[n] = mc[n] * ctf[n];
                }
                std::vector<double> xr(F2), xi(F2);
                std::vector<double> acc_re((size_t)NOR * F2), acc_im((size_t)NOR * F2), acc_w((size_t)NOR * F2);
                std::vector<char> used(NOR);
                for (int d = 0; d < D; ++d)
                    for (int s = 0; s < PSI; ++s) {
                        Sub& L = sub(i, k, d, s);
                        if (L.empty()) continue;
                        std::fill(used.begin(), used.end(), 0);
                        for (auto& e : L) {
                            if (e.second < sig_w[i]) continue;
                            double w = e.second / sum_w[i];
                            int ot = e.first % NOT, rest = e.first / NOT;
                            int orot = rest % NOR, t = rest / NOR;
                            if (!used[orot]) {
                                std::fill(acc_re.begin() + (size_t)orot * F2, acc_re.begin() + (size_t)(orot + 1) * F2, 0.);
                                std::fill(acc_im.begin() + (size_t)orot * F2, acc_im.begin() + (size_t)(orot + 1) * F2, 0.);
                                std::fill(acc_w.begin() + (size_t)orot * F2, acc_w.begin() + (size_t)(orot + 1) * F2, 0.);
                                used[orot] = 1;
                            }
                            shift_image(nr.data(), ni.data(), xr.data(), xi.data(), P.size, P.sx[t * NOT + ot],
                                        P.sy[t * NOT + ot], c.ori_size);
                            double* ar = &acc_re[(size_t)orot * F2];
                            double* ai = &acc_im[(size_t)orot * F2];
                            double* aw = &acc_w[(size_t)orot * F2];
                            for (int n = 0; n < F2; ++n) {
                                ar[n] += xr[n] * w;
                                ai[n] += xi[n] * w;
                                aw[n] += mc2[n] * w;
                            }
                        }
                        for (int orot = 0; orot < NOR; ++orot) {
                            if (!used[orot]) continue;
                            double A[3][3];
                            if (is3d) {
                                const double* pose = P.os ? c.pose_fine + 3 * (((size_t)d * PSI + s) * NOR + orot)
                                                          : c.pose_coarse + 3 * ((size_t)d * PSI + s);
                                euler_angles2matrix(pose[0], pose[1], pose[2], A);
                                backproject3d(th_data[tid].data(), th_weight[tid].data(), pad, r_max, A, P.size,
                                              &acc_re[(size_t)orot * F2], &acc_im[(size_t)orot * F2],
                                              &acc_w[(size_t)orot * F2]);
                            } else {
                                double psi = P.os ? c.pose_fine[(size_t)s * NOR + orot] : c.pose_coarse[s];
                                euler_angles2matrix(0, 0, psi, A);
                                backrotate2d(th_data[tid].data(), th_weight[tid].data(), pad, r_max, A, P.size,
                                             &acc_re[(size_t)orot * F2], &acc_im[(size_t)orot * F2],
                                             &acc_w[(size_t)orot * F2]);
                            }
                        }
                    }
            }
            // MAPModel::reduceThreadMap (src/map_model.cpp:209-230)
            double* bd = out->bp_data + 2 * gridN * k;
            double* bw = out->bp_weight + gridN * k;
            for (int tt = 0; tt < nthreads; ++tt) {
                if (th_data[tt].empty()) continue;
                for (size_t q = 0; q < 2 * gridN; ++q) bd[q] += th_data[tt][q];
                for (size_t q = 0; q < gridN; ++q) bw[q] += th_weight[tt][q];
                std::fill(th_data[tt].begin(), th_data[tt].end(), 0.);
                std::fill(th_weight[tt].begin(), th_weight[tt].end(), 0.);
            }
        }
    }

    // ---- storeWeightedSums (src/map3d_optimizer.cpp:1933-2057)
    for (int i = 0; i < N; ++i) {
        int g = im->group[i];
        const double* pw = im->power_img + (size_t)i * ori_Fsize;
        for (int ires = curF; ires < ori_Fsize; ++ires) {
            out->wsum_sigma2_noise[(size_t)g * ori_Fsize + ires] += pw[ires];
            wsum_norm[i] += pw[ires];
        }
        double* m = &meta[i * 9];
        m[7] = im->norm[i];
        if (c.do_norm_correction) {
            double old = im->norm[i] / c.avg_norm_correction;
            m[7] = old * sqrt(wsum_norm[i] * 2.);
            out->wsum_scalars[3] += old * sqrt(wsum_norm[i] * 2.);
        }
        double logsigma2 = 0.;
        for (int n = 0; n < curF2; ++n) {
            int ires = mres_fine[n];
            if (ires > 0) logsigma2 += log(2. * kPi * c.sigma2_noise[(size_t)g * ori_Fsize + ires]);
        }
        if (sum_w[i] == 0) return fail("exp_sum_weight[ipart]==0");
        double dLL = log(sum_w[i]) - min_diff2[i] - logsigma2;
        m[8] = dLL;
        out->wsum_scalars[1] += dLL;
        out->wsum_scalars[2] += m[0];
    }

    for (int i = 0; i < N; ++i) {
        if (out->min_diff2) out->min_diff2[i] = min_diff2[i];
        if (out->sum_weight) out->sum_weight[i] = sum_w[i];
        if (out->sig_weight) out->sig_weight[i] = sig_w[i];
        if (out->nr_sign) out->nr_sign[i] = nr_sign[i];
        if (out->argmax) std::memcpy(out->argmax + 6 * i, &best[6 * i], 6 * sizeof(int));
        if (out->meta) std::memcpy(out->meta + 9 * i, &meta[9 * i], 9 * sizeof(double));
        if (out->wsum_norm_correction) out->wsum_norm_correction[i] = wsum_norm[i];
    }
    return 0;
}

}  // extern "C"
