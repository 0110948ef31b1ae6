// Pose and translation tables of ROME's HealpixSampler / SamplingGrid, built on
// the host for the E-step (C1 symmetry).  The E-step kernels never regenerate
// poses: the caller (or these helpers) hand the perturbed tables across the
// C-ABI.
//
//   directions  : HEALPix NEST pixel centres, src/sampling.cpp:207-221
//                 (Healpix_Base::pix2ang_z_phi, Healpix_2.15a/healpix_base.cc:494-527)
//   psi         : src/sampling.cpp:180-191
//   fine poses  : 4 NEST children x 2 psi halves, src/sampling.cpp:543-580
//   perturbation: pose <- angles(A(pose) R(p,p,p)), src/sampling.cpp:510-528
//   translations: disc of radius offset_range, src/sampling.cpp:121-172
//   shifts      : src/sampling.cpp:312-349 (perturbation on every (over)shift)
#include "rome_common.h"

#include <cmath>
#include <vector>

namespace rome {

static const double kPi = 3.14159265358979323846;
static const double kHalfPi = 1.570796326794896619231321691639751442099;

static inline double deg(double r) { return r * 180. / kPi; }
static inline double rad(double d) { return d * kPi / 180.; }

// Morton de-interleave: even bits -> x, odd bits -> y.
static int unspread(int v) {
    v &= 0x55555555;
    v = (v | (v >> 1)) & 0x33333333;
    v = (v | (v >> 2)) & 0x0f0f0f0f;
    v = (v | (v >> 4)) & 0x00ff00ff;
    v = (v | (v >> 8)) & 0x0000ffff;
    return v;
}
static int spread(int v) {
    v &= 0x0000ffff;
    v = (v | (v << 8)) & 0x00ff00ff;
    v = (v | (v << 4)) & 0x0f0f0f0f;
    v = (v | (v << 2)) & 0x33333333;
    v = (v | (v << 1)) & 0x55555555;
    return v;
}

void healpix_nest2xyf(int order, int pix, int& ix, int& iy, int& face) {
    face = pix >> (2 * order);
    int in_face = pix & ((1 << (2 * order)) - 1);
    ix = unspread(in_face);
    iy = unspread(in_face >> 1);
}
int healpix_xyf2nest(int order, int ix, int iy, int face) {
    return (face << (2 * order)) + (spread(ix) | (spread(iy) << 1));
}

void healpix_pix2ang(int order, int pix, double& z, double& phi) {
    static const int jrll[12] = {2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4};
    static const int jpll[12] = {1, 3, 5, 7, 0, 2, 4, 6, 1, 3, 5, 7};
    const int nside = 1 << order;
    const double fact2 = 4. / (12 * (nside << order));
    const double fact1 = (nside << 1) * fact2;
    int ix, iy, face;
    healpix_nest2xyf(order, pix, ix, iy, face);
    const int jr = (jrll[face] << order) - ix - iy - 1;
    int nr, kshift = 0;
    if (jr < nside) {
        nr = jr;
        z = 1 - nr * nr * fact2;
    } else if (jr > 3 * nside) {
        nr = 4 * nside - jr;
        z = nr * nr * fact2 - 1;
    } else {
        nr = nside;
        z = (2 * nside - jr) * fact1;
        kshift = (jr - nside) & 1;
    }
    int jp = (jpll[face] * nr + ix - iy + 1 + kshift) / 2;
    if (jp > 4 * nside) jp -= 4 * nside;
    if (jp < 1) jp += 4 * nside;
    phi = (jp - (kshift + 1) * 0.5) * (kHalfPi / nr);
}

void euler_matrix(double rot, double tilt, double psi, double A[9]) {
    const double a = rad(rot), b = rad(tilt), g = rad(psi);
    const double ca = cos(a), cb = cos(b), cg = cos(g), sa = sin(a), sb = sin(b), sg = sin(g);
    const double cc = cb * ca, cs = cb * sa, sc = sb * ca, ss = sb * sa;
    A[0] = cg * cc - sg * sa;
    A[1] = cg * cs + sg * ca;
    A[2] = -cg * sb;
    A[3] = -sg * cc - cg * sa;
    A[4] = -sg * cs + cg * ca;
    A[5] = sg * sb;
    A[6] = sc;
    A[7] = ss;
    A[8] = cb;
}

// Euler_matrix2angles (src/math.cpp:149-195).  For a rotation matrix the sign
// logic there always selects sin(beta) >= 0.
void euler_angles(const double A[9], double& rot, double& tilt, double& psi) {
    const double eps = 1.19209e-07;
    const double abs_sb = sqrt(A[2] * A[2] + A[5] * A[5]);
    double alpha, beta, gamma;
    if (abs_sb > 16 * eps) {
        gamma = atan2(A[5], -A[2]);
        alpha = atan2(A[7], A[6]);
        double sign_sb;
        if (std::fabs(sin(gamma)) < eps)
            sign_sb = (-A[2] / cos(gamma)) >= 0 ? 1 : -1;
        else
            sign_sb = (sin(gamma) > 0) ? (A[5] >= 0 ? 1 : -1) : (A[5] >= 0 ? -1 : 1);
        beta = atan2(sign_sb * abs_sb, A[8]);
    } else if (A[8] >= 0) {
        alpha = 0;
        beta = 0;
        gamma = atan2(-A[3], A[0]);
    } else {
        alpha = 0;
        beta = kPi;
        gamma = atan2(A[3], -A[0]);
    }
    psi = deg(gamma);
    tilt = deg(beta);
    rot = deg(alpha);
}

static void perturb(double p, double& rot, double& tilt, double& psi) {
    if (!(std::fabs(p) > 0.)) return;
    double A[9], R[9], AR[9];
    euler_matrix(rot, tilt, psi, A);
    euler_matrix(p, p, p, R);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            AR[3 * i + j] = A[3 * i] * R[j] + A[3 * i + 1] * R[3 + j] + A[3 * i + 2] * R[6 + j];
    euler_angles(AR, rot, tilt, psi);
}

static void direction(int order, int pix, double& rot, double& tilt) {
    double z, phi;
    healpix_pix2ang(order, pix, z, phi);
    rot = deg(phi);
    tilt = acos(z) * 180. / kPi;
    if (tilt < 0.) {  // checkDirection, src/sampling.h:59-75
        tilt = -tilt;
        rot += 180.;
    }
    while (rot > 180. || rot < -180.) rot += (rot > 180.) ? -360. : 360.;
}

int pose_table3d(int order, int oversampling, double perturbation, double* out) {
    if (order < 0 || order > 12 || oversampling < 0 || oversampling > 1) return -1;
    const int ndir = 12 << (2 * order);
    const double ang = 360. / (6 * round(std::pow(2., order)));
    const int npsi = (int)ceil(360. / ang);
    const double psi_step = 360. / (double)npsi;
    const double p = perturbation * ang;
    if (oversampling == 0) {
        for (int d = 0; d < ndir; ++d) {
            double r0, t0;
            direction(order, d, r0, t0);
            for (int s = 0; s < npsi; ++s) {
                double r = r0, t = t0, ps = s * psi_step;
                perturb(p, r, t, ps);
                double* o = out + 3 * ((size_t)d * npsi + s);
                o[0] = r, o[1] = t, o[2] = ps;
            }
        }
        return ndir * npsi;
    }
    for (int d = 0; d < ndir; ++d) {
        int x, y, face;
        healpix_nest2xyf(order, d, x, y, face);
        double cr[4], ct[4];
        int c = 0;
        for (int j = 2 * y; j < 2 * y + 2; ++j)
            for (int i = 2 * x; i < 2 * x + 2; ++i, ++c) direction(order + 1, healpix_xyf2nest(order + 1, i, j, face), cr[c], ct[c]);
        for (int s = 0; s < npsi; ++s)
            for (int c2 = 0; c2 < 4; ++c2)
                for (int h = 0; h < 2; ++h) {
                    double r = cr[c2], t = ct[c2];
                    double ps = s * psi_step - 0.5 * psi_step + (0.5 + h) * psi_step / 2;
                    perturb(p, r, t, ps);
                    double* o = out + 3 * (((size_t)d * npsi + s) * 8 + c2 * 2 + h);
                    o[0] = r, o[1] = t, o[2] = ps;
                }
    }
    return ndir * npsi * 8;
}

int pose_table2d(double psi_step, int oversampling, double perturbation, double* out) {
    if (!(psi_step > 0) || oversampling < 0 || oversampling > 1) return -1;
    const int npsi = (int)ceil(360. / psi_step);
    psi_step = 360. / (double)npsi;
    const int nover = 1 << oversampling;
    const bool pert = std::fabs(perturbation) > 0.;
    for (int s = 0; s < npsi; ++s)
        for (int h = 0; h < nover; ++h) {
            double v = oversampling ? s * psi_step - 0.5 * psi_step + (0.5 + h) * psi_step / nover : s * psi_step;
            if (pert) v += perturbation * psi_step;
            out[s * nover + h] = v;
        }
    return npsi * nover;
}

int translations(double range, double step, double* tx, double* ty) {
    if (!(step > 0) || range < 0) return -1;
    const int maxr = (int)ceil(range / step);
    int n = 0;
    for (int ix = -maxr; ix <= maxr; ++ix)
        for (int iy = -maxr; iy <= maxr; ++iy) {
            const double xo = ix * step, yo = iy * step;
            if (xo * xo + yo * yo <= range * range) {
                if (tx) tx[n] = xo;
                if (ty) ty[n] = yo;
                ++n;
            }
        }
    return n;
}

int shift_table(double range, double step, int oversampling, double perturbation, double* ox, double* oy) {
    const int T = translations(range, step, nullptr, nullptr);
    if (T < 0 || oversampling < 0 || oversampling > 1) return -1;
    std::vector<double> tx(T), ty(T);
    translations(range, step, tx.data(), ty.data());
    const int nos = 1 << oversampling, nover = nos * nos;
    const bool pert = std::fabs(perturbation) > 0.;
    const double p = perturbation * step;
    for (int t = 0; t < T; ++t)
        for (int iy = 0; iy < nos; ++iy)
            for (int ix = 0; ix < nos; ++ix) {
                double x = tx[t], y = ty[t];
                if (oversampling) {
                    y = ty[t] - 0.5 * step + (0.5 + iy) * step / nos;
                    x = tx[t] - 0.5 * step + (0.5 + ix) * step / nos;
                }
                if (pert) {
                    x += p;
                    y += p;
                }
                ox[t * nover + iy * nos + ix] = x;
                oy[t * nover + iy * nos + ix] = y;
            }
    return T * nover;
}

}  // namespace rome
