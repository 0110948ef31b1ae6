/*
 * A C caller of include/rome_gpu.h (C99, no C++): compiling it against the header and linking it
 * against librome_gpu.so pins every declaration a drop-in caller uses, and running it drives the
 * boundary the way Map3dOptimizer_old would.
 *
 *   capi_driver          host-only calls (sampling tables, the context-free model update)
 *   capi_driver gpu      one MAP3D pool on device 0: a reference map installed through
 *                        rome_fourier_map, noise-free particles projected at known fine poses and
 *                        shifts, rome_expectation, the sums, the backprojector, rome_reconstruct and
 *                        rome_update_model; every particle's max-posterior sample must be its own pose.
 *
 * Prints one line per step and "capi_driver ok" at the end; exits non-zero on the first failure.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "rome_gpu.h"

#define CHECK(cond, ...)                          \
    do {                                          \
        if (!(cond)) {                            \
            fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
            fprintf(stderr, __VA_ARGS__);         \
            fprintf(stderr, "\n");                \
            exit(1);                              \
        }                                         \
    } while (0)

#define RC(ctx, call)                                                                        \
    do {                                                                                     \
        int rc_ = (call);                                                                    \
        CHECK(rc_ == ROME_OK, "%s -> %d (%s)", #call, rc_, (ctx) ? rome_last_error(ctx) : ""); \
    } while (0)

static const double kPi = 3.14159265358979323846;

/* ZYZ Euler matrix (rot, tilt, psi in degrees), the convention of the pose tables */
static void euler_matrix(double rot, double tilt, double psi, double A[9]) {
    const double a = rot * kPi / 180., b = tilt * kPi / 180., g = psi * kPi / 180.;
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

static void host_checks(void) {
    /* HEALPix order 1: 48 directions x 12 psi; oversampled x 8 */
    double* pc = malloc(sizeof(double) * 48 * 12 * 3);
    double* pf = malloc(sizeof(double) * 48 * 12 * 8 * 3);
    CHECK(rome_sampling_pose3d(1, 0, 0.0, pc) == 48 * 12, "pose3d count");
    CHECK(rome_sampling_pose3d(1, 1, 0.0, pf) == 48 * 12 * 8, "pose3d fine count");
    CHECK(fabs(pc[3 * 1 + 2] - 30.0) < 1e-12, "psi step of order 1 is 30 deg, got %g", pc[3 * 1 + 2]);
    /* the disc of translations: 81 for range 5 step 1, 29 for range 3 */
    CHECK(rome_sampling_translations(5.0, 1.0, NULL, NULL) == 81, "disc of radius 5");
    CHECK(rome_sampling_translations(3.0, 1.0, NULL, NULL) == 29, "disc of radius 3");
    double z = 0, phi = 0;
    CHECK(rome_healpix_pix2ang(0, 0, &z, &phi) == 0 && fabs(z - 2.0 / 3.0) < 1e-15, "pix2ang order 0");
    double pose2d[36];
    CHECK(rome_sampling_pose2d(10.0, 0, 0.0, pose2d) == 36, "pose2d count");
    /* maximizationOtherParameters on hand-made sums: K = 2, one group, 2D (ori 8, 5 shells) */
    double s2[5] = {0, 8, 16, 24, 32}, sumw[1] = {4}, pc2[2] = {3, 1}, pd[2] = {3, 1}, px[2] = {1.5, -1},
           py[2] = {0, 0.5}, sc[4] = {6, -10, 2, 4};
    rome_sums sums = {s2, sumw, pc2, pd, px, py, sc};
    double o_s2[5], o_pc[2], o_pd[2], o_px[2], o_py[2], o_sc[4];
    rome_model m = {o_s2, o_pc, o_pd, o_px, o_py, o_sc};
    const double zero[2] = {0, 0};
    CHECK(rome_maximization_other(2, 8, 2, 1, 1, 1, 1.0, zero, zero, &sums, &m) == ROME_OK, "maximization_other");
    CHECK(fabs(o_pc[0] - 0.75) < 1e-15 && fabs(o_pc[1] - 0.25) < 1e-15, "pdf_class %g %g", o_pc[0], o_pc[1]);
    CHECK(fabs(o_px[0] - 0.5) < 1e-15 && fabs(o_py[1] - 0.5) < 1e-15, "prior offsets %g %g", o_px[0], o_py[1]);
    CHECK(fabs(o_sc[0] - 6.0 / 8.0) < 1e-15, "sigma2_offset %g", o_sc[0]);
    CHECK(fabs(o_sc[1] - 1.0) < 1e-15 && fabs(o_sc[3] - 0.5) < 1e-15, "avg norm / ave_Pmax %g %g", o_sc[1], o_sc[3]);
    free(pc);
    free(pf);
    printf("host calls ok\n");
}

static void gpu_checks(void) {
    enum { ORI = 32, K = 1, ORDER = 1, NDIR = 48, NPSI = 12, NIMG = 6 };
    const int F = ORI / 2 + 1, NPX = ORI * F;
    rome_ctx* ctx = NULL;
    CHECK(rome_create(0, &ctx) == ROME_OK && ctx, "rome_create");
    CHECK(rome_configure(ctx, NULL) == ROME_ERR_ARG, "a NULL configuration is an argument error");
    const int T = rome_sampling_translations(2.0, 1.0, NULL, NULL);
    CHECK(T == 13, "disc of radius 2: %d", T);
    rome_config c;
    memset(&c, 0, sizeof c);
    c.ref_dim = 3;
    c.ori_size = ORI;
    c.current_size = ORI;
    c.coarse_size = 16;
    c.nr_classes = K;
    c.nr_groups = 1;
    c.adaptive_oversampling = 1;
    c.nr_dir = NDIR;
    c.nr_psi = NPSI;
    c.nr_trans = T;
    c.adaptive_fraction = 0.999;
    c.sigma2_fudge = 1.0;
    c.sigma2_offset = 4.0;
    c.avg_norm_correction = 1.0;
    c.do_ctf_correction = 1;
    c.refs_are_ctf_corrected = 1;
    c.do_norm_correction = 1;
    c.do_map = 1;
    c.precision = ROME_PREC_FP32;
    RC(ctx, rome_configure(ctx, &c));
    printf("configured\n");
    fflush(stdout);

    double* pc = malloc(sizeof(double) * NDIR * NPSI * 3);
    double* pf = malloc(sizeof(double) * NDIR * NPSI * 8 * 3);
    double tx[13], ty[13], sx[13], sy[13], fx[13 * 4], fy[13 * 4];
    CHECK(rome_sampling_pose3d(ORDER, 0, 0.0, pc) > 0 && rome_sampling_pose3d(ORDER, 1, 0.0, pf) > 0, "poses");
    CHECK(rome_sampling_translations(2.0, 1.0, tx, ty) == T, "translations");
    CHECK(rome_sampling_shifts(2.0, 1.0, 0, 0.0, sx, sy) == T && rome_sampling_shifts(2.0, 1.0, 1, 0.0, fx, fy) == 4 * T,
          "shift tables");
    RC(ctx, rome_set_sampling(ctx, pc, pf, sx, sy, fx, fy, tx, ty));

    /* reference: twelve Gaussian blobs (x, y, z, sigma, amplitude), through the device Fourier map (which
     * installs the projector grid); with fewer blobs, low-resolution views from distant directions can look
     * alike and the coarse pass of a sharp posterior keeps the wrong one */
    double* vol = calloc((size_t)ORI * ORI * ORI, sizeof(double));
    const double blobs[12][5] = {{6, -2, 3, 2.0, 1.2},  {-5, 4, -1, 2.5, 0.8}, {1, 7, -6, 1.5, 1.0}, {-7, -6, 2, 3.0, 0.6},
                                 {3, 1, -8, 2.0, 1.4},  {-2, -7, 6, 1.8, 0.9}, {8, 5, 4, 2.2, 0.7},  {0, -3, -2, 2.8, 1.1},
                                 {-6, 2, 7, 1.6, 1.3},  {5, -8, -5, 2.4, 0.5}, {-8, 8, 0, 2.0, 1.0},  {2, 3, 8, 1.7, 0.8}};
    for (int zz = 0; zz < ORI; ++zz)
        for (int yy = 0; yy < ORI; ++yy)
            for (int xx = 0; xx < ORI; ++xx) {
                double v = 0;
                for (int b = 0; b < 12; ++b) {
                    const double dx = xx - ORI / 2 - blobs[b][0], dy = yy - ORI / 2 - blobs[b][1],
                                 dz = zz - ORI / 2 - blobs[b][2], s = blobs[b][3];
                    v += blobs[b][4] * exp(-(dx * dx + dy * dy + dz * dz) / (2 * s * s));
                }
                vol[((size_t)zz * ORI + yy) * ORI + xx] = v;
            }
    double power[ORI / 2 + 1];
    RC(ctx, rome_fourier_map(ctx, 0, vol, 1, power, NULL, NULL));
    CHECK(power[1] > 0, "power spectrum");
    printf("reference installed\n");
    fflush(stdout);

    /* particles: exact projections at fine poses, shifted so that the estimate is a fine translation */
    const int want[NIMG][4] = {{5, 3, 2, 0}, {17, 8, 7, 6}, {30, 0, 4, 12}, {41, 11, 1, 3}, {12, 6, 5, 9}, {46, 2, 0, 1}};
    double A9[NIMG * 9];
    for (int i = 0; i < NIMG; ++i) {
        const double* e = pf + 3 * (((size_t)want[i][0] * NPSI + want[i][1]) * 8 + want[i][2]);
        euler_matrix(e[0], e[1], e[2], A9 + 9 * i);
    }
    double* pre = malloc(sizeof(double) * NIMG * NPX);
    double* pim = malloc(sizeof(double) * NIMG * NPX);
    RC(ctx, rome_project(ctx, 0, ORI, NIMG, A9, pre, pim));
    double* fre = malloc(sizeof(double) * NIMG * NPX);
    double* fim = malloc(sizeof(double) * NIMG * NPX);
    double* fctf = malloc(sizeof(double) * NIMG * NPX);
    int ot_of[NIMG];
    double sig = 0;
    for (int i = 0; i < NIMG; ++i) {
        const int t = want[i][3], ot = (i * 3) % 4;
        ot_of[i] = ot;
        const double shx = fx[4 * t + ot], shy = fy[4 * t + ot];
        for (int q = 0; q < NPX; ++q) {
            const int row = q / F, x = q % F, y = row < F ? row : row - ORI;
            const double ph = 2 * kPi * (x * shx + y * shy) / ORI, cr = cos(ph), ci = sin(ph);
            const double a = pre[(size_t)i * NPX + q], b = pim[(size_t)i * NPX + q];
            fre[(size_t)i * NPX + q] = a * cr - b * ci;
            fim[(size_t)i * NPX + q] = a * ci + b * cr;
            fctf[(size_t)i * NPX + q] = 1.0;
            sig += a * a + b * b;
        }
    }
    /* noise power 1/20 of the mean signal power per pixel: sharp posteriors */
    double sigma2[ORI / 2 + 1], pdf_class[K] = {1.0}, pdf_dir[NDIR], prior[K] = {0.0};
    for (int s = 0; s < F; ++s) sigma2[s] = 0.05 * sig / (NIMG * NPX);
    for (int d = 0; d < NDIR; ++d) pdf_dir[d] = 1.0 / NDIR;
    RC(ctx, rome_set_model(ctx, sigma2, pdf_class, pdf_dir, prior, prior));
    RC(ctx, rome_reset_sums(ctx));

    double hxi2[NIMG] = {0}, offx[NIMG] = {0}, offy[NIMG] = {0}, norm[NIMG];
    double* pw = calloc((size_t)NIMG * F, sizeof(double));
    int group[NIMG] = {0};
    for (int i = 0; i < NIMG; ++i) norm[i] = 1.0;
    rome_images im = {NIMG, fre, fim, fre, fim, fctf, hxi2, pw, offx, offy, norm, group};
    double meta[NIMG * 9], dmin[NIMG], sumw[NIMG];
    int nr_sign[NIMG], amax[NIMG * 6];
    rome_results res;
    memset(&res, 0, sizeof res);
    res.meta = meta;
    res.nr_sign = nr_sign;
    res.argmax = amax;
    res.min_diff2 = dmin;
    res.sum_weight = sumw;
    RC(ctx, rome_expectation(ctx, &im, &res));
    for (int i = 0; i < NIMG; ++i) {
        const int* a = amax + 6 * i;
        CHECK(a[0] == 0 && a[1] == want[i][0] && a[2] == want[i][1] && a[3] == want[i][2] && a[4] == want[i][3] &&
                  a[5] == ot_of[i],
              "image %d: argmax (%d %d %d %d %d %d), posed at (0 %d %d %d %d %d)", i, a[0], a[1], a[2], a[3], a[4],
              a[5], want[i][0], want[i][1], want[i][2], want[i][3], ot_of[i]);
        CHECK(nr_sign[i] >= 1 && meta[9 * i] > 0 && meta[9 * i] <= 1.0 + 1e-12 && meta[9 * i + 1] == 1.0,
              "image %d: NR_SIGN %d PMAX %g CLASS %g", i, nr_sign[i], meta[9 * i], meta[9 * i + 1]);
        const double* e = pf + 3 * (((size_t)want[i][0] * NPSI + want[i][1]) * 8 + want[i][2]);
        CHECK(fabs(meta[9 * i + 4] - e[2]) < 1e-9 || fabs(fabs(meta[9 * i + 4] - e[2]) - 360.0) < 1e-9,
              "image %d: PSI %g, posed %g", i, meta[9 * i + 4], e[2]);
        CHECK(fabs(meta[9 * i + 5] - fx[4 * want[i][3] + ot_of[i]]) < 1e-9, "image %d: XOFF %g", i, meta[9 * i + 5]);
    }
    printf("expectation ok: %d particles at their own poses\n", NIMG);
    fflush(stdout);

    double ws2[ORI / 2 + 1], wsg[1], wpc[K], wpd[NDIR], wpx[K], wpy[K], wsc[4];
    rome_sums sums = {ws2, wsg, wpc, wpd, wpx, wpy, wsc};
    RC(ctx, rome_get_sums(ctx, &sums));
    CHECK(fabs(wpc[0] - NIMG) < 1e-9 && fabs(wsg[0] - NIMG) < 1e-9, "sum of normalised weights %g", wpc[0]);

    /* backprojector: 259 would be ORI 128; here pad = 2 (2 r_max + 1) + 1 = 67 */
    const int pad = 2 * (2 * (ORI / 2) + 1) + 1, padF = pad / 2 + 1;
    const size_t ng = (size_t)pad * pad * padF;
    double* bd = malloc(sizeof(double) * 2 * ng);
    double* bw = malloc(sizeof(double) * ng);
    RC(ctx, rome_get_backprojector(ctx, 0, bd, bw));
    double wsum = 0;
    for (size_t q = 0; q < ng; ++q) wsum += bw[q];
    CHECK(wsum > 0, "backprojected weight");

    rome_recon_params rp;
    memset(&rp, 0, sizeof rp);
    rp.max_iter_preweight = 10;
    rp.do_map = 0;
    rp.minres_map = 5;
    rp.tau2_fudge = 1.0;
    rp.normalise = wpc[0];
    double tau2[ORI / 2 + 1], fsc[ORI / 2 + 1], rs2[ORI / 2 + 1], dvp[ORI / 2 + 1];
    for (int s = 0; s < F; ++s) {
        tau2[s] = 1.0;
        fsc[s] = 0.0;
    }
    double* rec = malloc(sizeof(double) * ORI * ORI * ORI);
    RC(ctx, rome_reconstruct(ctx, 0, &rp, NULL, NULL, tau2, fsc, rs2, dvp, rec));
    /* six views reconstruct a map positively correlated with the reference */
    double sab = 0, saa = 0, sbb = 0, ma = 0, mb = 0;
    const size_t nv = (size_t)ORI * ORI * ORI;
    for (size_t q = 0; q < nv; ++q) {
        CHECK(isfinite(rec[q]), "map value %zu", q);
        ma += rec[q];
        mb += vol[q];
    }
    ma /= nv;
    mb /= nv;
    for (size_t q = 0; q < nv; ++q) {
        sab += (rec[q] - ma) * (vol[q] - mb);
        saa += (rec[q] - ma) * (rec[q] - ma);
        sbb += (vol[q] - mb) * (vol[q] - mb);
    }
    const double corr = sab / sqrt(saa * sbb);
    CHECK(corr > 0.3, "map correlation %g", corr);
    printf("reconstruct ok: correlation with the reference %.3f\n", corr);

    double m_s2[ORI / 2 + 1], m_pc[K], m_pd[NDIR], m_px[K], m_py[K], m_sc[4];
    rome_model m = {m_s2, m_pc, m_pd, m_px, m_py, m_sc};
    RC(ctx, rome_update_model(ctx, 0, &m));
    CHECK(fabs(m_pc[0] - 1.0) < 1e-12, "pdf_class after the update %g", m_pc[0]);
    CHECK(rome_reconstruct(ctx, K, &rp, NULL, NULL, tau2, fsc, rs2, dvp, rec) == ROME_ERR_ARG, "class out of range");
    CHECK(strlen(rome_last_error(ctx)) > 0, "error message");
    rome_destroy(ctx);
    free(pc);
    free(pf);
    free(vol);
    free(pre);
    free(pim);
    free(fre);
    free(fim);
    free(fctf);
    free(pw);
    free(bd);
    free(bw);
    free(rec);
    printf("gpu calls ok\n");
}

int main(int argc, char** argv) {
    host_checks();
    if (argc > 1 && strcmp(argv[1], "gpu") == 0) gpu_checks();
    printf("capi_driver ok\n");
    return 0;
}
