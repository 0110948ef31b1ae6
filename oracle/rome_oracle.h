/*
 * rome_oracle.h -- FP64 CPU restatement of ROME's MAP E-step and Fourier
 * reconstruction hot path.  TEST INFRASTRUCTURE ONLY: this library is the
 * checker that the HIP path is compared against, and the CPU baseline timed
 * beside it.  Nothing in the product (rome_amd/) may link or load it.
 *
 * Every function cites the reference loop it restates (paths relative to
 * the ROME 1.1.2 source tree).  Arithmetic is IEEE double with FP
 * contraction disabled (-ffp-contract=off) so that results are reproducible.
 *
 * Layout conventions (shared with the product):
 *   half-plane Fourier image of size s: n = i*(s/2+1) + j, row i is y with
 *     wrap (i <= s/2 -> y=i, else y=i-s), column j is x in [0, s/2];
 *   projector / backprojector grid: [z][y][x], dims (pad, pad, pad/2+1),
 *     complex interleaved (re, im), origin at (pad/2, pad/2, 0).
 */
#ifndef ROME_ORACLE_H
#define ROME_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- sampling (src/sampling.cpp, Healpix_2.15a/healpix_base.cc) ---- */
int or_healpix_pix2ang_nest(int order, double* z, double* phi);     /* npix outputs */
int or_healpix_nest2xyf(int order, int pix, int* ix, int* iy, int* face);
int or_healpix_xyf2nest(int order, int ix, int iy, int face);
int or_translations(double offset_range, double offset_step, double* tx, double* ty); /* returns count */
void or_euler_angles2matrix(double rot, double tilt, double psi, double* A9);
void or_euler_matrix2angles(const double* A9, double* rot, double* tilt, double* psi);
/* 3D pose table: for oversampling 0, out[(dir*npsi+psi)*3 + {0,1,2}] = (rot,tilt,psi);
 * for oversampling 1: out[((dir*npsi+psi)*8 + iover)*3 + c].  Perturbation applied
 * (random_perturbation * angular sampling of the coarse grid). */
int or_pose_table3d(int order, int oversampling, double random_perturbation, double* out);
/* 2D psi table: out[psi*nover + iover] */
int or_pose_table2d(double psi_step, int oversampling, double random_perturbation, double* out);
/* shifts: out_x/out_y[t*nover + iover] perturbed by random_perturbation*offset_step */
int or_shift_table(double offset_range, double offset_step, int oversampling,
                   double random_perturbation, double* out_x, double* out_y);

/* ---- image helpers (src/image.h) ---- */
void or_shift_image(const double* in_re, const double* in_im, double* out_re, double* out_im,
                    int size, double shift_x, double shift_y, int ori_size);
void or_window_fourier(const double* in_re, const double* in_im, int in_size,
                       double* out_re, double* out_im, int out_size);
void or_mresol(int size, int* out);   /* Mresol_fine/coarse (src/map_model.cpp:545-575) */

/* ---- projector / backprojector (src/reconstruct.cpp) ---- */
void or_project3d(const double* grid, int pad, int r_max, const double* A9, int f2d_size,
                  double* out_re, double* out_im);
void or_rotate2d(const double* grid, int pad, int r_max, const double* A9, int f2d_size,
                 double* out_re, double* out_im);
void or_backproject3d(double* grid, double* weight, int pad, int r_max, const double* A9, int f2d_size,
                      const double* in_re, const double* in_im, const double* mweight);
void or_backrotate2d(double* grid, double* weight, int pad, int r_max, const double* A9, int f2d_size,
                     const double* in_re, const double* in_im, const double* mweight);

/* ---- E-step ---- */
typedef struct {
    int ori_size, current_size, coarse_size;
    int nr_classes, nr_groups;
    int is_3d;                    /* 1: map3d path, 0: map2d path */
    int adaptive_oversampling;    /* 0 or 1 */
    int nr_dir, nr_psi, nr_trans; /* coarse grid (nr_dir = 1 in 2D) */
    double adaptive_fraction, sigma2_fudge, sigma2_offset, avg_norm_correction;
    int do_ctf_correction, refs_are_ctf_corrected, do_norm_correction, do_map;
    int pad;                      /* projector pad size, r_max = current_size/2 */
    int nr_threads;               /* OpenMP threads (0 = default) */
    const double* sigma2_noise;   /* [G][ori/2+1] */
    const double* pdf_class;      /* [K] */
    const double* pdf_direction;  /* [K][nr_dir] */
    const double* prior_offx;     /* [K] */
    const double* prior_offy;     /* [K] */
    const double* grids;          /* [K][pad][pad][pad/2+1][2] */
    /* sampling tables */
    const double* pose_coarse;    /* 3D: [dir][psi][3]; 2D: [psi] (psi angle) */
    const double* pose_fine;      /* 3D: [dir][psi][8][3]; 2D: [psi][2] ; NULL if no oversampling */
    const double* shift_coarse_x, *shift_coarse_y;  /* [T] (perturbed) */
    const double* shift_fine_x, *shift_fine_y;      /* [T][4] (perturbed) */
    const double* trans_x, *trans_y;                /* [T] unperturbed (pdf_offset) */
} OrConfig;

typedef struct {
    int n;
    const double* fmask_re, *fmask_im;       /* [n][cur*(cur/2+1)] */
    const double* fnomask_re, *fnomask_im;   /* [n][cur*(cur/2+1)] */
    const double* fctf;                      /* [n][cur*(cur/2+1)] */
    const double* highres_Xi2;               /* [n] */
    const double* power_img;                 /* [n][ori/2+1] */
    const double* old_offx, *old_offy;       /* [n] rounded old offsets */
    const int* group;                        /* [n] 0-based */
    const double* norm;                      /* [n] metadata NORM */
    /* optional [n][K][dir][psi][T]: the coarse significance (exp_Mcoarse_significant) given instead of
     * computed; the coarse pass is skipped and the fine pass runs over these samples (NR_SIGN = their
     * count; the coarse-pass outputs are NaN).  Needs adaptive_oversampling = 1. */
    const unsigned char* coarse_sig_in;
} OrImages;

typedef struct {
    double* coarse_diff2;   /* optional [n][K][dir][psi][T] (NaN: skipped) */
    unsigned char* coarse_sig; /* optional [n][K][dir][psi][T] */
    double* fine_diff2;     /* optional [n][K][dir][psi][T][nover_rot][nover_trans] (NaN: not computed) */
    double* fine_weight;    /* optional, same shape, weights after conversion (NaN: none) */
    double* min_diff2_coarse, *sum_weight_coarse, *sig_weight_coarse; /* [n] */
    double* min_diff2, *sum_weight, *sig_weight;                      /* [n] last pass */
    int* nr_sign;           /* [n] */
    int* argmax;            /* [n][6] (k, dir, psi, iover_rot, t, iover_trans) of last pass */
    double* meta;           /* [n][9] PMAX CLASS ROT TILT PSI XOFF YOFF NORM DLL */
    double* wsum_norm_correction; /* [n] */
    double* wsum_sigma2_noise;    /* [G][ori/2+1] (accumulated) */
    double* wsum_sumw_group;      /* [G] */
    double* wsum_pdf_class;       /* [K] */
    double* wsum_pdf_direction;   /* [K][nr_dir] */
    double* wsum_scalars;         /* [4] sigma2_offset, LL, ave_Pmax, avg_norm_correction */
    double* wsum_prior_offx;      /* [K] 2D only (src/map2d_optimizer.cpp:1598-1599) */
    double* wsum_prior_offy;      /* [K] */
    double* bp_data;              /* [K][pad][pad][pad/2+1][2] (accumulated) */
    double* bp_weight;            /* [K][pad][pad][pad/2+1] (accumulated) */
} OrOutputs;

int or_estep(const OrConfig* cfg, const OrImages* img, OrOutputs* out, char* errbuf, int errlen);

/* ---- reference map (src/reconstruct.cpp:106-223) ---- */
void or_griddingcorrect(double* vol, int dim, int ndim, int padding_factor);
int or_fourier_map(const double* vol, int ori_size, int ndim, int current_size,
                   double* grid, double* power_spectrum);

#ifdef __cplusplus
}
#endif
#endif
