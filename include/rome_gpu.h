/*
 * rome_gpu.h -- C-ABI of the MI355X-native MAP E-step + Fourier backprojection
 * engine (librome_gpu.so).  Plain C types only; every pointer is a host pointer
 * unless the function name says "device".  Functions return 0 on success and a
 * negative code on failure; rome_last_error() then holds the message.  A context
 * is thread-compatible (one host thread at a time), never thread-safe.
 *
 * The reference (ROME 1.1.2) has no plugin API: its E-step is a set of functions in
 * namespace Map3dOptimizer_old / Map2dOptimizer_old reading namespace globals.  Each
 * entry point below names the reference interface it replaces; a drop-in caller sets
 * the same state the reference keeps in those globals and calls rome_expectation()
 * where the reference calls expectationSomeParticles() for a pool of images.
 */
#ifndef ROME_GPU_H
#define ROME_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rome_ctx rome_ctx;

enum {
    ROME_OK = 0,
    ROME_ERR_ARG = -1,        /* bad argument / not configured */
    ROME_ERR_HIP = -2,        /* HIP runtime error */
    ROME_ERR_NUMERIC = -3,    /* numeric guard (zero significant samples, zero weight sum) */
    ROME_ERR_UNSUPPORTED = -4 /* configuration outside what the engine implements */
};

/* Iteration-level configuration: the globals of src/map_optimizer_base.h:57-136 that
 * the E-step reads (current_size, coarse_size, adaptive_oversampling/fraction,
 * sigma2_fudge, do_ctf_correction, refs_are_ctf_corrected, do_norm_correction, do_map)
 * plus the sampling sizes of HealpixSampler (src/sampling.h:113-134). */
typedef struct {
    int ref_dim;               /* 3: rome_map3d path, 2: rome_map2d path */
    int ori_size, current_size, coarse_size;
    int nr_classes, nr_groups;
    int adaptive_oversampling; /* 0 or 1 */
    int nr_dir, nr_psi, nr_trans;  /* coarse grid; nr_dir = 1 in 2D */
    double adaptive_fraction, sigma2_fudge, sigma2_offset, avg_norm_correction;
    int do_ctf_correction, refs_are_ctf_corrected, do_norm_correction, do_map;
    /* Arithmetic of the coarse cross term Re(CF * conj(Minv CTF X S)) (no reference counterpart:
     * ROME computes it in FP64 on the CPU): 0 FP32 MFMA (default), 1 BF16 MFMA, 2 FP16 MFMA, both
     * with FP32 accumulation and power-of-two operand scaling.  Only the coarse pass (which picks the
     * significant samples) changes; the fine pass, weights, sums and maps keep FP32/FP64.  BF16/FP16
     * need adaptive_oversampling = 1 (ROME_ERR_UNSUPPORTED otherwise). */
    int precision;
} rome_config;

enum { ROME_PREC_FP32 = 0, ROME_PREC_BF16 = 1, ROME_PREC_FP16 = 2 };

/* One pool of prepared images, the state ParticleModel::setFourierTransforms leaves
 * (src/exp_model.h:397-639): masked / unmasked Fourier transforms windowed to
 * current_size (half-plane, n = i*(cur/2+1)+j), CTF image, highres_Xi2, power spectrum,
 * rounded old offsets, group (0-based), NORM.  FP64 host arrays, SoA like the reference. */
typedef struct {
    int n;
    const double *fmask_re, *fmask_im, *fnomask_re, *fnomask_im, *fctf; /* [n][cur*(cur/2+1)] */
    const double* highres_Xi2;  /* [n] */
    const double* power_img;    /* [n][ori/2+1] */
    const double *old_offx, *old_offy, *norm; /* [n] */
    const int* group;           /* [n] */
} rome_images;

/* Device-resident variant: interleaved FP32 complex / FP32 CTF, FP64 scalars.  The arrays are read in
 * place during the call (no copy): on the context's device, each aligned to its element size, and
 * not written by the caller until the call returns. */
typedef struct {
    int n;
    const void* fmask;   /* float2 [n][cur*(cur/2+1)] */
    const void* fnomask; /* float2 [n][cur*(cur/2+1)] */
    const float* fctf;   /* [n][cur*(cur/2+1)] */
    const double* highres_Xi2;
    const double* power_img;
    const double *old_offx, *old_offy, *norm;
    const int* group;
} rome_device_images;

/* ---- Image preparation of a pool: ParticleModel::setFourierTransforms (src/exp_model.h:397-639) ----
 * From real-space particle images [n][ori][ori] (row y, then x): norm correction (x avg_norm_correction
 * / NORM when do_norm_correction), translation by the rounded old offsets without wrapping, centred FFT
 * (divided by ori^2), window to current_size (unmasked transform); soft mask of radius
 * particle_diameter / (2 pixel_size) with a raised cosine over width_mask_edge pixels, filled with
 * `noise` (the background image the reference draws from its random generator) or, when noise is
 * NULL (-zero_mask), with the weighted background average; masked transform; power spectrum and
 * highres_Xi2 (when current_size < ori_size, else zero); CTF image (CTF::getFftwImage; all ones
 * without do_ctf_correction).  FP64 on the device with hipFFT.  Sizes and flags from rome_config. */
typedef struct {
    int n;
    const double* images;       /* [n][ori][ori] */
    const double* norm;         /* [n] NORM */
    const double *xoff, *yoff;  /* [n] XOFF, YOFF (rounded here) */
    const double* ctf;          /* [n][7] defocus U, defocus V (A), angle (deg), voltage (kV), Cs (mm), Q0, B-factor */
    const double* noise;        /* [n][ori][ori] or NULL */
} rome_raw_images;

typedef struct {
    double pixel_size, particle_diameter, width_mask_edge;
    int ctf_phase_flipped, only_flip_phases, intact_ctf_first_peak;
} rome_prep_params;

/* Host outputs (FP64, the layout of rome_images; NULL fields are skipped). */
typedef struct {
    double *fmask_re, *fmask_im, *fnomask_re, *fnomask_im, *fctf; /* [n][cur*(cur/2+1)] */
    double* highres_Xi2;                                        /* [n] */
    double* power_img;                                          /* [n][ori/2+1] */
    double *old_offx, *old_offy;                                /* [n] rounded */
} rome_prepared;

/* Device outputs (the layout of rome_device_images; NULL fields are skipped). */
typedef struct {
    void *fmask, *fnomask;      /* float2 [n][cur*(cur/2+1)] */
    float* fctf;                /* [n][cur*(cur/2+1)] */
    double *highres_Xi2, *power_img, *old_offx, *old_offy;
} rome_prepared_device;

int rome_prepare_images(rome_ctx* ctx, const rome_raw_images* in, const rome_prep_params* params,
                        rome_prepared* host_out, rome_prepared_device* device_out);

/* Per-image results (host, any pointer may be NULL).  meta columns follow the STAR
 * labels the reference writes (src/metadata.h:55): PMAX CLASS ROT TILT PSI XOFF YOFF
 * NORM DLL.  argmax = (class, dir, psi, over_rot, trans, over_trans) of the last pass. */
typedef struct {
    double* meta;            /* [n][9] */
    int* nr_sign;            /* [n] NR_SIGN of the coarse pass */
    int* argmax;             /* [n][6] */
    double* min_diff2;       /* [n] last pass */
    double* sum_weight;      /* [n] */
    double* sig_weight;      /* [n] */
    double* min_diff2_coarse;  /* [n] */
    double* sum_weight_coarse; /* [n] */
    double* sig_weight_coarse; /* [n] */
    double* wsum_norm_correction; /* [n] */
} rome_results;

/* Weighted sums of MLModel (src/ml_model.h) accumulated over rome_expectation calls. */
typedef struct {
    double* wsum_sigma2_noise;   /* [G][ori/2+1] */
    double* wsum_sumw_group;     /* [G] */
    double* wsum_pdf_class;      /* [K] */
    double* wsum_pdf_direction;  /* [K][nr_dir] */
    double* wsum_prior_offx;     /* [K] (2D) */
    double* wsum_prior_offy;     /* [K] (2D) */
    double* wsum_scalars;        /* [4] sigma2_offset, LL, ave_Pmax, avg_norm_correction */
} rome_sums;

/* ---- context ---- */
int rome_create(int device, rome_ctx** out);
void rome_destroy(rome_ctx* ctx);
const char* rome_last_error(const rome_ctx* ctx);

/* Map3dOptimizer_old::prepareExpData sizes (src/map3d_optimizer.cpp:615-714). */
int rome_configure(rome_ctx* ctx, const rome_config* cfg);

/* SamplingGrid tables (src/sampling.cpp:834-845, 927-958), perturbation already applied:
 * 3D pose_coarse [dir][psi][3] (rot, tilt, psi deg), pose_fine [dir][psi][8][3];
 * 2D pose_coarse [psi], pose_fine [psi][2].  Shifts [T] / [T][4] in pixels (perturbed);
 * trans_x/y [T] unperturbed (pdf_offset, src/map3d_optimizer.cpp:1704-1716).
 * pose_fine / shift_fine may be NULL when adaptive_oversampling == 0. */
int rome_set_sampling(rome_ctx* ctx, const double* pose_coarse, const double* pose_fine,
                      const double* shift_coarse_x, const double* shift_coarse_y,
                      const double* shift_fine_x, const double* shift_fine_y,
                      const double* trans_x, const double* trans_y);

/* MLModel state read by the E-step: sigma2_noise [G][ori/2+1], pdf_class [K],
 * pdf_direction [K][nr_dir], prior_offset{x,y}_class [K]. */
int rome_set_model(rome_ctx* ctx, const double* sigma2_noise, const double* pdf_class,
                   const double* pdf_direction, const double* prior_offx, const double* prior_offy);

/* Classes the following pools run: exp_iclass_min .. exp_iclass_max of expectationSomeParticles, which the
 * reference narrows to one class per pool in the seeding iteration (src/map3d_optimizer.cpp:823-840;
 * src/map2d_optimizer.cpp:731-758); other classes get no diff2, weight or sums, as classes without prior.
 * (-1, -1): every class (the default; rome_configure resets it). */
int rome_set_class_range(rome_ctx* ctx, int iclass_min, int iclass_max);

/* MyProjector::data of one class after computeFourierTransformMap
 * (src/reconstruct.cpp:106-223): complex interleaved FP64 [pad][pad][pad/2+1] (3D) or
 * [pad][pad/2+1] (2D), pad = 2*(2*min(cur/2,ori/2)+1)+1. */
int rome_set_reference(rome_ctx* ctx, int iclass, const double* grid);

/* MyProjector::computeFourierTransformMap (src/reconstruct.cpp:106-223) with griddingCorrect
 * (src/reconstruct.cpp:61-103) on the device, the way MAPModel::setFourierTransformMaps
 * (src/map_model.cpp:136-160) calls it: vol = the class map, [ori][ori][ori] (3D) or [ori][ori] (2D)
 * FP64; padding factor 2, trilinear, r_max = min(current_size, ori_size)/2 of rome_configure.  FFT in
 * FP64 (hipFFT D2Z).  The grid becomes reference iclass (what rome_set_reference would install);
 * power_spectrum [ori/2+1] receives the radial power (tau2 input), corrected_map the gridding-corrected
 * map (the reference corrects its input in place) and grid_out the [pad][pad][pad/2+1] (3D) or
 * [pad][pad/2+1] (2D) complex FP64 grid, interleaved; any of the three may be NULL. */
int rome_fourier_map(rome_ctx* ctx, int iclass, const double* vol, int do_gridding, double* power_spectrum,
                     double* corrected_map, double* grid_out);

/* MyBackProjector::reconstruct (src/reconstruct.cpp:1929-2522) as MAPModel::reconstruction
 * (src/map_model.cpp:233-239) calls it from maximization (src/map3d_optimizer.cpp:2507-2617): the class's
 * backprojector grids -> the new map, on the device in FP64 (hipFFT).  Grids: the context's own
 * accumulated backprojector of iclass when data_in and weight_in are NULL, else FP64 host grids in the
 * layout of rome_get_backprojector.  C1 symmetry, trilinear, padding 2, blob order 0 / radius 1.9 /
 * alpha 15 (MyBackProjector defaults, src/reconstruct.h:650-655).  tau2 [ori/2+1] is read (and
 * rewritten when update_tau2_with_fsc); fsc [ori/2+1] is read only with update_tau2_with_fsc; sigma2 and
 * data_vs_prior [ori/2+1] are written (data_vs_prior stays 0 without do_map and FSC update); vol_out is
 * [ori][ori][ori] (3D) or [ori][ori] (2D). */
typedef struct {
    int max_iter_preweight;     /* gridding iterations (rome_map3d: 10) */
    int do_map, minres_map;     /* MAPModel default minres_map 5 */
    int update_tau2_with_fsc, is_whole_instead_of_half;
    double tau2_fudge, normalise;  /* maximization passes normalise = wsum_pdf_class[iclass] */
} rome_recon_params;

int rome_reconstruct(rome_ctx* ctx, int iclass, const rome_recon_params* params, const double* data_in,
                     const double* weight_in, double* tau2, const double* fsc, double* sigma2,
                     double* data_vs_prior, double* vol_out);

/* MLModel::resetZero of the E-step sums + zeroed backprojector grids (MAPModel). */
int rome_reset_sums(rome_ctx* ctx);

/* The fine pass of a pool whose coarse significance is given (exp_Mcoarse_significant, which
 * findAllSignificantPoints sets at src/map3d_optimizer.cpp:1880-1915): getAllSquaredDifferencesFine
 * (src/map3d_optimizer.cpp:1261-1372) over those samples, then the weights, significance, sums,
 * backprojection and storeWeightedSums of the second pass of expectationSomeParticles
 * (src/map3d_optimizer.cpp:883-953), without the coarse pass.  sig_keys [n][cap] (device int32) holds
 * sig_cnt[i] (device int32, 1 .. cap, cap <= 8192) distinct keys per image, key = m * nr_trans + t with
 * m = (class * nr_dir + dir) * nr_psi + psi; an image outside that is ROME_ERR_ARG.  NR_SIGN = sig_cnt;
 * the coarse-pass results (min_diff2_coarse, sum_weight_coarse, sig_weight_coarse) are NaN.  Needs
 * adaptive_oversampling = 1. */
int rome_expectation_fine_device(rome_ctx* ctx, const rome_device_images* images, const int* sig_keys,
                                 const int* sig_cnt, int cap, rome_results* results);

/* Map3dOptimizer_old::expectationSomeParticles (src/map3d_optimizer.cpp:883-953) or
 * Map2dOptimizer_old::expectationSomeParticles (src/map2d_optimizer.cpp:802-854), with
 * storeWeightedSums, for one pool of images ("iteration >= 2" semantics: all classes). */
int rome_expectation(rome_ctx* ctx, const rome_images* images, rome_results* results);
int rome_expectation_device(rome_ctx* ctx, const rome_device_images* images, rome_results* results);

/* MLModel weighted sums (src/ml_model.cpp:91-114 layout). */
int rome_get_sums(rome_ctx* ctx, rome_sums* out);

/* MyBackProjector data / weight of one class (src/reconstruct.h:546-717), FP64 host. */
int rome_get_backprojector(rome_ctx* ctx, int iclass, double* data, double* weight);

/* Parity hooks: coarse-pass diff2 of the last pool [n][K][dir][psi][T] (NaN where the
 * reference stores no entry; needs ROME_COARSE_FULL, rome_set_coarse_mode), and the fine-pass
 * list as (key, diff2) with
 * key = ((((img*K + k)*D + dir)*PSI + psi)*T + t)*NOR*NOT + over_rot*NOT + over_trans. */
int rome_get_coarse_diff2(rome_ctx* ctx, double* out);
long long rome_get_fine_count(rome_ctx* ctx);
int rome_get_fine_list(rome_ctx* ctx, long long* keys, double* diff2);

/* MyProjector::projectOneTile / rotate2D of arbitrary rotations (A row-major 3x3 each)
 * into full half-planes of size `size`: out [nrot][size*(size/2+1)]. */
int rome_project(rome_ctx* ctx, int iclass, int size, int nrot, const double* A9,
                 double* out_re, double* out_im);

/* Projector into a device buffer (float2 [nrot][size*(size/2+1)]): synthetic particles. */
int rome_project_device(rome_ctx* ctx, int iclass, int size, int nrot, const double* A9, void* out_device);

/* Allocation-free form of rome_project_device (no counterpart in the reference): for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * The rotations are staged in a fixed device array of the library (16384 matrices at most; more is
 * ROME_ERR_ARG); no hipMalloc or hipFree is called.  Every call on one device shares that array, so
 * calls that overlap on one device (two host threads, two contexts) race. */
int rome_project_device_noalloc(rome_ctx* ctx, int iclass, int size, int nrot, const double* A9, void* out_device);

/* Execution of a pool (no counterpart in the reference, whose pools run as OpenMP loops):
 * ROME_EXEC_HOST   the host reads the fine pass's task totals back after the coarse pass and sizes every
 *                  later launch from them (two host synchronisations per pool);
 * ROME_EXEC_DEVICE the fine pass runs from buffers sized by capacities (significant samples per image x
 *                  pool size) with its counts read on the device: fixed grids (sized from the counts of
 *                  earlier pools or by rome_autotune) stride over them, the fine diff2's tasks past its
 *                  first launch go to a persistent work-queue kernel; one synchronisation per pool, when
 *                  the results are read back;
 * ROME_EXEC_GRAPH  (default) as DEVICE, and the pool is captured once as a hipGraph per (pool size, input
 *                  arrays) and replayed for later pools with the same inputs.
 * A pool whose coarse list or significant-sample keys overflow, or that needs the dense significance mask,
 * the learned proposal or profiling stage timers, runs (or is rerun) host-synchronised; the results are the
 * same in every mode. */
enum { ROME_EXEC_HOST = 0, ROME_EXEC_DEVICE = 1, ROME_EXEC_GRAPH = 2 };
int rome_set_execution(rome_ctx* ctx, int mode);
/* out6: how the last pool ran (0 host, 1 device, 2 graph replay; +8 when a device-driven pool was rerun
 * host-synchronised), graph launches, graph captures, device-driven reruns, and for the last device-driven
 * pool the host time (microseconds) spent issuing it and waiting for its results. */
int rome_get_execution_stats(rome_ctx* ctx, long long* out6);

/* Tile / grid autotuner of the device-driven kernels (no counterpart in the reference): runs the pool
 * `images` (with sig_keys / sig_cnt / cap: its fine pass from that significance, as
 * rome_expectation_fine_device; NULL: the whole E-step) device-driven with 4 candidate grids for each of the fine diff2's work-queue kernel, the
 * fine projection, the per-slice post kernel and the binned backprojection (blocks per CU: {1, 2, 3,
 * 4}, {4, 8, 12, 16}, {1, 2, 3, 4}, {8, 16, 32, 64}), twice each, times their stages with HIP events and
 * keeps the fastest grid of each kernel for the context.  chosen [4] receives the grids (blocks),
 * stage_ms [4][4] the measured stage times per candidate.  The runs add to the E-step sums: call
 * rome_reset_sums afterwards.  rome_set_persistent_grids sets the four grids directly (0: the default,
 * 2, 8, 2 and 32 blocks per CU) and forgets the task and rotation totals of earlier pools, from which
 * the first launches of the fine diff2, the projection and the post slices are sized (the next pool's
 * are then sized from 4 per image, and what exceeds them goes through the work queue / the tails). */
int rome_autotune(rome_ctx* ctx, const rome_device_images* images, const int* sig_keys, const int* sig_cnt, int cap,
                  int* chosen, double* stage_ms);
int rome_set_persistent_grids(rome_ctx* ctx, const int* blocks);

/* Measurement: stage times from HIP events recorded on the engine stream around each kernel
 * group (no counterpart in the reference, whose timers are TIMEPOINT macros, src/time.h:61-75).
 * Stages: 0 prep, 1 coarse norms, 2 coarse cross term, 3 coarse weights, 4 task build,
 * 5 fine projections, 6 fine diff2, 7 fine weights, 8 sigma2/pdf sums and slices (2D: also the
 * backprojection), 9 store sums, 10 binned 3D backprojection, 11 split three-product recompute of the
 * marked coarse samples (BF16 / FP16 arms; stage 3 then also holds the second weight pass).
 * rome_set_profiling(1) resets the accumulators; rome_get_profile returns the stage count. */
int rome_set_profiling(rome_ctx* ctx, int on);
int rome_get_profile(rome_ctx* ctx, double* stage_ms, long long* stage_calls, int nstages);
/* Last pool: n, fine tasks, fine blocks, unique rotations, coarse pixels, fine pixels,
 * rotations x classes, translations. */
int rome_get_batch_stats(rome_ctx* ctx, long long* out8);

/* Storage of the coarse pass (no counterpart in the reference, which keeps every coarse diff2 of a
 * particle in exp_Mweight, src/map3d_optimizer.cpp:1403-1608, before
 * convertSquaredDifferencesToWeights).  ROME_COARSE_LIST (default): the cross-term kernel keeps per
 * image only the entries whose u = diff2 - ln(pdf_direction pdf_offset) can lie within 64 nats of
 * the image's minimum u (the only ones whose weights reach the sums), so the full pose term never
 * goes to HBM; an image with more than list_cap such entries sends the whole pool through the full
 * path again (same results).  ROME_COARSE_FULL: the whole pose term in HBM, needed by
 * rome_get_coarse_diff2; also what a single pass (adaptive_oversampling = 0) always uses, since its
 * significant samples take their diff2 from it.  list_cap: entries per image, <= 0 for the default (all of an image's
 * coarse entries when the pool's lists fit in 1 GiB, else max(65536, a quarter of them)). */
enum { ROME_COARSE_LIST = 0, ROME_COARSE_FULL = 1 };
int rome_set_coarse_mode(rome_ctx* ctx, int mode, int list_cap);
/* Last pool: out3 = {0 list / 1 full / 2 list overflowed and the pool ran full, largest list,
 * total list entries}. */
int rome_get_coarse_stats(rome_ctx* ctx, long long* out3);

/* Form of the coarse significance the fine pass is built from (no counterpart in the reference, whose
 * Exp_Mweight stores it per sample): by default the keys of each image's significant samples, up to
 * `cap` (<= 8192) per image; a pool where some image has more goes through a dense per-sample mask
 * instead (and so do the next 8 pools).  cap = 0: the mask always; < 0: the default (4096).
 * Results are the same either way. */
int rome_set_significance_cap(rome_ctx* ctx, int cap);
/* Last pool: out2 = {0 keys / 1 mask / 2 keys overflowed and the pool used the mask, largest number of
 * significant samples of an image (keys form; -1 otherwise)}. */
int rome_get_significance_stats(rome_ctx* ctx, long long* out2);

/* ---- M-step: model parameters other than the maps ----
 * maximizationOtherParameters (src/map3d_optimizer.cpp:2620-2798; 2D src/map2d_optimizer.cpp:1985-2045)
 * for iterations without CC mode and without scale correction: from the E-step sums to
 *   pdf_class = wsum_pdf_class / sum_w, pdf_direction = wsum_pdf_direction / sum_w,
 *   prior offsets (2D) = wsum_prior_offset / wsum_pdf_class (0 for an empty class),
 *   sigma2_offset = wsum_sigma2_offset / (2 sum_w),
 *   sigma2_noise[g][n] = wsum_sigma2_noise[g][n] / (2 wsum_sumw_group[g] Npix_per_shell[n])
 *     (Npix_per_shell of MAPModel::updateImageSizeAndResolutionPointers, src/map_model.cpp:522-538),
 *   avg_norm_correction = wsum_avg_norm_correction / sum_w (when do_norm_correction),
 *   LL = wsum_LL, ave_Pmax = wsum_ave_Pmax / sum_w,
 * with sum_w = sum_k wsum_pdf_class[k].  Host FP64 arithmetic in the reference's order. */
typedef struct {
    double* sigma2_noise;   /* [G][ori/2+1] */
    double* pdf_class;      /* [K] */
    double* pdf_direction;  /* [K][nr_dir] */
    double* prior_offx;     /* [K] */
    double* prior_offy;     /* [K] */
    double* scalars;        /* [4] sigma2_offset, avg_norm_correction, LL, ave_Pmax */
} rome_model;

/* Context-free form: sums in (rome_sums layout, const), model out.  avg_norm_in and the prior
 * offsets in are kept where the reference keeps them (3D priors, do_norm_correction = 0). */
int rome_maximization_other(int ref_dim, int ori_size, int nr_classes, int nr_groups, int nr_dir,
                            int do_norm_correction, double avg_norm_in, const double* prior_offx_in,
                            const double* prior_offy_in, const rome_sums* sums, rome_model* out);

/* The same from the context's own sums and model; install != 0 makes the result the model of the
 * next rome_expectation (as rome_set_model would, plus sigma2_offset, avg_norm_correction and, with
 * CTF correction, refs_are_ctf_corrected = 1). */
int rome_update_model(rome_ctx* ctx, int install, rome_model* out);

/* Back-projection of whole images at given poses into class iclass's backprojector, the loop of the
 * offline reconstruction apps/rome_reconstruct.cpp:104-330 (MyBackProjector::backproject,
 * src/reconstruct.cpp:977-1164): each transform (rome_images layout at current_size, e.g. the fnomask of
 * rome_prepare_images) is shifted by its offset with the tabulated phase of shiftImageInFourierTransform,
 * multiplied by its CTF image (weights CTF^2; ones without do_ctf_correction or with fctf NULL) and
 * scattered into the FP32 grids.  euler [n][3] (rot, tilt, psi deg; 2D uses psi), shift [n][2] pixels.
 * rome_reconstruct then gives the map (no MAP term, as the offline tool). */
int rome_backproject_images(rome_ctx* ctx, int iclass, int n, const double* fre, const double* fim, const double* fctf,
                            const double* euler, const double* shift);

/* ---- Learned pose proposal (no counterpart in the reference, 3D two-pass only) ----
 * A network (2 * nr_coef low-resolution coefficients of the masked image -> 3 x hidden GELU -> K x C
 * cell logits + 2 shift outputs; C = the HEALPix directions one order finer than the coarse sampling x
 * nr_psi_bins in-plane bins; BF16 products, FP32 accumulation and parameters) learns the E-step's own
 * posteriors and then proposes, per image, the cells whose coarse rotations get exact diff2: the top_k
 * cells, their neighbouring coarse directions and in-plane bins.  It stands in for the exhaustive coarse
 * pass of getAllSquaredDifferencesCoarse (src/map3d_optimizer.cpp:1195-1259), whose significance
 * (src/map3d_optimizer.cpp:1812-1931) stays the judge: an image whose predicted captured mass is below
 * min_mass takes every rotation, and audit_fraction of every pool also runs the exhaustive pass; a pool
 * where an audited image's captured posterior mass is below min_audit_mass reruns exhaustively.  Needs
 * precision BF16 or FP16 (the proposed rotations' diff2 are computed as split three-product sums). */
typedef struct {
    int hidden;            /* hidden width, multiple of 32 (1024) */
    int nr_coef;           /* low-resolution coefficients of the input (812) */
    int nr_psi_bins;       /* in-plane bins per cell; divides nr_psi (4) */
    int top_k;             /* cells kept per image before the neighbours (64) */
    double min_mass;       /* predicted captured mass below which an image takes every rotation (0.9) */
    double audit_fraction; /* share of each pool that also runs the exhaustive coarse pass (0.02) */
    double min_audit_mass; /* audited captured mass below which the pool reruns exhaustively (0.99) */
    int target_cells;      /* cells per training target (32) */
    double shift_weight;   /* weight of the squared shift error in the loss (1.0) */
    int max_train;         /* capacity of the training set in images */
    unsigned long long seed;
} rome_proposal_config;

enum { ROME_PROPOSAL_OFF = 0, ROME_PROPOSAL_COLLECT = 1, ROME_PROPOSAL_USE = 2 };

/* (Re)creates the network with fresh parameters and an empty training set (needs rome_configure and
 * rome_set_sampling).  Mode COLLECT: every rome_expectation pool (exhaustive) adds its images and their
 * posteriors over cells to the training set; USE: pools run on the proposal; OFF: exhaustive. */
int rome_proposal_configure(rome_ctx* ctx, const rome_proposal_config* cfg);
int rome_proposal_mode(rome_ctx* ctx, int mode);
/* top_k, min_mass and min_audit_mass of a configured proposal, keeping its network and training set. */
int rome_proposal_set_thresholds(rome_ctx* ctx, int top_k, double min_mass, double min_audit_mass);
/* Adam over the training set: epochs x (training images / batch) steps of `batch` images (a multiple of
 * 32), shuffled per epoch; loss_per_epoch [epochs] (may be NULL) receives each epoch's mean loss. */
int rome_proposal_train(rome_ctx* ctx, int epochs, int batch, double lr, double* loss_per_epoch);
/* out8: pools, images, mean candidate rotations per image, images that took every rotation, audited
 * images, lowest audited captured mass, pools run exhaustively (an audited captured mass below
 * min_audit_mass, or more than half of the pool's images below min_mass), training images. */
int rome_proposal_stats(rome_ctx* ctx, double* out8);

/* Parameters in host FP32, row-major: W1 [hidden][2 nr_coef], W2, W3 [hidden][hidden],
 * W4 [K C + 2][hidden], biases [rows]; also the layout of gradients. */
typedef struct {
    float *W1, *b1, *W2, *b2, *W3, *b3, *W4, *b4;
} rome_proposal_params;
int rome_proposal_get_params(rome_ctx* ctx, rome_proposal_params* out);
int rome_proposal_set_params(rome_ctx* ctx, const rome_proposal_params* in);
/* Parity hooks: the network's outputs z [n][K C + 2] for host features x [n][2 nr_coef], and the loss and
 * gradients of one batch (n a multiple of 32) with dense targets [n][K C] (at most target_cells non-zero
 * per row) and shift targets [n][2]. */
int rome_proposal_forward(rome_ctx* ctx, int n, const float* x, float* z);
int rome_proposal_gradients(rome_ctx* ctx, int n, const float* x, const float* target, const float* shift,
                            double* loss, rome_proposal_params* grads);
/* Features of the last pool's images [n][2 nr_coef] and their training targets (dense [n][K C], shift
 * [n][2]) as the training set holds them, for rows row0 .. row0 + n - 1 of the set. */
int rome_proposal_training_rows(rome_ctx* ctx, int row0, int n, float* x, float* target, float* shift);

/* ---- host sampling helpers (HealpixSampler, C1) ---- */
int rome_sampling_pose3d(int healpix_order, int oversampling, double perturbation, double* out);
int rome_sampling_pose2d(double psi_step, int oversampling, double perturbation, double* out);
int rome_sampling_translations(double offset_range, double offset_step, double* tx, double* ty);
int rome_sampling_shifts(double offset_range, double offset_step, int oversampling, double perturbation,
                         double* out_x, double* out_y);
int rome_healpix_pix2ang(int order, int pix, double* z, double* phi);

#ifdef __cplusplus
}
#endif
#endif
