// Kernel argument blocks and launch wrappers of the E-step engine (kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rome_common.h"

namespace rome {

constexpr int kMaxKT = 1024;    // classes x translations held in LDS by the weight kernels
constexpr int kMaxT = 128;      // translations per task held in LDS by the post kernel
constexpr int kMaxShell = 512;  // ori_size/2+1 upper bound
constexpr int kMaxK = 64;       // classes

struct PrepArgs {
    // inputs
    const float2* fmask;    // [n][F2f]
    const float2* fnomask;  // [n][F2f]
    const float* fctf;      // [n][F2f]
    const int* group;       // [n]
    const double* hxi2;     // [n]
    const double* sigma2;   // [G][ori_Fsize]
    int ori_Fsize, F2f;
    double sigma2_fudge;
    int ctf_in_diff, do_ctf, do_map, have_fine;
    // coarse pixel set
    int Pc, Pc_pad, T, Npad;
    const int* c_nfine;
    const int* c_ires;
    const float2* Sc;       // [T][Pc]
    // fine pixel set
    int Pf;
    const int* f_nfine;
    const int* f_ires;
    // outputs
    float* WcT;             // [2*Pc_pad][Npad]
    float* MC2c;            // [Kq][mc2_ld] (transposed: A-term operand; rows >= Pc, columns >= n zero)
    int mc2_ld;
    double* C2c;            // [n]  1/2 (sum Minv|X|^2 + hxi2), coarse
    float2* Xf;             // [n][Pf]
    float* Minvf;           // [n][Pf]
    float* Ctff;            // [n][Pf]
    double* C2f;            // [n]
    float2* NMf;            // [n][Pf]
    float* MC2bp;           // [n][Pf]
};

// Window of the coarse weights: entries with u = d - ln(pdf_direction * pdf_offset) below
// min u + kWeightWindow nats are the only ones whose weights reach the sums (beyond it a weight is
// below exp(-48) of the largest).
constexpr double kWeightWindow = 64.0;

// Single pass: the FP32 cross term marks every sample whose weight is within this many nats of the
// significance threshold; those get their diff2 recomputed exactly before weights and sums are formed.
// The FP32 cross term's error (about 1e-4 nat, below 1e-2 nat at the largest sizes) is far inside it,
// so the exact significant set lies inside the marked one.
constexpr double kRecomputeMargin = 2.0;

// Mixed arms: the BF16 / FP16 cross term marks the samples within this many nats of its significance
// threshold for the split three-product recompute, which then decides significance.  The rounding of the
// low-precision operands moves a diff2 by an amount that grows with the image's signal: about 0.04 nat
// for BF16 (2^-8 per operand over a few hundred products) on a 128^2 image at SNR 0.05, 4.2 nats on a
// 64^2 image at SNR 1; FP16 an eighth of that.  The margins, 12 nats (BF16) and 2 nats (FP16), exceed
// twice the largest error measured (the threshold's and the sample's); the mixed-arm tests check them
// against the measured coarse error of every case.
constexpr double kLowpRecomputeMargin = 2.0;       // FP16
constexpr double kLowpRecomputeMarginBF16 = 12.0;  // BF16

// Coarse list output of the cross-term GEMMs (instead of the full pose term E): per image the
// entries e = t*Mtot + m that can lie within kWeightWindow of the image's minimum u, with their E
// value.  With v = E - ln pdf_direction(m) and D = max - min of ln pdf_offset over the image's
// (class, translation) pairs, u = v - ln pdf_offset lies in [v - max, v - min], so every entry with
// u < min u + W has v < min v + D + W: the kernel keeps those, against a running bound of min v
// (every tile lowers it), a superset of the window.  Entries past `cap` are counted but not stored
// (the caller then reruns with E).
struct ListArgs {
    const float* lpd;     // [Mpad] ln pdf_direction per rotation (-inf: no prior / padding)
    const float* reach;   // [n] D + W + slack per image (+inf when some pdf_offset is zero)
    int* vbound;          // [n] ordered-int key of an upper bound of min v (INT_MAX before the pass)
    unsigned* ent;        // [n][cap]
    float* val;           // [n][cap]
    int* count;           // [n] entries found (zero before the pass)
    int cap;
};

struct GemmArgs {
    const float* A;   // [K][Mpad]
    const float* B;   // [K][Npad]
    const float* At;  // [n][Mpad]
    float* E;         // [n][T][Mtot] (unused with a list)
    const unsigned char* valid;  // [Mpad] rotation has a positive prior (0 in the padding)
    int* emin_key;    // [n] ordered-int key of the per-image minimum of E (init INT_MAX)
    int K, Mpad, Npad, Mtot, Ncols, T;
    int sband;        // m-tiles per L2-resident sub-band of the XCD-banded tile order (0: plain order)
    ListArgs list;    // list.count != nullptr: list output instead of E
};

// Host side: the largest divisor of the per-XCD band of m-tiles whose slice of the rotation operand
// (bytes_per_mtile each) fits in `budget` bytes of L2; 0 disables the banded order.
inline int gemm_sub_band(int nbm, size_t bytes_per_mtile, size_t budget) {
    if (nbm % 8) return 0;
    const int band = nbm / 8;
    for (int s = band; s >= 1; --s)
        if (band % s == 0 && (size_t)s * bytes_per_mtile <= budget) return s;
    return 1;
}

// Mixed-precision coarse cross term (gemm_lowp.hip).  Operands are BF16 or FP16, K-contiguous rows.
struct LowpGemmArgs {
    const void* A;            // [Mpad][K16] scaled coarse projections
    const void* B;            // [Npad][K16] scaled Minv*CTF*X*S, row = img*T + t
    const float* At;          // [n][Mpad]
    const float* inv_scale;   // [n] 1 / (sF * s_img)
    float* E;                 // [n][T][Mtot] (unused with a list)
    const unsigned char* valid;
    int* emin_key;
    int K16, Mpad, Npad, Mtot, Ncols, T;
    int sband;
    ListArgs list;            // list.count != nullptr: list output instead of E
    // paired cross term (psih > 0): A holds only the rotations with psi in the first half turn, Mh rows;
    // rotation m + psih (psi + 180 deg) has the conjugate projection, so with P = sum Re*Re and
    // Q = sum Im*Im over the K16/2 halves, C(m) = P + Q and C(m + psih) = P - Q: half the MFMA work
    int Mh, psih;
};

struct PrepLowpArgs {
    const float2* fmask;
    const float* fctf;
    const int* group;
    const double* sigma2;
    int ori_Fsize, F2f;
    double sigma2_fudge;
    int ctf_in_diff;
    int Pc, Pc_pad, T, K16;
    const int* c_nfine;
    const int* c_ires;
    const float2* Sc;
    const int* fct_absmax;    // float bits of max |FcT| (sets the table's scale sF)
    void* W16;                // [Npad][K16]
    void* W16lo;              // [Npad][K16] rounding remainder of W16 (split recompute), or null
    float* inv_scale;         // [n]
};

// Split three-product recompute of the marked coarse samples (mixed arms): task lists of the marked
// samples as the fine pass builds them (tasks = (image, rotation m), blocks = (task, t)), operands hi / lo.
struct RecomputeArgs {
    const long long* task_off;  // [n + 1]
    const int* task_m;
    const int* task_nt;
    const long long* task_blk0;
    const int* blk_t;
    const void *Ahi, *Alo;      // [rows][K16] coarse table (half table when psih > 0)
    int psih;
    const void *Whi, *Wlo;      // [n*T][K16] image operand
    const float* inv_scale;     // [n]
    const float* At;            // [n][Mpad]
    const double* C2c;          // [n]
    int K16, Mpad, T;
    double* D2;                 // [blocks] diff2 of the marked samples
};

struct WeightParams {
    int K, D, PSI, R, Mtot, T;
    unsigned long long magicM;  // ceil(2^48 / Mtot): e / Mtot == (e * magicM) >> 48 (e < Mtot * T, Mtot^2 T < 2^48)
    const double* pdf_m;     // [Mtot] pdf_direction of the rotation's (class, direction); 0 if skipped
    const double* lpdf_m;    // [Mtot] log(pdf_m) or -inf
    double lpdf_max;         // max over rotations of lpdf_m
    const int* k_m;          // [Mtot] class of the rotation
    const float2* mrow;      // [Mtot] {(float) lpdf_m or -inf, k*T as int bits}: FP32 pre-test of the scan
    const double* trans_x;   // [T] unperturbed translations
    const double* trans_y;
    const double* old_offx;  // [n]
    const double* old_offy;
    const double* prior_x;   // [K]
    const double* prior_y;
    double sigma2_offset, adaptive_fraction;
    // Significance marking: 0 marks the samples with w >= threshold; m > 0 marks every sample with
    // w >= threshold * exp(-m) (the superset a single pass recomputes exactly).
    double mark_margin;
};

struct WeightOutArgs {
    double* dmin;
    double* sumw;
    double* maxw;
    long long* amax;
    double* thr;
    int* nsig;
    int* status;
    double* tail;  // optional: weight mass of the entries left unmarked (sum - marked), anchored at dmin
};

// Weight mass that a recomputed pass does not hold (the unmarked entries of the pass before it) and
// the minimum diff2 it is anchored at; added to the pass's own sum after re-anchoring.
struct TailArgs {
    const double* mass;
    const double* dmin;
};

struct FineArgs {
    const int* task_img;
    const int* task_m;
    const long long* task_blk0;
    const int* task_nt;
    const int* blk_t;
    const int* m2u;
    const float2* Pfine;  // [U][NOR][Pf]
    const float2* Xf;
    const float* Minvf;
    const float* Ctff;
    const float2* Sf;     // [T*NOT][Pf]
    const double* C2f;
    double* D2;           // [blocks][NOR*NOT]
    int Pf;
    // device-driven pools (all null otherwise): task total on the device, work-queue head of the tasks
    // past the grid (zero at launch), pool abort flag (non-zero: the pool is rerun, nothing is computed)
    const long long* ntask_dev = nullptr;
    int* queue = nullptr;
    const int* abort = nullptr;
};

struct PostArgs {
    int K, D, PSI, R, is3d, Pf, oriF, pad, padF, c, r_max;
    size_t grid_stride;
    const int* group;
    const double* sumw;
    const double* thr;
    const double* dmin;
    const long long* task_off;  // [n+1]
    const int* task_img;        // [tasks]
    const int* task_m;
    const int* task_nt;
    const long long* task_blk0;
    const int* blk_t;
    const int* m2u;
    const double* D2;
    const double* pdf_dir;
    const double* trans_x;
    const double* trans_y;
    const double* shift_x;      // [T*NOT] last-pass shifts
    const double* shift_y;
    const double* old_offx;
    const double* old_offy;
    const double* prior_x;
    const double* prior_y;
    double sigma2_offset;
    const double* Ai_fine;      // [R][NOR][6]
    const float2* Pfine;
    const float2* Xf;
    const float* Ctff;
    const float2* NMf;
    const float* MC2bp;
    const float2* Sf;
    const int* f_xy;
    const int* f_ires;
    // outputs
    float* bp_data;             // [K][grid][2]
    float* bp_weight;           // [K][grid]
    double* wsum_sigma2;        // [G][oriF]
    double* wsum_pdf_class;     // [K]
    double* wsum_pdf_dir;       // [K][D]
    double* wsum_sumw_group;    // [G]
    double* wsum_prior_x;       // [K]
    double* wsum_prior_y;
    double* wsum_scalars;       // [4]
    double* wnorm;              // [n]
    // binned 3D backprojection: slices go to HBM instead of the grids (null: direct scatter)
    float2* slice_val;          // [ntask * NOR][Pf]
    double* slice_w;            // [ntask * NOR]
    // device-driven pools (null otherwise): the task total on the device, the slice buffer's capacity
    // (pools with more slices take the direct scatter) and the pool abort flag
    const long long* ntask_dev = nullptr;
    long long slice_cap = 0;
    const int* abort = nullptr;
    unsigned long long* xq = nullptr;  // 8 per-XCD work counters of the per-slice kernel (zero at launch)
};

// Binned 3D backprojection (backproject.hip)
struct BpArgs {
    int K, Pf, NOR, R, pad, padF, c, r_max;
    size_t grid_stride;
    int nbx, nby, nbz, nb_class;  // bricks of 8 base voxels per axis (x over padF, y/z over pad)
    const int* task_img;
    const int* task_m;
    const double* Ai_fine;        // [R][NOR][6]
    const int* f_xy;
    const float* MC2bp;           // [n][Pf]
    const double* slice_w;        // [S] sum of the slice's normalised significant weights (0: none)
    const float2* slice_val;      // [S][Pf]
    unsigned* count;              // [nb]
    unsigned* off;                // [nb + 1]
    unsigned* coff;               // [nb + 1] chunks of the brick kernel per brick, prefix sums
    unsigned* cursor;             // [nb]
    unsigned long long* rec;      // [records] slice << 32 | pixel
    float* bp_data;
    float* bp_weight;
    // device-driven pools (null otherwise): slices = task total on the device x NOR; more than
    // slice_cap (the direct scatter took the pool) or a set abort flag: nothing to do
    const long long* ntask_dev = nullptr;
    long long slice_cap = 0;
    const int* abort = nullptr;
};

// Image preparation of a pool (setFourierTransforms), FP64; any output pointer may be null.
struct ImagePrepArgs {
    int ori, cur;
    const double* images;            // [n][ori][ori] real space, row y then x
    const double* norm;              // [n] NORM
    double avg_norm;
    int do_norm;
    const double* xoff;              // [n] old offsets (rounded here)
    const double* yoff;
    double* old_offx;                // [n] rounded offsets out
    double* old_offy;
    double* timg;                    // [n][ori][ori] translated image
    double* cbuf;                    // [n][ori][ori] centred FFT input
    const double* noise;             // [n][ori][ori] background fill outside the mask, or null
    double radius, cosine_width;     // mask: particle_diameter / (2 pixel_size), width_mask_edge
    const double2* F1;               // [n][ori][ori/2+1] unnormalised transform of the unmasked image
    const double2* F2;               // same, masked image
    double2* fnomask;                // [n][cur][cur/2+1]
    double2* fmask;
    float2* fnomask32;
    float2* fmask32;
    double* power;                   // [n][ori/2+1]
    double* hxi2;                    // [n]
    const double* ctf;               // [n][7] defocus U, V, angle, voltage, Cs, Q0, B-factor (null: CTF 1)
    double angpix;
    int ctf_abs, ctf_flip, intact_first_peak;
    double* fctf;                    // [n][cur][cur/2+1]
    float* fctf32;
};

hipError_t launch_imageprep_pre(hipStream_t s, int n, const ImagePrepArgs& a);
hipError_t launch_imageprep_mask(hipStream_t s, int n, const ImagePrepArgs& a);
hipError_t launch_imageprep_post(hipStream_t s, int n, const ImagePrepArgs& a);

struct StoreArgs {
    int n, curF, oriF, do_norm;
    double avg_norm;
    const int* group;
    const double* power;      // [n][oriF]
    const double* norm;       // [n]
    const double* sumw;
    const double* dmin;
    const double* logsigma2;  // [G]
    double* wnorm;            // [n] in/out
    double* meta;             // [n][9]
    double* wsum_sigma2;
    double* wsum_scalars;
    const int* abort = nullptr;  // device-driven pools: non-zero, the pool adds nothing
};

hipError_t launch_phase_table(hipStream_t s, const double* sx, const double* sy, int nshift, const int* xy, int P, int ori,
                              const double* sintab, const double* costab, float2* out);
hipError_t launch_project_coarse(hipStream_t s, const float2* grids, size_t grid_stride, int pad, int padF, int c, int is3d,
                                 const double* Ai, int R, int Mtot, const int* xy, int Pc, int Pc_pad, int my_r_max,
                                 int Mpad, float* FcT);
// cube[K][grid_n] records of 4 float4: the 2x2x2 trilinear neighbourhood of every voxel of a 3D grid
// records per class of the neighbourhood grid (8 x 8 x 8 bricks covering pad x pad x padF)
__host__ __device__ inline size_t cube_records(int pad, int padF) {
    const size_t nbx = (padF + 7) / 8, nby = (pad + 7) / 8;
    return nbx * nby * nby * 512;
}
hipError_t launch_build_cube(hipStream_t s, const float2* grids, size_t grid_n, int K, int pad, int padF, float4* cube);
// U_dev non-null (device-driven): the unique-rotation count is read on the device (at most U, the buffer's
// capacity); one block per item covers U_guess rotations and persist_blocks blocks take the rest from 8
// per-XCD counters xq (zero at launch)
hipError_t launch_project_fine(hipStream_t s, const float2* grids, const float4* cube, size_t grid_stride, int pad,
                               int padF, int c, int is3d, const double* Ai_fine, int R, int NOR, const int* ulist, int U,
                               const int* U_dev, unsigned long long* xq, int persist_blocks, int U_guess,
                               const int* xy, int Pf, int r_max, float2* Pfine);
hipError_t launch_project_plane(hipStream_t s, const float2* grid, int pad, int padF, int c, int is3d, const double* Ai,
                                int nrot, int size, int my_r_max, float2* out);
// Allocation-free projection (noalloc.hip): rotations staged in a file-scope device array of
// kNoallocMaxRot matrices; noalloc_init(device) takes its per-device address once (at context creation).
constexpr int kNoallocMaxRot = 16384;
hipError_t noalloc_init(int device);
hipError_t project_plane_noalloc(hipStream_t s, int device, const double* A9_host, int nrot, const float2* grid,
                                 int pad, int padF, int c, int is3d, int size, int my_r_max, float2* out);
hipError_t launch_prep_images(hipStream_t s, int n, const PrepArgs& a);
// Qc[p][m] = |FcT[p][m]|^2 (Re and Im rows), zero for p in [Pc, Kq); Kq = Pc rounded up to 16
hipError_t launch_fct_norm(hipStream_t s, const float* FcT, int Pc, int Pc_pad, int Mpad, int Kq, float* Qc);
// At[img][m] = 0.5 * sum_p Qc[p][m] * MC2cT[p][img] (FP32 MFMA); Mpad and ldn multiples of 128
hipError_t launch_coarse_aterm(hipStream_t s, int n, const float* Qc, int Kq, int Mtot, int Mpad, const float* MC2cT,
                               int ldn, float* At);
hipError_t launch_coarse_gemm(hipStream_t s, const GemmArgs& a);
struct CandArgs {  // per-image candidate list of the weight pass: [n][cap]
    double* w;
    double* du;
    long long* e;
    int cap;
};

// Significant coarse samples (weight >= the image's threshold), one of two forms:
//   keys != nullptr: per image the keys m*T + t of its significant samples, unordered, in
//                    keys[img][0 .. cap); cnt[img] = how many there are (may exceed cap: the caller
//                    then redoes the pass with the mask);
//   otherwise      : mask[img][m][TW] bit t of rotation m.
constexpr int kSigCap = 8192;  // most keys the sparse task build sorts per image (in LDS)
constexpr int kSigCapDefault = 4096;  // keys per image the significance pass keeps by default
struct SigArgs {
    uint32_t* mask;
    int TW;
    int* keys;
    int* cnt;
    int cap;
};
hipError_t launch_weights_coarse(hipStream_t s, int n, const WeightParams& wp, const float* E, const double* C2c,
                                 const int* emin_key, const WeightOutArgs& wo, const SigArgs& sig,
                                 const CandArgs& cand);
// The same weights and significance from the coarse list (ListArgs); an image whose list
// overflowed gets status 8 and nothing else.
hipError_t launch_weights_list(hipStream_t s, int n, const WeightParams& wp, const ListArgs& list, const double* C2c,
                               const int* emin_key, const WeightOutArgs& wo, const SigArgs& sig,
                               const CandArgs& cand);
// reach[img] = kWeightWindow + (max - min of ln pdf_offset over (class, translation)) + slack,
// +inf when a pdf_offset is zero
hipError_t launch_list_reach(hipStream_t s, int n, const WeightParams& wp, float* reach);
hipError_t launch_weights_fine(hipStream_t s, int n, const WeightParams& wp, const double* D2, const int* blk_task,
                               const int* blk_t, const int* task_m, const long long* img_blk, int NE,
                               const WeightOutArgs& wo, const CandArgs& cand, TailArgs tail = TailArgs{nullptr, nullptr},
                               SigArgs sig = SigArgs{nullptr, 1, nullptr, nullptr, 0});
// Task lists of the fine pass from the significant samples: per image the number of tasks
// (rotations with any significant translation) and blocks (significant samples), their exclusive
// scans (task_off, blk_off [n + 1]), used[m] = 1 for every rotation some image needs (zero before)
// and the compaction of those: ulist[u] = m, m2u[m] = u, *ucount = U.  The sparse form also sorts
// each image's keys in place.
hipError_t launch_task_build(hipStream_t s, int n, const SigArgs& sig, int M, int T, int* ntask, int* nblk,
                             long long* task_off, long long* blk_off, unsigned char* used, int* ulist, int* m2u,
                             int* ucount);
// Ordered emission (rotations ascending per image, translations ascending per rotation):
// task_img/task_m/task_blk0/task_nt per task, blk_task/blk_t per block.
hipError_t launch_task_emit(hipStream_t s, int n, const SigArgs& sig, int M, int T, const long long* task_off,
                            const long long* blk_off, int* task_img, int* task_m, long long* task_blk0, int* task_nt,
                            int* blk_task, int* blk_t);
// Device-driven (a.ntask_dev set): k_fine_diff2 over `first` blocks, then the persistent work-queue kernel
// k_fine_diff2_queue (persist_blocks blocks) over the tasks past them.
hipError_t launch_fine_diff2(hipStream_t s, int ntask, int NOR, int NOT, const FineArgs& a, int persist_blocks = 0,
                             int first = 0);
hipError_t launch_post(hipStream_t s, int n, int NOR, int NOT, const PostArgs& a);
// The binned 3D form with one block per slice (hipErrorNotSupported outside NOR = 8, NOT = 4,
// Pf <= 8192): wnorm must be zero before the launch.
// The per-slice post kernel handles NOR 8 x NOT 4, 3D, Pf <= 32768 (else hipErrorNotSupported).  With
// a.ntask_dev (device-driven pools) a first launch of one block per item covers `guess` tasks and
// persist_blocks blocks take the rest from per-XCD counters.
bool post_slices_supported(int NOR, int NOT, int Pf, int is3d);
hipError_t launch_post_slices(hipStream_t s, long long ntask, int NOR, int NOT, const PostArgs& a,
                              int persist_blocks = 0, long long guess = 0);
// Coarse significance given by a caller (keys m * T + t): an image with a count outside 1 .. cap or a key
// outside [0, MT) keeps none (cnt[img] = 0) and sets flags[2] and flags[0]; after the task build has
// sorted the keys, a repeated key sets the same flags.
hipError_t launch_check_given_keys(hipStream_t s, int n, const int* keys, int cap, int* cnt, long long MT,
                                   int* flags);
hipError_t launch_check_sorted_dups(hipStream_t s, int n, const int* keys, int cap, const int* cnt, int* flags);
// Device-driven pools: flags[0] |= 1 (abort, rerun host-synchronised) on a list overflow (status 8), an
// image without significant samples (status 1), more keys than cap, or more unique rotations than ucap.
hipError_t launch_pool_check(hipStream_t s, int n, const int* status, const int* cnt, int cap, const int* ucount,
                             int ucap, int* flags);
// p[0 .. words) = value (a kernel: safe to record into a graph, unlike hipMemsetAsync nodes)
hipError_t launch_fill32(hipStream_t s, void* p, uint32_t value, size_t words);
// Gathers up to kPackSegs device arrays into one contiguous staging buffer (one launch), so that a
// group of small read-backs is one copy to pinned host memory.  Sizes and offsets are multiples of 4.
constexpr int kPackSegs = 16;
struct PackArgs {
    const void* src[kPackSegs];
    unsigned long long off[kPackSegs + 1];  // byte offsets in dst (16-aligned); off[nseg] = total
    unsigned long long len[kPackSegs];      // bytes of each source array
    int nseg;
    void* dst;
};
hipError_t launch_pack(hipStream_t s, const PackArgs& a);

// Per-image metadata of the last pass (the max-posterior sample): meta columns 0-6 (PMAX, CLASS,
// ROT, TILT, PSI, XOFF, YOFF; 7 and 8 zero, filled by the sums kernel) and argmax = (class, dir, psi,
// over_rot, trans, over_trans).  src/map3d_optimizer.cpp:1786-1805, src/map2d_optimizer.cpp:1200-1218.
// The last pass is always a block list (the fine pass, or the single pass's exact recompute of its
// significant samples, NOR = NOT = 1).
struct MetaArgs {
    int n, is3d, R, PSI, NOR, NOT, T, Mtot;
    const long long* amax;     // [n] argmax key: entry (b - img_blk[i]) * NE + o * NOT + ot of the last pass
    const long long* img_blk;  // [n + 1] blocks per image
    const int* blk_task;
    const int* blk_t;
    const int* task_m;
    const double *maxw, *sumw, *offx, *offy;
    const double* pose;        // poses of the last pass [R][NOR][3 or 1]
    const double *sx, *sy;     // shifts of the last pass [T][NOT]
    double* meta;              // [n][9]
    int* argmax;               // [n][6]
};
hipError_t launch_meta(hipStream_t s, const MetaArgs& a);
// Whole images scattered into one class grid at their own poses (the offline reconstruction).
struct BpImageArgs {
    const float2* X;      // [n][F2] transforms (current_size half-plane)
    const float* ctf;     // [n][F2]
    const float2* phase;  // [n][Pf] tabulated shift phase of each image's offset
    const double* Ai;     // [n][6] 2 A^T restricted to (x, y)
    int F2, Pf;
    const int* f_nfine;
    const int* f_xy;
    int is3d, pad, padF, c, r_max;
    float* bp_data;       // the class's grid [pad][pad][padF] float2 (interleaved)
    float* bp_weight;
};
hipError_t launch_bp_images(hipStream_t s, int n, const BpImageArgs& a);
hipError_t launch_store_sums(hipStream_t s, const StoreArgs& a);
// With a.ntask_dev (device-driven pools) S is read on the device and persist_blocks blocks stride over
// the slices (one pixel block per grid row) and over the brick chunks.
hipError_t launch_backproject_binned(hipStream_t s, long long S, const BpArgs& a, int persist_blocks = 0);
// M-step reconstruction (reconstruct.hip)
hipError_t rc_fold(hipStream_t s, double2* data, double* weight, int Dz, int P, int PF);
hipError_t rc_decenter_real(hipStream_t s, const double* cen, double* out, int Dz, int P, int PF, int max_r2);
hipError_t rc_spectrum(hipStream_t s, const double* Fw, int Dz, int P, int PF, int max_r2, int pf, double osc, int oriF,
                       double* s0, double* s1);
hipError_t rc_map(hipStream_t s, double* Fw, int Dz, int P, int PF, int max_r2, int pf, double osc, double tau2_fudge,
                  const double* tau2, int minres, int oriF, double* dvp, double* cnt, int* bad);
hipError_t rc_normalise(hipStream_t s, double* Fw, double2* data, size_t n, double norm);
hipError_t rc_newweight_init(hipStream_t s, double* Fnw, int Dz, int P, int PF, int max_r2);
hipError_t rc_fconv(hipStream_t s, const double* Fnw, const double* Fw, double2* Fc, size_t n);
hipError_t rc_blob(hipStream_t s, double* M, int Dz, int P, int ori, int pf, const double* tab, int ntab,
                   double sampling, double norm);
hipError_t rc_newweight(hipStream_t s, const double2* Fc, double* Fnw, int Dz, int P, int PF, int max_r2,
                        double size3);
hipError_t rc_final_fconv(hipStream_t s, const double2* data, const double* Fnw, double2* Fc, int Dz, int P, int PF,
                          int max_r2);
hipError_t rc_window_ft(hipStream_t s, const double2* Fc, double2* Ft, int Dz, int P, int PF, int NZ, int N, int NF);
hipError_t rc_center_window(hipStream_t s, const double* Mout, double* vol, int NZ, int N, int OZ, int ori,
                            double normfft);
hipError_t rc_bgsum(hipStream_t s, const double* vol, int OZ, int ori, double radius, double width, double* sums);
hipError_t rc_mask_grid(hipStream_t s, double* vol, int OZ, int ori, int pf, double radius, double width,
                        const double* sums);
hipError_t rc_map_spectrum(hipStream_t s, const double2* F, int OZ, int ori, double size3, double* spec, double* cnt);
hipError_t rc_from_float(hipStream_t s, const float2* d32, const float* w32, double2* d64, double* w64, size_t n);
// projector grid of a map (fourier_map.hip)
hipError_t launch_fm_pad(hipStream_t s, const double* vol, int ori, int padori, int dim3, int do_gridding, int pf,
                         double* Mc, double* corrected);
hipError_t launch_fm_window(hipStream_t s, const double2* F, int padori, int dim3, int pf, int r_max, int pad,
                            int oriF, double normfft, double2* grid64, float2* grid32, double* pw, double* cnt);

// precision: 1 BF16, 2 FP16
hipError_t launch_absmax(hipStream_t s, const float* x, size_t n, int* out);
// A16 rows from FcT (Re at k = p, Im at k = K16/2 + p): rows = Mpad (row = rotation) with psih = 0, or
// rows = Mtot/2 of the paired cross term (row h = rotation (h / psih) 2 psih + h % psih) with psih > 0
hipError_t launch_fct_to_lowp(hipStream_t s, int precision, const float* FcT, int Pc_pad, int Mpad, int Mtot,
                              int rows, int psih, const int* fct_absmax, void* A16, void* A16lo, int K16);
hipError_t launch_recompute_split(hipStream_t s, int precision, int n, const RecomputeArgs& a);
hipError_t launch_prep_lowp(hipStream_t s, int precision, int n, const PrepLowpArgs& a);
hipError_t launch_gemm_lowp(hipStream_t s, int precision, const LowpGemmArgs& a);

}  // namespace rome
