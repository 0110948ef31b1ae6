// Host orchestration of the MI355X E-step (C++ above the C-ABI).  Mirrors the call
// order of Map3dOptimizer_old::expectationSomeParticles (src/map3d_optimizer.cpp:883-953):
//   coarse pass  : preShiftedImages -> getAllSquaredDifferencesCoarse -> convert -> significance
//   fine pass    : getAllSquaredDifferencesFine -> convert -> significance
//   updateOtherParams + backProjection (one fused kernel), storeWeightedSums.
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/rome_gpu.h"
#include "kernels.h"
#include "proposal.h"
#include "rome_common.h"

namespace rome {

// Device allocations move when a buffer grows; a captured pool graph holds the addresses it was captured
// with, so every (re)allocation advances this epoch and graphs of an older epoch are not replayed.  While a
// pool is being captured no buffer may grow (hipMalloc is not part of a graph): reserve() then fails.
std::atomic<unsigned long long> g_alloc_epoch{0};
thread_local bool g_capturing = false;

// Pinned host memory (hipHostMalloc), grown on demand: read-backs and uploads go through it as one
// DMA each instead of staged copies of pageable memory.
struct PinnedBuf {
    void* p = nullptr;
    size_t n = 0;
    PinnedBuf() = default;
    PinnedBuf(const PinnedBuf&) = delete;
    PinnedBuf& operator=(const PinnedBuf&) = delete;
    ~PinnedBuf() {
        if (p) (void)hipHostFree(p);
    }
    hipError_t reserve(size_t bytes) {
        if (bytes <= n && p) return hipSuccess;
        if (g_capturing) return hipErrorStreamCaptureUnsupported;
        g_alloc_epoch.fetch_add(1);
        if (p) (void)hipHostFree(p);
        p = nullptr;
        n = 0;
        hipError_t e = hipHostMalloc(&p, std::max<size_t>(bytes, 64), hipHostMallocDefault);
        if (e == hipSuccess) n = std::max<size_t>(bytes, 64);
        return e;
    }
};

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    // A buffer that has to grow again (its size follows the data, e.g. the number of significant
    // poses) gets 25 % headroom: hipFree waits for the device, so every regrowth stalls the stream.
    hipError_t reserve(size_t count) {
        if (count <= n && p) return hipSuccess;
        if (g_capturing) return hipErrorStreamCaptureUnsupported;
        g_alloc_epoch.fetch_add(1);
        if (p) count += count / 4;
        release();
        size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
        hipError_t e = hipMalloc(&p, bytes);
        if (e == hipSuccess) n = std::max<size_t>(count, 1);
        return e;
    }
    // capacities (device-driven pools): exactly the size asked for, no headroom
    hipError_t reserve_exact(size_t count) {
        if (count <= n && p) return hipSuccess;
        if (g_capturing) return hipErrorStreamCaptureUnsupported;
        g_alloc_epoch.fetch_add(1);
        release();
        const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
        hipError_t e = hipMalloc(&p, bytes);
        if (e == hipSuccess) n = std::max<size_t>(count, 1);
        return e;
    }
    // for sizes that follow the data: 25 % headroom from the first allocation on
    hipError_t grow(size_t count) {
        if (count <= n && p) return hipSuccess;
        return reserve(count + count / 4);
    }
    hipError_t upload(const T* h, size_t count, hipStream_t s) {
        hipError_t e = reserve(count);
        if (e != hipSuccess) return e;
        if (count == 0) return hipSuccess;
        return hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s);
    }
};

struct Error {
    int code;
};

// Host-side results of a device-driven pool: every array the pool reads back, in one pinned block.
struct PoolOut {
    std::vector<double> meta, dmin, sumw, thr, dmin_c, sumw_c, thr_c, wnorm;
    std::vector<int> argmax, nsig, cstatus, lcount, scount;
    long long tot[2] = {0, 0};
    int U = 0, flags[2] = {0, 0};
    void size(int n) {
        for (auto* v : {&dmin, &sumw, &thr, &dmin_c, &sumw_c, &thr_c, &wnorm}) v->assign(n, 0.);
        meta.assign((size_t)9 * n, 0.);
        argmax.assign((size_t)6 * n, 0);
        for (auto* v : {&nsig, &cstatus, &lcount, &scount}) v->assign(n, 0);
    }
};

// State of the learned pose proposal (rome_proposal_*): parameters (FP32 masters, BF16 copies and their
// transposes for the backward products, Adam moments, gradients), the input and neighbour tables, the
// training set, activations of the current batch, and the per-pool candidate and audit buffers.
struct ProposalState {
    bool ready = false;
    rome_proposal_config cfg{};
    int mode = ROME_PROPOSAL_OFF;
    int K = 0, D = 0, PSI = 0, cur = 0;  // shape the network was built for
    int ncoef = 0, Fin = 0, Fin_pad = 0, H = 0, ncell = 0, O = 0, O_pad = 0, C = 0, nbins = 4, gwords = 0;
    int rows[4] = {}, cols[4] = {};        // W_l [rows][cols] (W1's columns padded to Fin_pad)
    DevBuf<float> w[8], m[8], v[8], g[8];  // 2l: W_l+1, 2l+1: b_l+1
    DevBuf<uint16_t> wq[4], wqT[4];        // BF16 copies of W_l, and of W_l^T (l >= 2)
    int adam_t = 0;
    DevBuf<int> coef_n, coef_ires, nbr;
    int ntrain = 0;
    DevBuf<uint16_t> tx;  // [max_train][Fin_pad]
    DevBuf<int> tcell, tcnt;
    DevBuf<float> tprob, tshift;
    DevBuf<uint16_t> x, h[3], xT, hT[3], dzq, dzqT, daq, daqT;
    DevBuf<float> a[3], z, dz, dh;
    DevBuf<int> bidx, bcell, bcnt;
    DevBuf<float> bprob, bshift, loss;
    DevBuf<unsigned> gmask;
    DevBuf<int> cand_m, cand_cnt, cand_all;
    DevBuf<float> pmass;
    DevBuf<float> E_audit;
    DevBuf<int> audit_emin;
    DevBuf<double> audit_mass;
    long long audit_pos = 0;
    long long st_pools = 0, st_images = 0, st_all = 0, st_audited = 0, st_fallback = 0, st_uncertain = 0;
    double st_cand = 0., st_min_mass = 1.;
    unsigned long long rng = 0;
};

}  // namespace rome

using namespace rome;

struct rome_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    rome_config cfg{};
    bool configured = false, have_sampling = false, have_model = false;
    std::vector<char> have_ref;
    bool fct_dirty = true;

    // derived sizes
    int is3d = 1, K = 0, D = 0, PSI = 0, T = 0, R = 0, Mtot = 0, Mpad = 0;
    int NORf = 1, NOTf = 1, oriF = 0, cs = 0, fs = 0, F2f = 0;
    int pad = 0, padF = 0, cgrid = 0, r_max = 0;
    size_t grid_n = 0;
    int Pc = 0, Pc_pad = 0, Pf = 0, TW = 1;
    bool two_pass = true, post_enabled = true;

    // host copies of tables
    std::vector<double> pose_c, pose_f, sxc, syc, sxf, syf, tx, ty;
    std::vector<double> sigma2_h, pdf_class_h, pdf_dir_h, prior_x_h, prior_y_h;
    std::vector<int> f_nfine_h, f_xy_h;

    // device persistent
    DevBuf<float2> grids;
    DevBuf<float4> cube;  // 3D: 2x2x2 neighbourhood records of grids, rebuilt when a grid changes
    DevBuf<double> Ai_c, Ai_f, sintab, costab, d_sxc, d_syc, d_sxf, d_syf, d_tx, d_ty, d_pose_c, d_pose_f;
    DevBuf<int> c_nfine, c_ires, c_xy, f_nfine, f_ires, f_xy;
    DevBuf<float2> Sc, Sf;
    DevBuf<double> sigma2, pdf_dir, lpdf_dir, prior_x, prior_y, logsigma2, pdf_m, lpdf_m;
    DevBuf<int> k_m;
    DevBuf<float2> mrow;
    double lpdf_max = 0.;
    DevBuf<int> emin_key;
    // coarse list (rome_set_coarse_mode): entries within reach of the weights, per image
    int coarse_mode = ROME_COARSE_LIST;
    int list_cap = 0;  // 0: all of an image's coarse entries if the pool's lists fit in 1 GiB, else max(65536, a quarter)
    DevBuf<float> lpd_row, reach, list_val;  // lpd_row [Mpad]: ln pdf_direction per rotation (-inf: none)
    DevBuf<int> vbound, list_cnt;
    DevBuf<unsigned> list_ent;
    long long last_list_max = 0, last_list_total = 0;
    int last_coarse_full = 0;  // last batch: 0 list, 1 full pose term, 2 list overflowed -> full
    DevBuf<double> cand_w, cand_du;
    DevBuf<long long> cand_e;
    static constexpr int kCandCap = 8192;
    DevBuf<unsigned char> valid;
    DevBuf<float> FcT;
    DevBuf<float> Qc;  // [Kq][Mpad] |FcT|^2, the A-term operand
    int Kq = 16;
    // mixed-precision arm: scaled BF16/FP16 copies of the coarse table and of the image operand
    DevBuf<int> fct_absmax;
    DevBuf<uint16_t> A16, W16;  // A16: every rotation, or the half table of the paired cross term
    DevBuf<uint16_t> A16lo, W16lo;  // rounding remainders (split three-product recompute of marked samples)
    DevBuf<double> D2c, c_dmin1;    // recomputed diff2 of the marked coarse samples; first pass's minimum
    long long last_marked = 0;      // last pool: marked coarse samples recomputed (mixed arms)
    bool pq = false;            // the rotation table pairs psi with psi + 180 deg (conjugate projections)
    DevBuf<float> inv_scale;
    int K16() const { return (2 * Pc_pad + 63) / 64 * 64; }
    // sums
    DevBuf<float> bp_data, bp_weight;
    DevBuf<double> wsum_sigma2, wsum_pdf_class, wsum_pdf_dir, wsum_sumw_group, wsum_prior_x, wsum_prior_y, wsum_scalars;
    // per batch
    DevBuf<float2> in_fmask, in_fnomask;
    DevBuf<float> in_fctf;
    DevBuf<double> in_hxi2, in_power, in_offx, in_offy, in_norm;
    DevBuf<int> in_group;
    // the batch's inputs as the E-step reads them: the buffers above (host entry) or the caller's own
    // device arrays (device entry, read in place)
    struct {
        const float2 *fmask, *fnomask;
        const float* fctf;
        const double *hxi2, *power, *offx, *offy, *norm;
        const int* group;
    } in{};
    DevBuf<float> WcT, MC2c, At, E, Minvf, Ctff, MC2bp;
    DevBuf<double> C2c, C2f;
    DevBuf<float2> Xf, NMf;
    DevBuf<double> c_dmin, c_sumw, c_maxw, c_thr, c_tail, f_dmin, f_sumw, f_maxw, f_thr, wnorm, meta;
    DevBuf<long long> c_amax, f_amax;
    DevBuf<int> c_nsig, c_status, f_status;
    DevBuf<uint32_t> sigmask;
    DevBuf<int> sig_keys, sig_cnt;
    int sig_cap = kSigCapDefault;  // significant samples per image kept as keys (0: the mask always)
    int sig_dense_left = 0;  // batches that go straight to the significance mask after a key overflow
    long long last_sig_form = 0, last_sig_max = -1;
    DevBuf<int> ntask, nblk;
    DevBuf<long long> task_off, blk_off;
    DevBuf<int> task_img, task_m, task_nt, blk_task, blk_t, ulist, m2u, ucount;
    DevBuf<long long> task_blk0;
    DevBuf<unsigned char> used;
    DevBuf<float2> Pfine;
    DevBuf<double> D2;
    DevBuf<int> argmax;
    // binned backprojection
    static constexpr size_t kBinnedBytes = (size_t)40 << 30;  // slice values + records per batch
    DevBuf<float2> slice_val;
    DevBuf<double> slice_w;
    DevBuf<unsigned> bp_count, bp_off, bp_coff, bp_cursor;
    DevBuf<unsigned long long> bp_rec;
    // hipFFT plans by (type, rank, size), kept for the context's lifetime
    struct FftPlan {
        int type, rank, n, batch;
        hipfftHandle h;
    };
    // image preparation (rome_prepare_images)
    DevBuf<double> ip_img, ip_timg, ip_cbuf, ip_noise, ip_scal, ip_ctf, ip_fctf, ip_power;
    DevBuf<double2> ip_F1, ip_F2, ip_fn, ip_fm;
    std::vector<FftPlan> fft_plans;
    // projector grids built on the device (rome_fourier_map)
    DevBuf<double> fm_vol, fm_Mc, fm_corr, fm_pw, fm_cnt;
    DevBuf<double2> fm_F, fm_grid64;
    // M-step reconstruction buffers (rome_reconstruct)
    DevBuf<double2> rc_data, rc_Fc, rc_Ft;
    DevBuf<double> rc_weight, rc_Fw, rc_Fnw, rc_M, rc_Mout, rc_vol, rc_tab, rc_sums, rc_tau2;
    DevBuf<int> rc_bad;
    std::vector<double> rc_tab_h;
    // stage timing with HIP events on the engine stream (rome_set_profiling)
    static constexpr int kStages = 13;
    static constexpr int kStageRecords = 4;  // a stage can run more than once per batch (a coarse rerun)
    bool profiling = false;
    hipEvent_t ev[kStages][kStageRecords][2] = {};
    int ev_used[kStages] = {};
    double stage_ms[kStages] = {};
    long long stage_calls[kStages] = {};
    long long last_U = 0;
    ProposalState prop;
    // batched read-backs / uploads (pack kernel + pinned staging)
    DevBuf<unsigned char> rb_dev;
    PinnedBuf rb_host, up_host;
    // last batch bookkeeping (parity hooks)
    int last_n = 0;
    long long last_blocks = 0, last_tasks = 0;
    // execution of a pool (rome_set_execution): host-synchronised launches, device-driven launches (the
    // fine pass sized by capacities, its counts read on the device, one synchronisation per pool), or the
    // device-driven pool captured once as a hipGraph and replayed
    int exec_mode = ROME_EXEC_GRAPH;
    int last_exec = 0;              // how the last pool ran (0 host, 1 device, 2 graph replay, +8 rerun)
    int n_cu = 256;
    int persist[4] = {0, 0, 0, 0};  // fixed grids of the device-driven kernels: fine diff2, projection,
                                    // post slices, backprojection (0: from n_cu; rome_autotune sets them)
    DevBuf<int> pool_flags;         // [0] abort (list or key overflow, no significant sample), [1] queue
    unsigned long long state_epoch = 1;  // advanced by every call that changes what a pool's kernels read
    static constexpr int kGraphKeys = 12;
    struct PoolGraph {
        hipGraphExec_t exec = nullptr;
        unsigned long long alloc_epoch = 0, state_epoch = 0;
        int n = 0, prec = 0;
        const void* in[kGraphKeys] = {};  // inputs, given keys and counts, key cap
        size_t rb_bytes = 0;
        long long hits = 0;
        bool failed = false;        // the pool could not be captured: no further attempts for this key
    };
    PoolOut pool_out;  // host side of device-driven pools' read-backs
    // coarse significance of the current pool given by the caller (rome_expectation_fine_device)
    struct {
        const int* keys = nullptr;  // device [n][cap] keys m * T + t
        const int* cnt = nullptr;   // device [n]
        int cap = 0;
    } given;
    int class_lo = 0, class_hi = 1 << 30;  // classes a pool runs (rome_set_class_range)
    std::vector<PoolGraph> graphs;  // most recently used last
    long long graph_launches = 0, graph_captures = 0, dd_reruns = 0;
    double last_launch_ms = 0., last_wait_ms = 0.;  // host time issuing the last device-driven pool / waiting
    static constexpr size_t kPfineBytes = (size_t)32 << 30;  // fine projections of a device-driven pool
    // pool_flags: [0] abort, [1] fine diff2 queue head, [2] invalid given keys, [3] unused, then 64-bit
    // work counters from word 4: 8 per-XCD of the fine projection's tail, 8 of the post slices' tail
    static constexpr int kPoolFlagWords = 4 + 2 * 16;
    unsigned long long* work_counters(int first) const {
        return reinterpret_cast<unsigned long long*>(pool_flags.p + 4) + first;
    }
    // Device-driven launches: the first launch of the fine diff2, the projection and the post slices
    // covers the counts of the pools before (10 % above the largest of them, decaying; the blocks past a
    // pool's own count return at once), a tail kernel or the work queue the rest.  Without history: 4
    // tasks / rotations per image.  Bounded by the capacities.
    double hist_tasks = 0., hist_U = 0.;
    void note_counts(long long tasks, long long u) {
        hist_tasks = std::max((double)tasks, 0.9 * hist_tasks);
        hist_U = std::max((double)u, 0.9 * hist_U);
    }
    long long guess_tasks(int n, long long cap) const {
        const double g = hist_tasks > 0. ? 1.1 * hist_tasks + 64 : 4.0 * n;
        return std::max<long long>(1, std::min<long long>(cap, (long long)g));
    }
    int guess_U(int n, long long cap) const {
        const double g = hist_U > 0. ? 1.1 * hist_U + 64 : 4.0 * n;
        return (int)std::max<long long>(1, std::min<long long>(cap, (long long)g));
    }
    int persist_blocks(int i) const {
        static const int per_cu[4] = {2, 8, 2, 32};
        return persist[i] > 0 ? persist[i] : n_cu * per_cu[i];
    }
    // unique rotations a device-driven pool holds fine projections for (a pool that needs more aborts)
    int dd_ucap(int n) const {
        const long long by_mem = (long long)(kPfineBytes / ((size_t)NORf * Pf * sizeof(float2)));
        return (int)std::max<long long>(1, std::min<long long>({(long long)Mtot, (long long)n * sig_cap, by_mem}));
    }
    // slices of the binned backprojection a device-driven pool can hold
    long long dd_slice_cap(int n) const {
        const long long by_mem = (long long)(kBinnedBytes / ((size_t)Pf * 16));
        const long long by_index = (long long)(((1ULL << 32) - 1) / (unsigned long long)Pf);
        return std::max<long long>(1, std::min<long long>({(long long)n * sig_cap * NORf, by_mem, by_index}));
    }
};

namespace {

constexpr int kTuneCands = 4;  // candidate grids per kernel tried by rome_autotune

int fail(rome_ctx* c, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code;
}

#define HIPCHK(expr)                                                                                   \
    do {                                                                                               \
        hipError_t _e = (expr);                                                                        \
        if (_e != hipSuccess)                                                                          \
            return fail(ctx, ROME_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e),       \
                        __FILE__, __LINE__);                                                           \
    } while (0)

// Every run of a stage within a batch gets its own event pair (up to kStageRecords; a batch whose
// coarse list overflowed runs the coarse stages twice), and the batch's time of the stage is their sum.
void stage_begin(rome_ctx* ctx, int i) {
    if (!ctx->profiling || ctx->ev_used[i] >= rome_ctx::kStageRecords) return;
    (void)hipEventRecord(ctx->ev[i][ctx->ev_used[i]][0], ctx->stream);
}
void stage_end(rome_ctx* ctx, int i) {
    if (!ctx->profiling || ctx->ev_used[i] >= rome_ctx::kStageRecords) return;
    (void)hipEventRecord(ctx->ev[i][ctx->ev_used[i]][1], ctx->stream);
    ++ctx->ev_used[i];
}
void stage_collect(rome_ctx* ctx) {
    if (!ctx->profiling) return;
    for (int i = 0; i < rome_ctx::kStages; ++i) {
        if (!ctx->ev_used[i]) continue;
        double sum = 0.;
        for (int q = 0; q < ctx->ev_used[i]; ++q) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, ctx->ev[i][q][0], ctx->ev[i][q][1]) == hipSuccess) sum += ms;
        }
        ctx->stage_ms[i] += sum;
        ctx->stage_calls[i] += 1;
        ctx->ev_used[i] = 0;
    }
}

inline int pack_xy(int x, int y) { return (int)(((uint32_t)(y & 0xffff) << 16) | (uint32_t)(x & 0xffff)); }

// Mresol of src/map_model.cpp:545-575
int mresol(int size, int i, int j) {
    const int Fsize = size / 2 + 1;
    const int ip = (i < Fsize) ? i : i - size;
    const int ires = (int)round(sqrt((double)(ip * ip + j * j)));
    return (ires < Fsize && !(j == 0 && ip < 0)) ? ires : -1;
}

void rot6(const double A[9], double* out) {
    // Ainv = 2 * A^T restricted to the (x, y) columns (src/reconstruct.cpp:456-466)
    out[0] = 2.0 * A[0];
    out[1] = 2.0 * A[3];
    out[2] = 2.0 * A[1];
    out[3] = 2.0 * A[4];
    out[4] = 2.0 * A[2];
    out[5] = 2.0 * A[5];
}

int upload_poses(rome_ctx* ctx) {
    const int nc = ctx->R, nf = ctx->R * ctx->NORf;
    std::vector<double> a(6 * (size_t)nc), b(6 * (size_t)std::max(nf, 1));
    double A[9];
    for (int r = 0; r < nc; ++r) {
        if (ctx->is3d) euler_matrix(ctx->pose_c[3 * r], ctx->pose_c[3 * r + 1], ctx->pose_c[3 * r + 2], A);
        else euler_matrix(0, 0, ctx->pose_c[r], A);
        rot6(A, &a[6 * r]);
    }
    if (ctx->two_pass) {
        for (int r = 0; r < nf; ++r) {
            if (ctx->is3d) euler_matrix(ctx->pose_f[3 * r], ctx->pose_f[3 * r + 1], ctx->pose_f[3 * r + 2], A);
            else euler_matrix(0, 0, ctx->pose_f[r], A);
            rot6(A, &b[6 * r]);
        }
    } else {
        b = a;  // single pass: the "fine" stage reuses the pass's own rotations
    }
    // Paired cross term: rotation r + PSI/2 of a direction maps pixel (x, y) to minus the coordinates of
    // rotation r (psi + 180 deg), so its projection is the conjugate of r's (Friedel symmetry of the
    // real maps' transforms); used when every pair has that form and the half table tiles evenly.
    ctx->pq = ctx->PSI % 8 == 0 && (ctx->Mtot / 2) % 128 == 0 && ctx->T >= 8;  // T >= 8: <= 18 images per tile
    for (int d = 0; ctx->pq && d < ctx->R / ctx->PSI; ++d)
        for (int p = 0; ctx->pq && p < ctx->PSI / 2; ++p) {
            const double* u = &a[6 * ((size_t)d * ctx->PSI + p)];
            const double* v = &a[6 * ((size_t)d * ctx->PSI + p + ctx->PSI / 2)];
            for (int e = 0; e < 6; ++e)
                if (!(std::fabs(u[e] + v[e]) <= 1e-12)) ctx->pq = false;
        }
    HIPCHK(ctx->Ai_c.upload(a.data(), a.size(), ctx->stream));
    HIPCHK(ctx->Ai_f.upload(b.data(), b.size(), ctx->stream));
    return ROME_OK;
}

int build_phase_tables(rome_ctx* ctx) {
    const rome_config& c = ctx->cfg;
    HIPCHK(ctx->Sc.reserve((size_t)ctx->T * ctx->Pc));
    HIPCHK(launch_phase_table(ctx->stream, ctx->d_sxc.p, ctx->d_syc.p, ctx->T, ctx->c_xy.p, ctx->Pc, c.ori_size,
                              ctx->sintab.p, ctx->costab.p, ctx->Sc.p));
    const int nsf = ctx->T * ctx->NOTf;
    HIPCHK(ctx->Sf.reserve((size_t)nsf * ctx->Pf));
    const double* sx = ctx->two_pass ? ctx->d_sxf.p : ctx->d_sxc.p;
    const double* sy = ctx->two_pass ? ctx->d_syf.p : ctx->d_syc.p;
    HIPCHK(launch_phase_table(ctx->stream, sx, sy, nsf, ctx->f_xy.p, ctx->Pf, c.ori_size, ctx->sintab.p, ctx->costab.p,
                              ctx->Sf.p));
    return ROME_OK;
}

int ensure_ready(rome_ctx* ctx) {
    if (!ctx->configured) return fail(ctx, ROME_ERR_ARG, "rome_configure has not been called");
    if (!ctx->have_sampling) return fail(ctx, ROME_ERR_ARG, "rome_set_sampling has not been called");
    if (!ctx->have_model) return fail(ctx, ROME_ERR_ARG, "rome_set_model has not been called");
    for (int k = 0; k < ctx->K; ++k)
        if (!ctx->have_ref[k]) return fail(ctx, ROME_ERR_ARG, "reference of class %d not set", k);
    if (ctx->fct_dirty) {
        HIPCHK(ctx->FcT.reserve((size_t)2 * ctx->Pc_pad * ctx->Mpad));
        HIPCHK(hipMemsetAsync(ctx->FcT.p, 0, (size_t)2 * ctx->Pc_pad * ctx->Mpad * sizeof(float), ctx->stream));
        const int my_r_max = std::min(ctx->r_max, ctx->cs / 2);
        HIPCHK(launch_project_coarse(ctx->stream, ctx->grids.p, ctx->grid_n, ctx->pad, ctx->padF, ctx->cgrid, ctx->is3d,
                                     ctx->Ai_c.p, ctx->R, ctx->Mtot, ctx->c_xy.p, ctx->Pc, ctx->Pc_pad, my_r_max,
                                     ctx->Mpad, ctx->FcT.p));
        ctx->Kq = (ctx->Pc + 15) / 16 * 16;
        HIPCHK(ctx->Qc.reserve((size_t)ctx->Kq * ctx->Mpad));
        HIPCHK(launch_fct_norm(ctx->stream, ctx->FcT.p, ctx->Pc, ctx->Pc_pad, ctx->Mpad, ctx->Kq, ctx->Qc.p));
        if (ctx->is3d) {
            HIPCHK(ctx->cube.reserve(4 * cube_records(ctx->pad, ctx->padF) * ctx->K));
            HIPCHK(launch_build_cube(ctx->stream, ctx->grids.p, ctx->grid_n, ctx->K, ctx->pad, ctx->padF, ctx->cube.p));
        }
        if (ctx->cfg.precision != ROME_PREC_FP32) {
            const int K16 = ctx->K16();
            const size_t nf = (size_t)2 * ctx->Pc_pad * ctx->Mpad;
            HIPCHK(ctx->fct_absmax.reserve(1));
            HIPCHK(hipMemsetAsync(ctx->fct_absmax.p, 0, sizeof(int), ctx->stream));
            HIPCHK(launch_absmax(ctx->stream, ctx->FcT.p, nf, ctx->fct_absmax.p));
            const int rows = ctx->pq ? ctx->Mtot / 2 : ctx->Mpad;
            HIPCHK(ctx->A16.reserve((size_t)rows * K16));
            HIPCHK(ctx->A16lo.reserve((size_t)rows * K16));
            HIPCHK(launch_fct_to_lowp(ctx->stream, ctx->cfg.precision, ctx->FcT.p, ctx->Pc_pad, ctx->Mpad, ctx->Mtot,
                                      rows, ctx->pq ? ctx->PSI / 2 : 0, ctx->fct_absmax.p, ctx->A16.p, ctx->A16lo.p,
                                      K16));
        }
        ctx->fct_dirty = false;
    }
    return ROME_OK;
}

struct Seg {
    const void* dev;
    void* host;
    size_t bytes;  // a multiple of 4
};

// Device arrays -> host arrays in two halves: readback_begin packs the segments (one pack kernel per
// kPackSegs arrays, all into one device block) and copies the block into pinned host memory, without
// waiting; readback_end synchronises and scatters into the host arrays.  `all`: segments with a null host
// pointer are packed too (a captured pool keeps one layout whatever the caller asks for).
struct RbLayout {
    std::vector<size_t> off;  // per segment, SIZE_MAX when not packed
    size_t bytes = 0;
};

void rb_layout(const std::vector<Seg>& segs, RbLayout& L, bool all) {
    L.off.assign(segs.size(), SIZE_MAX);
    L.bytes = 0;
    for (size_t g = 0; g < segs.size(); ++g) {
        if (!segs[g].bytes || !segs[g].dev || !(all || segs[g].host)) continue;
        L.off[g] = L.bytes;
        L.bytes += (segs[g].bytes + 15) & ~(size_t)15;
    }
}

int readback_begin(rome_ctx* ctx, const std::vector<Seg>& segs, RbLayout& L, bool all) {
    hipStream_t s = ctx->stream;
    rb_layout(segs, L, all);
    if (!L.bytes) return ROME_OK;
    if (g_capturing && (L.bytes > ctx->rb_dev.n || L.bytes > ctx->rb_host.n))
        return fail(ctx, ROME_ERR_HIP, "read-back buffers would grow during capture");
    HIPCHK(ctx->rb_dev.reserve(L.bytes));
    HIPCHK(ctx->rb_host.reserve(L.bytes));
    PackArgs pa{};
    size_t base = 0;
    auto flush = [&](size_t end) -> int {
        if (!pa.nseg) return ROME_OK;
        pa.off[pa.nseg] = end - base;
        pa.dst = ctx->rb_dev.p + base;
        HIPCHK(launch_pack(s, pa));
        pa = PackArgs{};
        base = end;
        return ROME_OK;
    };
    for (size_t g = 0; g < segs.size(); ++g) {
        if (L.off[g] == SIZE_MAX) continue;
        if (pa.nseg == kPackSegs) {
            const int r = flush(L.off[g]);
            if (r) return r;
        }
        pa.src[pa.nseg] = segs[g].dev;
        pa.off[pa.nseg] = L.off[g] - base;
        pa.len[pa.nseg] = segs[g].bytes;
        ++pa.nseg;
    }
    const int r = flush(L.bytes);
    if (r) return r;
    HIPCHK(hipMemcpyAsync(ctx->rb_host.p, ctx->rb_dev.p, L.bytes, hipMemcpyDeviceToHost, s));
    return ROME_OK;
}

int readback_end(rome_ctx* ctx, const std::vector<Seg>& segs, const RbLayout& L) {
    if (L.bytes) HIPCHK(hipStreamSynchronize(ctx->stream));
    for (size_t g = 0; g < segs.size(); ++g)
        if (L.off[g] != SIZE_MAX && segs[g].host)
            std::memcpy(segs[g].host, static_cast<const char*>(ctx->rb_host.p) + L.off[g], segs[g].bytes);
    return ROME_OK;
}

// returns once the data is on the host; segments with a null host pointer are skipped
int readback(rome_ctx* ctx, const std::vector<Seg>& segs) {
    RbLayout L;
    const int r = readback_begin(ctx, segs, L, false);
    return r ? r : readback_end(ctx, segs, L);
}

WeightParams weight_params(rome_ctx* ctx) {
    WeightParams wp;
    wp.K = ctx->K;
    wp.D = ctx->D;
    wp.PSI = ctx->PSI;
    wp.R = ctx->R;
    wp.Mtot = ctx->Mtot;
    wp.T = ctx->T;
    wp.magicM = ((1ULL << 48) + (unsigned long long)ctx->Mtot - 1) / (unsigned long long)ctx->Mtot;
    wp.pdf_m = ctx->pdf_m.p;
    wp.lpdf_m = ctx->lpdf_m.p;
    wp.k_m = ctx->k_m.p;
    wp.mrow = ctx->mrow.p;
    wp.lpdf_max = ctx->lpdf_max;
    wp.trans_x = ctx->d_tx.p;
    wp.trans_y = ctx->d_ty.p;
    wp.old_offx = ctx->in.offx;
    wp.old_offy = ctx->in.offy;
    wp.prior_x = ctx->prior_x.p;
    wp.prior_y = ctx->prior_y.p;
    wp.sigma2_offset = ctx->cfg.sigma2_offset;
    wp.adaptive_fraction = ctx->cfg.adaptive_fraction;
    // a single pass marks its significant samples with a margin: the superset gets exact diff2
    wp.mark_margin = ctx->two_pass ? 0.0 : kRecomputeMargin;
    return wp;
}

// ------------------------------------------------------------------ learned pose proposal (host side)
unsigned long long splitmix(unsigned long long& st) {
    unsigned long long z = (st += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
double uniform01(unsigned long long& st) { return ((splitmix(st) >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
double gauss(unsigned long long& st) {
    const double u = uniform01(st), v = uniform01(st);
    return std::sqrt(-2.0 * std::log(u)) * std::cos(2.0 * 3.14159265358979323846 * v);
}

// Activation buffers for `rows` images.
int prop_acts(rome_ctx* ctx, int rows) {
    ProposalState& P = ctx->prop;
    const size_t r = (size_t)std::max(rows, 1), r32 = (r + 31) / 32 * 32;
    HIPCHK(P.x.grow(r * P.Fin_pad));
    HIPCHK(P.xT.grow((size_t)P.Fin_pad * r32));
    for (int l = 0; l < 3; ++l) {
        HIPCHK(P.a[l].grow(r * P.H));
        HIPCHK(P.h[l].grow(r * P.H));
        HIPCHK(P.hT[l].grow((size_t)P.H * r32));
    }
    HIPCHK(P.z.grow(r * P.O_pad));
    HIPCHK(P.dz.grow(r * P.O_pad));
    HIPCHK(P.dzq.grow(r * P.O_pad));
    HIPCHK(P.dzqT.grow((size_t)P.O_pad * r32));
    HIPCHK(P.dh.grow(r * P.H));
    HIPCHK(P.daq.grow(r * P.H));
    HIPCHK(P.daqT.grow((size_t)P.H * r32));
    const size_t tc = (size_t)P.cfg.target_cells;
    HIPCHK(P.bcell.grow(r * tc));
    HIPCHK(P.bprob.grow(r * tc));
    HIPCHK(P.bcnt.grow(r));
    HIPCHK(P.bshift.grow(2 * r));
    return ROME_OK;
}

// BF16 copies of the weight matrices and the transposes the backward products read.
int prop_refresh(rome_ctx* ctx) {
    ProposalState& P = ctx->prop;
    hipStream_t s = ctx->stream;
    for (int l = 0; l < 4; ++l) {
        HIPCHK(mlp_to_bf16(s, P.w[2 * l].p, P.wq[l].p, (size_t)P.rows[l] * P.cols[l]));
        if (l > 0) HIPCHK(mlp_transpose(s, P.wq[l].p, false, P.rows[l], P.cols[l], P.cols[l], P.wqT[l].p, l == 3 ? P.O_pad : P.rows[l]));
    }
    return ROME_OK;
}

// Forward pass of `rows` images whose features are in P.x: pre-activations a, activations h, outputs z.
int prop_forward(rome_ctx* ctx, int rows) {
    ProposalState& P = ctx->prop;
    hipStream_t s = ctx->stream;
    const uint16_t* in = P.x.p;
    int ldin = P.Fin_pad, kin = P.Fin_pad;
    for (int l = 0; l < 3; ++l) {
        HIPCHK(mlp_gemm(s, in, ldin, P.wq[l].p, P.cols[l], P.a[l].p, P.H, P.w[2 * l + 1].p, rows, P.H, kin));
        HIPCHK(mlp_gelu_fwd(s, P.a[l].p, P.H, P.h[l].p, P.H, rows, P.H));
        in = P.h[l].p;
        ldin = kin = P.H;
    }
    HIPCHK(mlp_gemm(s, in, P.H, P.wq[3].p, P.H, P.z.p, P.O_pad, P.w[7].p, rows, P.O, P.H));
    return ROME_OK;
}

// Loss and gradients of a batch of B images (features in P.x, forward done, targets in P.b*); the loss
// is added to P.loss.
int prop_backward(rome_ctx* ctx, int B) {
    ProposalState& P = ctx->prop;
    hipStream_t s = ctx->stream;
    HIPCHK(mlp_softmax_ce(s, P.z.p, P.O_pad, P.ncell, B, P.bcell.p, P.bprob.p, P.bcnt.p, P.cfg.target_cells,
                          P.bshift.p, (float)P.cfg.shift_weight, P.dz.p, P.dzq.p, P.O_pad, P.loss.p));
    // layer 4: dW4 = dz^T h3, db4 = sum dz, dh3 = dz W4
    HIPCHK(mlp_transpose(s, P.dzq.p, false, B, P.O_pad, P.O_pad, P.dzqT.p, B));
    HIPCHK(mlp_transpose(s, P.h[2].p, false, B, P.H, P.H, P.hT[2].p, B));
    HIPCHK(mlp_gemm(s, P.dzqT.p, B, P.hT[2].p, B, P.g[6].p, P.H, nullptr, P.O, P.H, B));
    HIPCHK(mlp_colsum(s, P.dz.p, B, P.O, P.O_pad, P.g[7].p));
    HIPCHK(mlp_gemm(s, P.dzq.p, P.O_pad, P.wqT[3].p, P.O_pad, P.dh.p, P.H, nullptr, B, P.H, P.O_pad));
    for (int l = 2; l >= 0; --l) {
        // da = dh * gelu'(a); dW = da^T input; db = sum da; dh(prev) = da W
        HIPCHK(mlp_gelu_bwd(s, P.dh.p, P.a[l].p, P.daq.p, (size_t)B * P.H));
        HIPCHK(mlp_transpose(s, P.daq.p, false, B, P.H, P.H, P.daqT.p, B));
        const uint16_t* inT;
        int ncol;
        if (l == 0) {
            HIPCHK(mlp_transpose(s, P.x.p, false, B, P.Fin_pad, P.Fin_pad, P.xT.p, B));
            inT = P.xT.p;
            ncol = P.Fin_pad;
        } else {
            HIPCHK(mlp_transpose(s, P.h[l - 1].p, false, B, P.H, P.H, P.hT[l - 1].p, B));
            inT = P.hT[l - 1].p;
            ncol = P.H;
        }
        HIPCHK(mlp_gemm(s, P.daqT.p, B, inT, B, P.g[2 * l].p, ncol, nullptr, P.H, ncol, B));
        HIPCHK(mlp_colsum(s, P.dh.p, B, P.H, P.H, P.g[2 * l + 1].p));
        if (l > 0) HIPCHK(mlp_gemm(s, P.daq.p, P.H, P.wqT[l].p, P.H, P.dh.p, P.H, nullptr, B, P.H, P.H));
    }
    return ROME_OK;
}

int prop_adam(rome_ctx* ctx, double lr) {
    ProposalState& P = ctx->prop;
    hipStream_t s = ctx->stream;
    ++P.adam_t;
    const double b1 = 0.9, b2 = 0.999;
    const float bc1 = (float)(1.0 / (1.0 - std::pow(b1, P.adam_t))), bc2 = (float)(1.0 / (1.0 - std::pow(b2, P.adam_t)));
    for (int q = 0; q < 8; ++q) {
        const int l = q / 2;
        const size_t n = (q & 1) ? (size_t)P.rows[l] : (size_t)P.rows[l] * P.cols[l];
        HIPCHK(mlp_adam(s, P.w[q].p, P.m[q].p, P.v[q].p, P.g[q].p, (q & 1) ? nullptr : P.wq[l].p, n, (float)lr,
                        (float)b1, (float)b2, 1e-8f, bc1, bc2));
    }
    for (int l = 1; l < 4; ++l)
        HIPCHK(mlp_transpose(s, P.wq[l].p, false, P.rows[l], P.cols[l], P.cols[l], P.wqT[l].p, l == 3 ? P.O_pad : P.rows[l]));
    return ROME_OK;
}

// The last pool's images (fine pass done) -> training rows: features and posterior targets.
int prop_collect(rome_ctx* ctx, int n) {
    ProposalState& P = ctx->prop;
    const int rows = std::min(n, P.cfg.max_train - P.ntrain);
    if (rows <= 0) return ROME_OK;
    hipStream_t s = ctx->stream;
    HIPCHK(mlp_features(s, rows, ctx->in.fmask, ctx->F2f, P.coef_n.p, P.coef_ires.p, P.ncoef, ctx->sigma2.p,
                        ctx->in.group, ctx->oriF, P.tx.p + (size_t)P.ntrain * P.Fin_pad, P.Fin_pad));
    ProposalTargetArgs ta;
    ta.K = ctx->K;
    ta.C = P.C;
    ta.nbins = P.nbins;
    ta.R = ctx->R;
    ta.PSI = ctx->PSI;
    ta.NOR = ctx->NORf;
    ta.NOT = ctx->NOTf;
    ta.tcap = P.cfg.target_cells;
    ta.img_blk = ctx->blk_off.p;
    ta.blk_task = ctx->blk_task.p;
    ta.blk_t = ctx->blk_t.p;
    ta.task_m = ctx->task_m.p;
    ta.D2 = ctx->D2.p;
    ta.dmin = ctx->f_dmin.p;
    ta.sumw = ctx->f_sumw.p;
    ta.pdf_m = ctx->pdf_m.p;
    ta.trans_x = ctx->d_tx.p;
    ta.trans_y = ctx->d_ty.p;
    ta.old_offx = ctx->in.offx;
    ta.old_offy = ctx->in.offy;
    ta.prior_x = ctx->prior_x.p;
    ta.prior_y = ctx->prior_y.p;
    ta.sigma2_offset = ctx->cfg.sigma2_offset;
    ta.shift_x = ctx->d_sxf.p;
    ta.shift_y = ctx->d_syf.p;
    double rng = 1.0;
    for (size_t t = 0; t < ctx->tx.size(); ++t) rng = std::max(rng, std::hypot(ctx->tx[t], ctx->ty[t]));
    ta.shift_scale = (float)rng;
    ta.row0 = P.ntrain;
    ta.tcell = P.tcell.p;
    ta.tprob = P.tprob.p;
    ta.tcnt = P.tcnt.p;
    ta.tshift = P.tshift.p;
    HIPCHK(proposal_targets(s, rows, ta));
    P.ntrain += rows;
    return ROME_OK;
}

// The E-step on the device inputs ctx->in points to.
// Pool modes of run_pool
enum { kPoolHost = 0, kPoolDevice = 1, kPoolCapture = 2 };

// The segments a device-driven pool reads back (the same list, in the same order, for a given pool
// size: a captured pool's read-back layout).
std::vector<Seg> pool_segs(rome_ctx* ctx, int n, PoolOut& o) {
    const size_t nd = n * sizeof(double), ni = n * sizeof(int);
    const bool tp = ctx->two_pass;
    return {{ctx->f_sumw.p, o.sumw.data(), nd},
            {ctx->meta.p, o.meta.data(), 9 * nd},
            {ctx->argmax.p, o.argmax.data(), 6 * ni},
            {ctx->c_nsig.p, o.nsig.data(), ni},
            {ctx->f_dmin.p, o.dmin.data(), nd},
            {ctx->f_thr.p, o.thr.data(), nd},
            {tp ? ctx->c_dmin.p : ctx->f_dmin.p, o.dmin_c.data(), nd},
            {tp ? ctx->c_sumw.p : ctx->f_sumw.p, o.sumw_c.data(), nd},
            {tp ? ctx->c_thr.p : ctx->f_thr.p, o.thr_c.data(), nd},
            {ctx->wnorm.p, o.wnorm.data(), nd},
            {ctx->c_status.p, o.cstatus.data(), ni},
            {ctx->list_cnt.p, o.lcount.data(), ni},
            {ctx->sig_cnt.p, o.scount.data(), ni},
            {ctx->task_off.p + n, &o.tot[0], sizeof(long long)},
            {ctx->blk_off.p + n, &o.tot[1], sizeof(long long)},
            {ctx->ucount.p, &o.U, sizeof(int)},
            {ctx->pool_flags.p, o.flags, 2 * sizeof(int)}};
}

int run_pool(rome_ctx* ctx, int n, rome_results* res, int pool_mode, PoolOut* dout, RbLayout* dlayout) {
    const bool dd = pool_mode != kPoolHost;  // device-driven: no host synchronisation until the read-back
    // buffers whose size follows the pool: 25 % headroom on the host path, exact capacities device-driven
    auto sized = [dd](auto& buf, size_t count) { return dd ? buf.reserve_exact(count) : buf.grow(count); };
    int rc = ensure_ready(ctx);
    if (rc) return rc;
    const rome_config& c = ctx->cfg;
    hipStream_t s = ctx->stream;
    if (dd || ctx->given.keys) {
        HIPCHK(ctx->pool_flags.reserve(rome_ctx::kPoolFlagWords));
        // abort flag, fine diff2 queue head, invalid keys; the work counters of the device-driven kernels' tails
        HIPCHK(launch_fill32(s, ctx->pool_flags.p, 0u, rome_ctx::kPoolFlagWords));
    }
    if (dd) dout->size(n);
    const int T = ctx->T, M = ctx->Mtot;
    const int Ncols = n * T;
    const int Npad = (Ncols + 255) / 256 * 256;
    const int Kdim = 2 * ctx->Pc_pad;
    const int prec = c.precision;
    const int K16 = ctx->K16();
    // ---- image preparation
    if (prec == ROME_PREC_FP32) {
        HIPCHK(ctx->WcT.reserve((size_t)Kdim * Npad));
        HIPCHK(launch_fill32(s, ctx->WcT.p, 0u, (size_t)Kdim * Npad));
    } else {
        // 256 rows of slack: the audit slice of a learned-proposal pool reads whole 256-row tiles from its start
        HIPCHK(ctx->W16.reserve((size_t)(Npad + 256) * K16));
        HIPCHK(ctx->W16lo.reserve((size_t)(Npad + 256) * K16));
        HIPCHK(ctx->inv_scale.reserve(n));
        // rows past the last column only feed discarded outputs; keep them finite
        HIPCHK(launch_fill32(s, ctx->W16.p + (size_t)Ncols * K16, 0u, (size_t)(Npad + 256 - Ncols) * K16 / 2));
    }
    const int ldn = (n + 127) / 128 * 128;  // images of the A-term operand, padded with zero columns
    HIPCHK(ctx->MC2c.reserve((size_t)ctx->Kq * ldn));
    HIPCHK(launch_fill32(s, ctx->MC2c.p, 0u, (size_t)ctx->Kq * ldn));
    HIPCHK(ctx->C2c.reserve(n));
    HIPCHK(ctx->C2f.reserve(n));
    const size_t nPf = (size_t)n * ctx->Pf;
    HIPCHK(ctx->Xf.reserve(nPf));
    HIPCHK(ctx->Minvf.reserve(nPf));
    HIPCHK(ctx->Ctff.reserve(nPf));
    HIPCHK(ctx->NMf.reserve(nPf));
    HIPCHK(ctx->MC2bp.reserve(nPf));
    PrepArgs pa;
    pa.fmask = ctx->in.fmask;
    pa.fnomask = ctx->in.fnomask;
    pa.fctf = ctx->in.fctf;
    pa.group = ctx->in.group;
    pa.hxi2 = ctx->in.hxi2;
    pa.sigma2 = ctx->sigma2.p;
    pa.ori_Fsize = ctx->oriF;
    pa.F2f = ctx->F2f;
    pa.sigma2_fudge = c.sigma2_fudge;
    pa.ctf_in_diff = (c.do_ctf_correction && c.refs_are_ctf_corrected) ? 1 : 0;
    pa.do_ctf = c.do_ctf_correction ? 1 : 0;
    pa.do_map = c.do_map ? 1 : 0;
    pa.have_fine = 1;
    pa.Pc = ctx->Pc;
    pa.Pc_pad = ctx->Pc_pad;
    pa.T = T;
    pa.Npad = Npad;
    pa.c_nfine = ctx->c_nfine.p;
    pa.c_ires = ctx->c_ires.p;
    pa.Sc = ctx->Sc.p;
    pa.Pf = ctx->Pf;
    pa.f_nfine = ctx->f_nfine.p;
    pa.f_ires = ctx->f_ires.p;
    pa.WcT = prec == ROME_PREC_FP32 ? ctx->WcT.p : nullptr;
    pa.MC2c = ctx->MC2c.p;
    pa.mc2_ld = ldn;
    pa.C2c = ctx->C2c.p;
    pa.Xf = ctx->Xf.p;
    pa.Minvf = ctx->Minvf.p;
    pa.Ctff = ctx->Ctff.p;
    pa.C2f = ctx->C2f.p;
    pa.NMf = ctx->NMf.p;
    pa.MC2bp = ctx->MC2bp.p;
    stage_begin(ctx, 0);
    HIPCHK(launch_prep_images(s, n, pa));
    if (prec != ROME_PREC_FP32) {
        PrepLowpArgs pl;
        pl.fmask = ctx->in.fmask;
        pl.fctf = ctx->in.fctf;
        pl.group = ctx->in.group;
        pl.sigma2 = ctx->sigma2.p;
        pl.ori_Fsize = ctx->oriF;
        pl.F2f = ctx->F2f;
        pl.sigma2_fudge = c.sigma2_fudge;
        pl.ctf_in_diff = pa.ctf_in_diff;
        pl.Pc = ctx->Pc;
        pl.Pc_pad = ctx->Pc_pad;
        pl.T = T;
        pl.K16 = K16;
        pl.c_nfine = ctx->c_nfine.p;
        pl.c_ires = ctx->c_ires.p;
        pl.Sc = ctx->Sc.p;
        pl.fct_absmax = ctx->fct_absmax.p;
        pl.W16 = ctx->W16.p;
        pl.W16lo = ctx->W16lo.p;
        pl.inv_scale = ctx->inv_scale.p;
        HIPCHK(launch_prep_lowp(s, prec, n, pl));
    }
    stage_end(ctx, 0);
    // ---- coarse cross term + norms (a pool with its coarse significance given has no coarse pass)
    const bool given = ctx->given.keys != nullptr;
    if (given && !ctx->two_pass) return fail(ctx, ROME_ERR_ARG, "given coarse significance needs oversampling 1");
    HIPCHK(ctx->At.reserve(given ? 1 : (size_t)n * ctx->Mpad));
    WeightParams wp = weight_params(ctx);
    stage_begin(ctx, 1);
    if (!given) HIPCHK(launch_coarse_aterm(s, n, ctx->Qc.p, ctx->Kq, M, ctx->Mpad, ctx->MC2c.p, ldn, ctx->At.p));
    if (!given && ctx->coarse_mode == ROME_COARSE_LIST && ctx->two_pass) {
        HIPCHK(ctx->reach.reserve(n));
        HIPCHK(launch_list_reach(s, n, wp, ctx->reach.p));
    }
    stage_end(ctx, 1);
    for (auto* b : {&ctx->c_dmin, &ctx->c_sumw, &ctx->c_maxw, &ctx->c_thr, &ctx->c_tail, &ctx->f_dmin, &ctx->f_sumw,
                    &ctx->f_maxw, &ctx->f_thr, &ctx->wnorm})
        HIPCHK(b->reserve(n));
    HIPCHK(ctx->meta.reserve((size_t)9 * n));
    HIPCHK(ctx->c_amax.reserve(n));
    HIPCHK(ctx->f_amax.reserve(n));
    HIPCHK(ctx->c_nsig.reserve(n));
    HIPCHK(ctx->c_status.reserve(n));
    HIPCHK(ctx->f_status.reserve(n));
    HIPCHK(ctx->used.reserve(((size_t)M + 3) / 4 * 4));
    HIPCHK(ctx->ulist.reserve(M));
    HIPCHK(ctx->m2u.reserve(M));
    HIPCHK(ctx->ucount.reserve(1));
    HIPCHK(ctx->emin_key.reserve(n));
    HIPCHK(ctx->cand_w.reserve((size_t)n * rome_ctx::kCandCap));
    HIPCHK(ctx->cand_du.reserve((size_t)n * rome_ctx::kCandCap));
    HIPCHK(ctx->cand_e.reserve((size_t)n * rome_ctx::kCandCap));
    HIPCHK(ctx->ntask.reserve(n));
    HIPCHK(ctx->nblk.reserve(n));
    HIPCHK(ctx->task_off.reserve(n + 1));
    HIPCHK(ctx->blk_off.reserve(n + 1));
    // The coarse pass marks a superset of its significant samples (mark_margin > 0) whenever their diff2 is
    // recomputed before the fine pass: a single pass (exact recompute, the pass's weights and sums come
    // from it) or a mixed-precision arm (split three-product recompute, significance from it).
    // A pool on the learned proposal computes its candidates' diff2 as split three-product sums already;
    // it needs the recompute only when it falls back to the exhaustive low-precision pass.
    ProposalState& P = ctx->prop;
    const bool learned = !given && P.ready && P.mode == ROME_PROPOSAL_USE && prec != ROME_PREC_FP32 && ctx->two_pass &&
                         ctx->is3d && ctx->coarse_mode == ROME_COARSE_LIST && P.K == ctx->K && P.D == ctx->D &&
                         P.PSI == ctx->PSI && P.cur == c.current_size;
    bool lowp_recompute = ctx->two_pass && prec != ROME_PREC_FP32 && !learned && !given;
    if (lowp_recompute) wp.mark_margin = prec == ROME_PREC_BF16 ? kLowpRecomputeMarginBF16 : kLowpRecomputeMargin;
    WeightParams wp_exact = wp;
    wp_exact.mark_margin = 0.0;
    WeightOutArgs wc{ctx->c_dmin.p, ctx->c_sumw.p,  ctx->c_maxw.p,
                     ctx->c_amax.p, ctx->c_thr.p,   ctx->c_nsig.p,
                     ctx->c_status.p, wp.mark_margin > 0. ? ctx->c_tail.p : nullptr};
    const WeightOutArgs wc_exact{ctx->c_dmin.p, ctx->c_sumw.p, ctx->c_maxw.p, ctx->c_amax.p, ctx->c_thr.p,
                                 ctx->c_nsig.p, ctx->c_status.p, nullptr};
    const CandArgs cand{ctx->cand_w.p, ctx->cand_du.p, ctx->cand_e.p, rome_ctx::kCandCap};
    long long tot[2];
    int U = 0;
    std::vector<int> cstatus(n), lcount, scount(n);
    bool sig_dense = false;
    std::vector<Seg> extra_rb;  // further device arrays read back with the significance results
    // weights and significance -> fine-pass task counts and the rotations they need; returns after
    // one host synchronisation.  Sparse: per-image keys of the significant samples; dense: the mask
    // (when some image has more than kSigCap significant samples).
    // from_d2c: the second significance pass of a mixed arm, over the recomputed marked samples (D2c, one
    // block per sample in the first pass's task lists) plus the mass the first pass left unmarked.
    auto significance = [&](const ListArgs* L, bool dense, bool from_d2c = false) -> int {
        SigArgs sg{nullptr, ctx->TW, nullptr, nullptr, 0};
        if (dense) {
            HIPCHK(ctx->sigmask.reserve((size_t)n * M * ctx->TW));
            sg.mask = ctx->sigmask.p;
        } else {
            HIPCHK(ctx->sig_keys.reserve((size_t)n * ctx->sig_cap));
            HIPCHK(ctx->sig_cnt.reserve(n));
            sg.keys = ctx->sig_keys.p;
            sg.cnt = ctx->sig_cnt.p;
            sg.cap = ctx->sig_cap;
        }
        HIPCHK(launch_fill32(s, ctx->used.p, 0u, ((size_t)M + 3) / 4));
        stage_begin(ctx, 3);
        if (from_d2c)
            HIPCHK(launch_weights_fine(s, n, wp_exact, ctx->D2c.p, ctx->blk_task.p, ctx->blk_t.p, ctx->task_m.p,
                                       ctx->blk_off.p, 1, wc_exact, cand, TailArgs{ctx->c_tail.p, ctx->c_dmin1.p}, sg));
        else if (L)
            HIPCHK(launch_weights_list(s, n, wp, *L, ctx->C2c.p, ctx->emin_key.p, wc, sg, cand));
        else
            HIPCHK(launch_weights_coarse(s, n, wp, ctx->E.p, ctx->C2c.p, ctx->emin_key.p, wc, sg, cand));
        stage_end(ctx, 3);
        stage_begin(ctx, 4);
        HIPCHK(launch_task_build(s, n, sg, M, T, ctx->ntask.p, ctx->nblk.p, ctx->task_off.p, ctx->blk_off.p,
                                 ctx->used.p, ctx->ulist.p, ctx->m2u.p, ctx->ucount.p));
        stage_end(ctx, 4);
        if (L) lcount.resize(n);
        sig_dense = dense;
        if (dd) {  // no read-back: a list or key overflow, or an image without significant samples, aborts
            HIPCHK(launch_pool_check(s, n, ctx->c_status.p, sg.cnt, sg.cap, ctx->ucount.p, ctx->dd_ucap(n),
                                     ctx->pool_flags.p));
            return ROME_OK;
        }
        std::vector<Seg> segs{{ctx->task_off.p + n, &tot[0], sizeof(long long)},
                              {ctx->blk_off.p + n, &tot[1], sizeof(long long)},
                              {ctx->ucount.p, &U, sizeof(int)},
                              {ctx->c_status.p, cstatus.data(), n * sizeof(int)},
                              {ctx->list_cnt.p, L ? lcount.data() : nullptr, n * sizeof(int)},
                              {ctx->sig_cnt.p, dense ? nullptr : scount.data(), n * sizeof(int)}};
        segs.insert(segs.end(), extra_rb.begin(), extra_rb.end());
        extra_rb.clear();
        return readback(ctx, segs);
    };
    // cross term -> list of the entries within reach of the weights (or the whole pose term E) ->
    // weights and significance -> fine-pass tasks; returns after one host synchronisation
    // the coarse list's buffers: every entry when the pool's lists fit in 1 GiB, else a quarter of them (a
    // list past its cap sends the pool through the full pose term)
    auto list_setup = [&](ListArgs& L) -> int {
        const size_t per = (size_t)M * T;
        const int cap = ctx->list_cap > 0                       ? ctx->list_cap
                        : (size_t)n * per * 8 <= ((size_t)1 << 30) ? (int)per
                                                                  : (int)std::max<size_t>(65536, (per + 3) / 4);
        HIPCHK(ctx->vbound.reserve(n));
        HIPCHK(ctx->list_cnt.reserve(n));
        HIPCHK(ctx->list_ent.reserve((size_t)n * cap));
        HIPCHK(ctx->list_val.reserve((size_t)n * cap));
        HIPCHK(launch_fill32(s, ctx->vbound.p, 0x7fffffffu, (size_t)n));
        HIPCHK(launch_fill32(s, ctx->list_cnt.p, 0u, (size_t)n));
        L = ListArgs{ctx->lpd_row.p, ctx->reach.p, ctx->vbound.p, ctx->list_ent.p, ctx->list_val.p, ctx->list_cnt.p, cap};
        return ROME_OK;
    };
    // weights and significance of the coarse pass (keys, or the mask past their cap)
    auto finish = [&](const ListArgs* L) -> int {
        if (dd) return significance(L, false);  // keys only (the pool was chosen for it)
        const bool dense = ctx->sig_cap == 0 || ctx->sig_dense_left > 0 || (long long)M * T > 0x7fffffffLL;
        int r = significance(L, dense);
        ctx->last_sig_form = dense ? 1 : 0;
        ctx->last_sig_max = dense ? -1 : *std::max_element(scount.begin(), scount.end());
        if (r || dense) return r;
        if (ctx->last_sig_max > ctx->sig_cap) {
            ctx->sig_dense_left = 8;  // more significant samples than the keys hold: the mask, for a while
            ctx->last_sig_form = 2;
            return significance(L, true);
        }
        return 0;
    };
    auto coarse = [&](bool list) -> int {
        HIPCHK(launch_fill32(s, ctx->emin_key.p, 0x7fffffffu, (size_t)n));
        ListArgs L{};
        if (list) {
            const int lr = list_setup(L);
            if (lr) return lr;
        } else {
            HIPCHK(ctx->E.reserve((size_t)n * M * T));
        }
        float* Ep = list ? nullptr : ctx->E.p;
        stage_begin(ctx, 2);
        if (prec == ROME_PREC_FP32) {
            // an L2-resident slice of the rotation operand per XCD: about 2.5 MB of its 4 MB L2
            const int nbm = ctx->Mpad / 128;
            GemmArgs ga{ctx->FcT.p, ctx->WcT.p, ctx->At.p, Ep, ctx->valid.p, ctx->emin_key.p, Kdim, ctx->Mpad, Npad, M,
                        Ncols, T, gemm_sub_band(nbm, (size_t)128 * Kdim * 4, (size_t)5 << 19), L};
            HIPCHK(launch_coarse_gemm(s, ga));
        } else {
            LowpGemmArgs la{ctx->A16.p, ctx->W16.p, ctx->At.p, ctx->inv_scale.p, Ep, ctx->valid.p,
                            ctx->emin_key.p, K16, ctx->Mpad, Npad, M, Ncols, T,
                            gemm_sub_band(ctx->Mpad / 256, (size_t)256 * K16 * 2, (size_t)5 << 19), L, 0, 0};
            if (ctx->pq) {  // the paired cross term over the half table (list or E output)
                la.Mh = M / 2;
                la.psih = ctx->PSI / 2;
                la.sband = gemm_sub_band(la.Mh / 128, (size_t)128 * K16 * 2, (size_t)5 << 19);
                if (!list) la.list.lpd = ctx->lpd_row.p;
            }
            HIPCHK(launch_gemm_lowp(s, prec, la));
        }
        stage_end(ctx, 2);
        return finish(list ? &L : nullptr);
    };
    // Learned proposal: network -> candidate rotations per image -> their diff2 (split three-product) into
    // the coarse list; the audit slice also runs the exhaustive pass (full pose term) for its captured mass.
    int na = 0;
    std::vector<double> audit_h;
    std::vector<int> cand_h, all_h;
    bool uncertain = false;  // the learned pool's images mostly take every rotation: exhaustive instead
    auto coarse_learned = [&]() -> int {
        HIPCHK(P.gmask.reserve((size_t)n * P.gwords));
        HIPCHK(P.cand_m.reserve((size_t)n * M));
        HIPCHK(P.cand_cnt.reserve(n));
        HIPCHK(P.cand_all.reserve(n));
        HIPCHK(P.pmass.reserve(n));
        int pr = prop_acts(ctx, n);
        if (pr) return pr;
        stage_begin(ctx, 12);
        HIPCHK(mlp_features(s, n, ctx->in.fmask, ctx->F2f, P.coef_n.p, P.coef_ires.p, P.ncoef, ctx->sigma2.p,
                            ctx->in.group, ctx->oriF, P.x.p, P.Fin_pad));
        pr = prop_forward(ctx, n);
        if (pr) return pr;
        CandidateArgs ca{P.z.p,     P.O_pad,   ctx->K,       P.C,          ctx->D,     P.nbins,
                         ctx->PSI,  P.cfg.top_k, (float)P.cfg.min_mass, P.nbr.p, P.gmask.p, P.gwords,
                         P.cand_m.p, M,         P.cand_cnt.p, P.pmass.p,   P.cand_all.p};
        HIPCHK(select_candidates(s, n, ca));
        {
            // a pool where most images take every rotation gains nothing from the per-image candidate pass:
            // it runs the exhaustive coarse pass instead
            std::vector<int> all(n);
            pr = readback(ctx, {{P.cand_all.p, all.data(), n * sizeof(int)}});
            if (pr) return pr;
            const long long nall = std::accumulate(all.begin(), all.end(), 0LL);
            if (nall * 2 > n) {
                uncertain = true;
                stage_end(ctx, 12);
                return ROME_OK;
            }
        }
        na = P.cfg.audit_fraction > 0. ? std::min(n, (int)std::ceil(P.cfg.audit_fraction * n)) : 0;
        const int a0 = na ? (int)(P.audit_pos % (long long)(n - na + 1)) : 0;
        P.audit_pos += na;
        if (na) {
            HIPCHK(P.E_audit.reserve((size_t)na * M * T));
            HIPCHK(P.audit_emin.reserve(na));
            HIPCHK(P.audit_mass.reserve(na));
            HIPCHK(launch_fill32(s, P.audit_emin.p, 0x7fffffffu, (size_t)na));
            const int ncol = na * T, npad = (ncol + 255) / 256 * 256;
            LowpGemmArgs la{ctx->A16.p, ctx->W16.p + (size_t)a0 * T * K16, ctx->At.p + (size_t)a0 * ctx->Mpad,
                            ctx->inv_scale.p + a0, P.E_audit.p, ctx->valid.p, P.audit_emin.p, K16, ctx->Mpad, npad,
                            M, ncol, T, gemm_sub_band(ctx->Mpad / 256, (size_t)256 * K16 * 2, (size_t)5 << 19),
                            ListArgs{}, 0, 0};
            if (ctx->pq) {
                la.Mh = M / 2;
                la.psih = ctx->PSI / 2;
                la.sband = gemm_sub_band(la.Mh / 128, (size_t)128 * K16 * 2, (size_t)5 << 19);
                la.list.lpd = ctx->lpd_row.p;
            }
            HIPCHK(launch_gemm_lowp(s, prec, la));
            AuditArgs aa{P.E_audit.p, a0, T, M, P.nbins, P.gwords, P.gmask.p, wp, P.audit_mass.p};
            HIPCHK(audit_mass(s, na, aa));
        }
        stage_end(ctx, 12);
        ListArgs L{};
        pr = list_setup(L);
        if (pr) return pr;
        HIPCHK(ctx->reach.reserve(n));
        CandCoarseArgs cc{P.cand_m.p,  P.cand_cnt.p, M,          ctx->A16.p, ctx->A16lo.p, ctx->pq ? ctx->PSI / 2 : 0,
                          ctx->W16.p,  ctx->W16lo.p, ctx->inv_scale.p, ctx->At.p, ctx->reach.p, K16,
                          ctx->Mpad,   M,            T,          L,          ctx->emin_key.p};
        stage_begin(ctx, 2);
        HIPCHK(cand_coarse(s, prec, n, cc));
        stage_end(ctx, 2);
        audit_h.assign(na, 0.);
        cand_h.assign(n, 0);
        all_h.assign(n, 0);
        extra_rb = {{P.audit_mass.p, na ? audit_h.data() : nullptr, na * sizeof(double)},
                    {P.cand_cnt.p, cand_h.data(), n * sizeof(int)},
                    {P.cand_all.p, all_h.data(), n * sizeof(int)}};
        return finish(&L);
    };
    // the single pass (no oversampling) takes its significant diff2 back from the full pose term
    const bool use_list = ctx->coarse_mode == ROME_COARSE_LIST && ctx->two_pass;
    bool fell_back = false;
    if (learned) {
        rc = coarse_learned();
        if (rc) return rc;
        P.st_pools += 1;
        P.st_images += n;
        if (uncertain) {  // every image of the pool is counted as taking every rotation
            P.st_cand += (double)n * M;
            P.st_all += n;
            P.st_uncertain += 1;
            fell_back = true;
        } else {
            double mn = 1.0;
            for (double v : audit_h) mn = std::min(mn, v);
            for (int i = 0; i < n; ++i) {
                P.st_cand += cand_h[i];
                P.st_all += all_h[i];
            }
            P.st_audited += na;
            P.st_min_mass = std::min(P.st_min_mass, mn);
            if (mn < P.cfg.min_audit_mass) fell_back = true;
        }
    }
    // a pool off the proposal (or one the audit rejected): the exhaustive low-precision pass + recompute
    auto exhaustive_setup = [&]() {
        if (prec != ROME_PREC_FP32 && ctx->two_pass && !lowp_recompute) {
            lowp_recompute = true;
            wp.mark_margin = prec == ROME_PREC_BF16 ? kLowpRecomputeMarginBF16 : kLowpRecomputeMargin;
            wc.tail = ctx->c_tail.p;
        }
    };
    // The coarse significance given as keys (rome_expectation_fine_device): the task build reads them as
    // the weights pass would have left them; NR_SIGN = their count, the coarse outputs are NaN.
    auto given_sig = [&]() -> int {
        const int gc = ctx->given.cap;
        HIPCHK(ctx->sig_keys.reserve((size_t)n * gc));
        HIPCHK(ctx->sig_cnt.reserve(n));
        HIPCHK(hipMemcpyAsync(ctx->sig_keys.p, ctx->given.keys, (size_t)n * gc * sizeof(int),
                              hipMemcpyDeviceToDevice, s));
        HIPCHK(hipMemcpyAsync(ctx->sig_cnt.p, ctx->given.cnt, (size_t)n * sizeof(int), hipMemcpyDeviceToDevice, s));
        HIPCHK(hipMemcpyAsync(ctx->c_nsig.p, ctx->given.cnt, (size_t)n * sizeof(int), hipMemcpyDeviceToDevice, s));
        HIPCHK(launch_fill32(s, ctx->c_status.p, 0u, (size_t)n));
        for (auto* b : {&ctx->c_dmin, &ctx->c_sumw, &ctx->c_thr})  // all-ones words: NaN
            HIPCHK(launch_fill32(s, b->p, 0xffffffffu, 2 * (size_t)n));
        HIPCHK(launch_fill32(s, ctx->used.p, 0u, ((size_t)M + 3) / 4));
        const SigArgs sg{nullptr, ctx->TW, ctx->sig_keys.p, ctx->sig_cnt.p, gc};
        stage_begin(ctx, 4);
        HIPCHK(launch_check_given_keys(s, n, ctx->sig_keys.p, gc, ctx->sig_cnt.p, (long long)M * T, ctx->pool_flags.p));
        HIPCHK(launch_task_build(s, n, sg, M, T, ctx->ntask.p, ctx->nblk.p, ctx->task_off.p, ctx->blk_off.p,
                                 ctx->used.p, ctx->ulist.p, ctx->m2u.p, ctx->ucount.p));
        HIPCHK(launch_check_sorted_dups(s, n, ctx->sig_keys.p, gc, ctx->sig_cnt.p, ctx->pool_flags.p));
        stage_end(ctx, 4);
        sig_dense = false;
        if (dd) {
            HIPCHK(launch_pool_check(s, n, ctx->c_status.p, ctx->sig_cnt.p, gc, ctx->ucount.p, ctx->dd_ucap(n),
                                     ctx->pool_flags.p));
            return ROME_OK;
        }
        int flags[4] = {0, 0, 0, 0};
        rc = readback(ctx, {{ctx->task_off.p + n, &tot[0], sizeof(long long)},
                            {ctx->blk_off.p + n, &tot[1], sizeof(long long)},
                            {ctx->ucount.p, &U, sizeof(int)},
                            {ctx->pool_flags.p, flags, sizeof(flags)}});
        if (rc) return rc;
        if (flags[2])
            return fail(ctx, ROME_ERR_ARG,
                        "given coarse significance: an image with no key, more than %d, a key outside [0, %lld) "
                        "or a repeated key", gc, (long long)M * T);
        return ROME_OK;
    };
    if (given) {
        rc = given_sig();
        if (rc) return rc;
    } else if (!learned || fell_back) {
        if (fell_back) P.st_fallback += 1;
        exhaustive_setup();
        rc = coarse(use_list);
        if (rc) return rc;
    }
    if (ctx->sig_dense_left > 0) --ctx->sig_dense_left;
    ctx->last_list_max = 0;
    ctx->last_list_total = 0;
    for (int v : lcount) {
        ctx->last_list_max = std::max<long long>(ctx->last_list_max, v);
        ctx->last_list_total += v;
    }
    ctx->last_coarse_full = use_list ? 0 : 1;
    // Device-driven pools size the fine pass by capacities: at most sig_cap significant (or marked) samples,
    // hence tasks and blocks, per image; the counts stay on the device.
    const long long cap_tb = (long long)n * ctx->sig_cap;
    if (!dd && use_list && std::any_of(cstatus.begin(), cstatus.end(), [](int v) { return (v & 8) != 0; })) {
        ctx->last_coarse_full = 2;  // a list overflowed: the whole batch again with the full pose term
        exhaustive_setup();
        rc = coarse(false);
        if (rc) return rc;
    }
    ctx->last_marked = 0;
    if (lowp_recompute) {
        // tasks of the marked samples -> split three-product diff2 -> exact weights and significance
        if (!dd)
            for (int i = 0; i < n; ++i)
                if (cstatus[i] & 1) return fail(ctx, ROME_ERR_NUMERIC, "image %d: no coarse sample marked", i);
        const long long mt = dd ? cap_tb : tot[0], mb = dd ? cap_tb : tot[1];
        ctx->last_marked = dd ? -1 : mb;
        HIPCHK(sized(ctx->task_img, mt));
        HIPCHK(sized(ctx->task_m, mt));
        HIPCHK(sized(ctx->task_nt, mt));
        HIPCHK(sized(ctx->task_blk0, mt));
        HIPCHK(sized(ctx->blk_task, mb));
        HIPCHK(sized(ctx->blk_t, mb));
        HIPCHK(sized(ctx->D2c, mb));
        HIPCHK(ctx->c_dmin1.reserve(n));
        const SigArgs sg1 = sig_dense ? SigArgs{ctx->sigmask.p, ctx->TW, nullptr, nullptr, 0}
                                      : SigArgs{nullptr, ctx->TW, ctx->sig_keys.p, ctx->sig_cnt.p, ctx->sig_cap};
        HIPCHK(launch_task_emit(s, n, sg1, M, T, ctx->task_off.p, ctx->blk_off.p, ctx->task_img.p, ctx->task_m.p,
                                ctx->task_blk0.p, ctx->task_nt.p, ctx->blk_task.p, ctx->blk_t.p));
        RecomputeArgs ra{ctx->task_off.p, ctx->task_m.p, ctx->task_nt.p, ctx->task_blk0.p, ctx->blk_t.p,
                         ctx->A16.p,      ctx->A16lo.p,  ctx->pq ? ctx->PSI / 2 : 0,
                         ctx->W16.p,      ctx->W16lo.p,  ctx->inv_scale.p, ctx->At.p, ctx->C2c.p,
                         K16,             ctx->Mpad,     T,               ctx->D2c.p};
        stage_begin(ctx, 11);
        HIPCHK(launch_recompute_split(s, prec, n, ra));
        stage_end(ctx, 11);
        HIPCHK(hipMemcpyAsync(ctx->c_dmin1.p, ctx->c_dmin.p, n * sizeof(double), hipMemcpyDeviceToDevice, s));
        const bool dense =
            !dd && (ctx->sig_cap == 0 || ctx->sig_dense_left > 0 || (long long)M * T > 0x7fffffffLL);
        rc = significance(nullptr, dense, true);
        if (rc) return rc;
        // The significant samples are drawn from the marked ones, whose keys fit the cap when this pass
        // is sparse; a rerun here could not read D2c anyway (the task build above replaced its block lists).
        if (!dd && !dense && *std::max_element(scount.begin(), scount.end()) > ctx->sig_cap)
            return fail(ctx, ROME_ERR_NUMERIC, "recomputed significant set exceeds the marked set's key cap");
    }
    if (!dd)
        for (int i = 0; i < n; ++i)
            if (cstatus[i] & 1)
                return fail(ctx, ROME_ERR_NUMERIC, "image %d: my_nr_significant_coarse_samples == 0", i);
    const long long ntasks = dd ? cap_tb : tot[0], nblocks = dd ? cap_tb : tot[1];
    if (dd) U = ctx->dd_ucap(n);  // capacity of the unique rotations (more aborts the pool)
    ctx->last_n = n;
    ctx->last_tasks = ntasks;
    ctx->last_blocks = nblocks;
    const int NOR = ctx->NORf, NOT = ctx->NOTf, NE = NOR * NOT;
    HIPCHK(sized(ctx->task_img, ntasks));
    HIPCHK(sized(ctx->task_m, ntasks));
    HIPCHK(sized(ctx->task_nt, ntasks));
    HIPCHK(sized(ctx->task_blk0, ntasks));
    HIPCHK(sized(ctx->blk_task, nblocks));
    HIPCHK(sized(ctx->blk_t, nblocks));
    const SigArgs sg = sig_dense ? SigArgs{ctx->sigmask.p, ctx->TW, nullptr, nullptr, 0}
                                 : SigArgs{nullptr, ctx->TW, ctx->sig_keys.p, ctx->sig_cnt.p, ctx->sig_cap};
    HIPCHK(launch_task_emit(s, n, sg, M, T, ctx->task_off.p, ctx->blk_off.p, ctx->task_img.p, ctx->task_m.p,
                            ctx->task_blk0.p, ctx->task_nt.p, ctx->blk_task.p, ctx->blk_t.p));
    HIPCHK(sized(ctx->D2, (size_t)nblocks * NE));
    HIPCHK(sized(ctx->Pfine, (size_t)U * NOR * ctx->Pf));
    stage_begin(ctx, 5);
    HIPCHK(launch_project_fine(s, ctx->grids.p, ctx->is3d ? ctx->cube.p : nullptr, ctx->grid_n, ctx->pad, ctx->padF,
                               ctx->cgrid, ctx->is3d, ctx->Ai_f.p, ctx->R, NOR, ctx->ulist.p, U,
                               dd ? ctx->ucount.p : nullptr, ctx->work_counters(0), ctx->persist_blocks(1),
                               ctx->guess_U(n, U), ctx->f_xy.p, ctx->Pf, ctx->r_max, ctx->Pfine.p));
    stage_end(ctx, 5);
    FineArgs fa;
    fa.task_img = ctx->task_img.p;
    fa.task_m = ctx->task_m.p;
    fa.task_blk0 = ctx->task_blk0.p;
    fa.task_nt = ctx->task_nt.p;
    fa.blk_t = ctx->blk_t.p;
    fa.m2u = ctx->m2u.p;
    fa.Pfine = ctx->Pfine.p;
    fa.Xf = ctx->Xf.p;
    fa.Minvf = ctx->Minvf.p;
    fa.Ctff = ctx->Ctff.p;
    fa.Sf = ctx->Sf.p;
    fa.C2f = ctx->C2f.p;
    fa.D2 = ctx->D2.p;
    fa.Pf = ctx->Pf;
    if (dd) {
        fa.ntask_dev = ctx->task_off.p + n;
        fa.queue = ctx->pool_flags.p + 1;  // zeroed with the other flags at the start of the pool
        fa.abort = ctx->pool_flags.p;
    }
    // Two passes: the fine pass over the coarse-significant samples (src/map3d_optimizer.cpp:1261-1372).
    // Single pass (oversampling 0): the same kernels recompute, with NOR = NOT = 1, the diff2 of the
    // samples the FP32 cross term marked significant (with kRecomputeMargin nats of slack); their weights,
    // threshold and NR_SIGN are then formed from those values plus the mass of the unmarked rest, so the
    // sums see the diff2 of per-pixel products with an FP64 reduction instead of the cross term's long
    // FP32 accumulation (src/map2d_optimizer.cpp:1006-1047 computes them once, directly, in FP64).
    WeightOutArgs wf{ctx->f_dmin.p, ctx->f_sumw.p, ctx->f_maxw.p, ctx->f_amax.p, ctx->f_thr.p,
                     ctx->two_pass ? nullptr : ctx->c_nsig.p, ctx->f_status.p, nullptr};
    const TailArgs tail = ctx->two_pass ? TailArgs{nullptr, nullptr} : TailArgs{ctx->c_tail.p, ctx->c_dmin.p};
    stage_begin(ctx, 6);
    // device-driven: one block per task up to the expected task total, the rest through the work queue
    const long long first = dd ? std::min<long long>(0x7fffffffLL, ctx->guess_tasks(n, ntasks)) : 0;
    HIPCHK(launch_fine_diff2(s, (int)ntasks, NOR, NOT, fa, ctx->persist_blocks(0), (int)first));
    stage_end(ctx, 6);
    stage_begin(ctx, 7);
    HIPCHK(launch_weights_fine(s, n, wp, ctx->D2.p, ctx->blk_task.p, ctx->blk_t.p, ctx->task_m.p, ctx->blk_off.p, NE, wf,
                               cand, tail));
    stage_end(ctx, 7);
    if (!given && P.ready && P.mode == ROME_PROPOSAL_COLLECT && ctx->two_pass && ctx->is3d && P.K == ctx->K && P.D == ctx->D &&
        P.PSI == ctx->PSI && P.cur == c.current_size) {
        rc = prop_collect(ctx, n);  // the pool's posteriors become training targets
        if (rc) return rc;
    }
    const double* thr_p = ctx->f_thr.p;
    const double* sum_p = ctx->f_sumw.p;
    const double* dmin_p = ctx->f_dmin.p;
    // ---- updateOtherParams + backProjection
    if (ctx->post_enabled) {
        PostArgs pp;
        pp.K = ctx->K;
        pp.D = ctx->D;
        pp.PSI = ctx->PSI;
        pp.R = ctx->R;
        pp.is3d = ctx->is3d;
        pp.Pf = ctx->Pf;
        pp.oriF = ctx->oriF;
        pp.pad = ctx->pad;
        pp.padF = ctx->padF;
        pp.c = ctx->cgrid;
        pp.r_max = ctx->r_max;
        pp.grid_stride = ctx->grid_n;
        pp.group = ctx->in.group;
        pp.sumw = sum_p;
        pp.thr = thr_p;
        pp.dmin = dmin_p;
        pp.task_off = ctx->task_off.p;
        pp.task_img = ctx->task_img.p;
        pp.task_m = ctx->task_m.p;
        pp.task_nt = ctx->task_nt.p;
        pp.task_blk0 = ctx->task_blk0.p;
        pp.blk_t = ctx->blk_t.p;
        pp.m2u = ctx->m2u.p;
        pp.D2 = ctx->D2.p;
        pp.pdf_dir = ctx->pdf_dir.p;
        pp.trans_x = ctx->d_tx.p;
        pp.trans_y = ctx->d_ty.p;
        pp.shift_x = ctx->two_pass ? ctx->d_sxf.p : ctx->d_sxc.p;
        pp.shift_y = ctx->two_pass ? ctx->d_syf.p : ctx->d_syc.p;
        pp.old_offx = ctx->in.offx;
        pp.old_offy = ctx->in.offy;
        pp.prior_x = ctx->prior_x.p;
        pp.prior_y = ctx->prior_y.p;
        pp.sigma2_offset = c.sigma2_offset;
        pp.Ai_fine = ctx->Ai_f.p;
        pp.Pfine = ctx->Pfine.p;
        pp.Xf = ctx->Xf.p;
        pp.Ctff = ctx->Ctff.p;
        pp.NMf = ctx->NMf.p;
        pp.MC2bp = ctx->MC2bp.p;
        pp.Sf = ctx->Sf.p;
        pp.f_xy = ctx->f_xy.p;
        pp.f_ires = ctx->f_ires.p;
        pp.bp_data = ctx->bp_data.p;
        pp.bp_weight = ctx->bp_weight.p;
        pp.wsum_sigma2 = ctx->wsum_sigma2.p;
        pp.wsum_pdf_class = ctx->wsum_pdf_class.p;
        pp.wsum_pdf_dir = ctx->wsum_pdf_dir.p;
        pp.wsum_sumw_group = ctx->wsum_sumw_group.p;
        pp.wsum_prior_x = ctx->wsum_prior_x.p;
        pp.wsum_prior_y = ctx->wsum_prior_y.p;
        pp.wsum_scalars = ctx->wsum_scalars.p;
        pp.wnorm = ctx->wnorm.p;
        // 3D: slices to HBM, then the binned scatter (bounded by a slice buffer of kBinnedBytes).
        // Device-driven pools: the slice buffer has a fixed capacity; the per-slice kernel and the binned
        // scatter take a pool whose slices fit it, the direct scatter (k_post) one whose slices do not.
        long long S = ntasks * NOR;
        const long long slice_cap = ctx->dd_slice_cap(n);
        if (dd) S = post_slices_supported(NOR, NOT, ctx->Pf, ctx->is3d) ? slice_cap : 0;
        const bool binned = ctx->is3d && S > 0 && (unsigned long long)S * ctx->Pf < (1ULL << 32) &&
                            (size_t)S * ctx->Pf * 16 <= rome_ctx::kBinnedBytes;
        if (dd) {
            pp.ntask_dev = ctx->task_off.p + n;
            pp.abort = ctx->pool_flags.p;
            pp.slice_cap = binned ? slice_cap : -1;  // -1: the direct scatter takes every pool
            pp.xq = ctx->work_counters(8);
        }
        pp.slice_val = nullptr;
        pp.slice_w = nullptr;
        if (binned) {
            HIPCHK(sized(ctx->slice_val, (size_t)S * ctx->Pf));
            HIPCHK(sized(ctx->slice_w, (size_t)S));
            pp.slice_val = ctx->slice_val.p;
            pp.slice_w = ctx->slice_w.p;
        }
        stage_begin(ctx, 8);
        hipError_t pe = hipErrorNotSupported;
        if (dd) {
            HIPCHK(launch_fill32(s, ctx->wnorm.p, 0u, 2 * (size_t)n));
            if (binned)
                HIPCHK(launch_post_slices(s, ntasks, NOR, NOT, pp, ctx->persist_blocks(2), ctx->guess_tasks(n, ntasks)));
            PostArgs pd = pp;  // the direct scatter of pools past the slice buffer
            pd.slice_val = nullptr;
            pd.slice_w = nullptr;
            pe = launch_post(s, n, NOR, NOT, pd);
        } else {
            if (binned) {
                HIPCHK(launch_fill32(s, ctx->wnorm.p, 0u, 2 * (size_t)n));
                pe = launch_post_slices(s, ntasks, NOR, NOT, pp);
            }
            if (pe == hipErrorNotSupported) pe = launch_post(s, n, NOR, NOT, pp);
        }
        HIPCHK(pe);
        stage_end(ctx, 8);
        if (binned) {
            BpArgs ba;
            ba.K = ctx->K;
            ba.Pf = ctx->Pf;
            ba.NOR = NOR;
            ba.R = ctx->R;
            ba.pad = ctx->pad;
            ba.padF = ctx->padF;
            ba.c = ctx->cgrid;
            ba.r_max = ctx->r_max;
            ba.grid_stride = ctx->grid_n;
            ba.nbx = (ctx->padF + 7) / 8;
            ba.nby = (ctx->pad + 7) / 8;
            ba.nbz = (ctx->pad + 7) / 8;
            ba.nb_class = ba.nbx * ba.nby * ba.nbz;
            const size_t nb = (size_t)ba.nb_class * ctx->K;
            HIPCHK(ctx->bp_count.reserve(nb));
            HIPCHK(ctx->bp_off.reserve(nb + 1));
            HIPCHK(ctx->bp_coff.reserve(nb + 1));
            HIPCHK(ctx->bp_cursor.reserve(nb));
            HIPCHK(sized(ctx->bp_rec, (size_t)S * ctx->Pf));
            ba.task_img = ctx->task_img.p;
            ba.task_m = ctx->task_m.p;
            ba.Ai_fine = ctx->Ai_f.p;
            ba.f_xy = ctx->f_xy.p;
            ba.MC2bp = ctx->MC2bp.p;
            ba.slice_w = ctx->slice_w.p;
            ba.slice_val = ctx->slice_val.p;
            ba.count = ctx->bp_count.p;
            ba.off = ctx->bp_off.p;
            ba.coff = ctx->bp_coff.p;
            ba.cursor = ctx->bp_cursor.p;
            ba.rec = ctx->bp_rec.p;
            ba.bp_data = ctx->bp_data.p;
            ba.bp_weight = ctx->bp_weight.p;
            if (dd) {
                ba.ntask_dev = ctx->task_off.p + n;
                ba.slice_cap = slice_cap;
                ba.abort = ctx->pool_flags.p;
            }
            stage_begin(ctx, 10);
            HIPCHK(launch_backproject_binned(s, S, ba, ctx->persist_blocks(3)));
            stage_end(ctx, 10);
        }
    } else {
        HIPCHK(launch_fill32(s, ctx->wnorm.p, 0u, 2 * (size_t)n));
    }
    // ---- metadata of the last pass (on the device; read back with the results)
    MetaArgs ma;
    ma.n = n;
    ma.is3d = ctx->is3d;
    ma.R = ctx->R;
    ma.PSI = ctx->PSI;
    ma.NOR = NOR;
    ma.NOT = NOT;
    ma.T = T;
    ma.Mtot = M;
    ma.amax = ctx->f_amax.p;
    ma.img_blk = ctx->blk_off.p;
    ma.blk_task = ctx->blk_task.p;
    ma.blk_t = ctx->blk_t.p;
    ma.task_m = ctx->task_m.p;
    ma.maxw = ctx->f_maxw.p;
    ma.sumw = sum_p;
    ma.offx = ctx->in.offx;
    ma.offy = ctx->in.offy;
    ma.pose = ctx->two_pass ? ctx->d_pose_f.p : ctx->d_pose_c.p;
    ma.sx = ctx->two_pass ? ctx->d_sxf.p : ctx->d_sxc.p;
    ma.sy = ctx->two_pass ? ctx->d_syf.p : ctx->d_syc.p;
    ma.meta = ctx->meta.p;
    HIPCHK(ctx->argmax.reserve((size_t)6 * n));
    ma.argmax = ctx->argmax.p;
    HIPCHK(launch_meta(s, ma));
    // ---- storeWeightedSums
    StoreArgs sa;
    sa.n = n;
    sa.curF = ctx->fs / 2 + 1;
    sa.oriF = ctx->oriF;
    sa.do_norm = c.do_norm_correction;
    sa.avg_norm = c.avg_norm_correction;
    sa.group = ctx->in.group;
    sa.power = ctx->in.power;
    sa.norm = ctx->in.norm;
    sa.sumw = sum_p;
    sa.dmin = dmin_p;
    sa.logsigma2 = ctx->logsigma2.p;
    sa.wnorm = ctx->wnorm.p;
    sa.meta = ctx->meta.p;
    sa.wsum_sigma2 = ctx->wsum_sigma2.p;
    sa.wsum_scalars = ctx->wsum_scalars.p;
    if (dd) sa.abort = ctx->pool_flags.p;
    stage_begin(ctx, 9);
    HIPCHK(launch_store_sums(s, sa));
    stage_end(ctx, 9);
    if (dd) {  // everything the host needs in one read-back; the caller synchronises (or replays)
        ctx->last_U = U;
        return readback_begin(ctx, pool_segs(ctx, n, *dout), *dlayout, true);
    }
    std::vector<double> sumw(n);
    {
        const size_t nd = n * sizeof(double);
        rc = readback(ctx, {{sum_p, sumw.data(), nd},
                            {ctx->meta.p, res ? res->meta : nullptr, 9 * nd},
                            {ctx->argmax.p, res ? res->argmax : nullptr, 6 * n * sizeof(int)},
                            {ctx->c_nsig.p, res ? res->nr_sign : nullptr, n * sizeof(int)},
                            {dmin_p, res ? res->min_diff2 : nullptr, nd},
                            {sum_p, res ? res->sum_weight : nullptr, nd},
                            {thr_p, res ? res->sig_weight : nullptr, nd},
                            // a single pass is its own coarse pass: its exact values
                            {ctx->two_pass ? ctx->c_dmin.p : dmin_p, res ? res->min_diff2_coarse : nullptr, nd},
                            {ctx->two_pass ? ctx->c_sumw.p : sum_p, res ? res->sum_weight_coarse : nullptr, nd},
                            {ctx->two_pass ? ctx->c_thr.p : thr_p, res ? res->sig_weight_coarse : nullptr, nd},
                            {ctx->wnorm.p, res ? res->wsum_norm_correction : nullptr, nd}});
        if (rc) return rc;
    }
    for (int i = 0; i < n; ++i)
        if (!(sumw[i] > 0)) return fail(ctx, ROME_ERR_NUMERIC, "image %d: exp_sum_weight == 0", i);
    HIPCHK(hipStreamSynchronize(s));
    stage_collect(ctx);
    ctx->last_U = U;
    ctx->note_counts(ntasks, U);
    return ROME_OK;
}

void graph_release(rome_ctx::PoolGraph& g) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    g.exec = nullptr;
}

void graphs_clear(rome_ctx* ctx) {
    for (auto& g : ctx->graphs) graph_release(g);
    ctx->graphs.clear();
}

// The cached graph of a pool with these inputs (nullptr: none).  Graphs of older allocation or state
// epochs are released.
rome_ctx::PoolGraph* graph_find(rome_ctx* ctx, int n, const void* const in[rome_ctx::kGraphKeys], bool create) {
    const unsigned long long ae = g_alloc_epoch.load();
    for (size_t i = 0; i < ctx->graphs.size(); ++i) {
        auto& g = ctx->graphs[i];
        if (g.alloc_epoch != ae || g.state_epoch != ctx->state_epoch) {
            graph_release(g);
            ctx->graphs.erase(ctx->graphs.begin() + (long)i);
            --i;
            continue;
        }
        if (g.n == n && g.prec == ctx->cfg.precision && std::equal(in, in + rome_ctx::kGraphKeys, g.in)) return &g;
    }
    if (!create) return nullptr;
    if (ctx->graphs.size() >= 16) {  // least recently used first
        graph_release(ctx->graphs.front());
        ctx->graphs.erase(ctx->graphs.begin());
    }
    ctx->graphs.emplace_back();
    auto& g = ctx->graphs.back();
    g.n = n;
    g.prec = ctx->cfg.precision;
    std::copy(in, in + rome_ctx::kGraphKeys, g.in);
    g.alloc_epoch = ae;
    g.state_epoch = ctx->state_epoch;
    return &g;
}

// Records the device-driven pool (same buffers, same inputs as the pool that just ran) into a graph.  A
// failure leaves no graph and marks the key so that it is not tried again; it is not an error of the pool.
void graph_capture(rome_ctx* ctx, int n, rome_ctx::PoolGraph& g) {
    hipStream_t s = ctx->stream;
    if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        (void)hipGetLastError();
        g.failed = true;
        return;
    }
    g_capturing = true;
    PoolOut tmp;
    RbLayout L;
    const std::string err = ctx->err;
    const int rc = run_pool(ctx, n, nullptr, kPoolCapture, &tmp, &L);
    g_capturing = false;
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &graph);
    ctx->err = err;
    if (rc || e != hipSuccess || !graph ||
        hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGetLastError();
        g.exec = nullptr;
        g.failed = true;
    } else {
        g.rb_bytes = L.bytes;
        ++ctx->graph_captures;
    }
    if (graph) (void)hipGraphDestroy(graph);
    // the epochs the graph was captured in (capture itself allocates nothing)
    g.alloc_epoch = g_alloc_epoch.load();
    g.state_epoch = ctx->state_epoch;
}

// One pool: host-synchronised, device-driven, or a replay of its captured graph (rome_set_execution).
int run_estep(rome_ctx* ctx, int n, rome_results* res) {
    const ProposalState& P = ctx->prop;
    const bool prop_shape = P.ready && ctx->two_pass && ctx->is3d && P.K == ctx->K && P.D == ctx->D &&
                            P.PSI == ctx->PSI && P.cur == ctx->cfg.current_size;
    const bool prop_on = prop_shape && (P.mode == ROME_PROPOSAL_COLLECT ||
                                        (P.mode == ROME_PROPOSAL_USE && ctx->cfg.precision != ROME_PREC_FP32));
    const bool dd_ok = ctx->exec_mode != ROME_EXEC_HOST && ctx->configured && ctx->two_pass &&
                       ctx->coarse_mode == ROME_COARSE_LIST && ctx->sig_cap > 0 && ctx->sig_dense_left == 0 &&
                       (long long)ctx->Mtot * ctx->T <= 0x7fffffffLL && !prop_on;
    if (!dd_ok) {
        ctx->last_exec = 0;
        return run_pool(ctx, n, res, kPoolHost, nullptr, nullptr);
    }
    PoolOut& o = ctx->pool_out;
    hipStream_t s = ctx->stream;
    const bool use_graph = ctx->exec_mode == ROME_EXEC_GRAPH && !ctx->profiling;
    const void* in[rome_ctx::kGraphKeys] = {ctx->in.fmask, ctx->in.fnomask, ctx->in.fctf, ctx->in.hxi2, ctx->in.power,
                                  ctx->in.offx, ctx->in.offy, ctx->in.norm, ctx->in.group, ctx->given.keys,
                                  ctx->given.cnt, reinterpret_cast<const void*>((uintptr_t)ctx->sig_cap)};
    rome_ctx::PoolGraph* g = use_graph ? graph_find(ctx, n, in, false) : nullptr;
    RbLayout L;
    int rc;
    bool capture_after = false;
    if (g && g->exec) {
        o.size(n);
        const std::vector<Seg> segs = pool_segs(ctx, n, o);
        rb_layout(segs, L, true);
        if (L.bytes != g->rb_bytes) return fail(ctx, ROME_ERR_HIP, "captured pool read-back layout changed");
        const auto t0 = std::chrono::steady_clock::now();
        HIPCHK(hipGraphLaunch(g->exec, s));
        const auto t1 = std::chrono::steady_clock::now();
        rc = readback_end(ctx, segs, L);
        ctx->last_launch_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        ctx->last_wait_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t1).count();
        if (rc) return rc;
        ++g->hits;
        ++ctx->graph_launches;
        ctx->last_exec = 2;
    } else {
        const auto t0 = std::chrono::steady_clock::now();
        rc = run_pool(ctx, n, res, kPoolDevice, &o, &L);
        if (rc) {
            (void)hipStreamSynchronize(s);  // launches before the failure may still run
            return rc;
        }
        const auto t1 = std::chrono::steady_clock::now();
        rc = readback_end(ctx, pool_segs(ctx, n, o), L);
        ctx->last_launch_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        ctx->last_wait_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t1).count();
        if (rc) return rc;
        stage_collect(ctx);
        ctx->last_exec = 1;
        capture_after = use_graph && !(g && g->failed);
    }
    if (o.flags[0]) {  // a list or key overflow, no significant sample, or too many rotations: host path
        ++ctx->dd_reruns;
        const int le = ctx->last_exec;
        rc = run_pool(ctx, n, res, kPoolHost, nullptr, nullptr);
        ctx->last_exec = le | 8;
        return rc;
    }
    for (int i = 0; i < n; ++i)
        if (!(o.sumw[i] > 0)) return fail(ctx, ROME_ERR_NUMERIC, "image %d: exp_sum_weight == 0", i);
    ctx->note_counts(o.tot[0], o.U);  // the launch sizes of the next pools (and of a capture) follow them
    if (capture_after) {  // before the bookkeeping below: a capture pass sets it from capacities
        g = graph_find(ctx, n, in, true);
        if (!g->exec && !g->failed) graph_capture(ctx, n, *g);
    }
    if (res) {
        auto put = [n](void* dst, const void* src, size_t per) {
            if (dst) std::memcpy(dst, src, per * n);
        };
        put(res->meta, o.meta.data(), 9 * sizeof(double));
        put(res->argmax, o.argmax.data(), 6 * sizeof(int));
        put(res->nr_sign, o.nsig.data(), sizeof(int));
        put(res->min_diff2, o.dmin.data(), sizeof(double));
        put(res->sum_weight, o.sumw.data(), sizeof(double));
        put(res->sig_weight, o.thr.data(), sizeof(double));
        put(res->min_diff2_coarse, o.dmin_c.data(), sizeof(double));
        put(res->sum_weight_coarse, o.sumw_c.data(), sizeof(double));
        put(res->sig_weight_coarse, o.thr_c.data(), sizeof(double));
        put(res->wsum_norm_correction, o.wnorm.data(), sizeof(double));
    }
    ctx->last_n = n;
    ctx->last_tasks = o.tot[0];
    ctx->last_blocks = o.tot[1];
    ctx->last_U = o.U;
    ctx->last_list_max = 0;
    ctx->last_list_total = 0;
    for (int v : o.lcount) {
        ctx->last_list_max = std::max<long long>(ctx->last_list_max, v);
        ctx->last_list_total += v;
    }
    ctx->last_coarse_full = 0;
    ctx->last_sig_form = 0;
    ctx->last_sig_max = *std::max_element(o.scount.begin(), o.scount.end());
    return ROME_OK;
}

}  // namespace

static int install_model(rome_ctx* ctx);

extern "C" {

int rome_set_execution(rome_ctx* ctx, int mode) {
    if (!ctx || mode < ROME_EXEC_HOST || mode > ROME_EXEC_GRAPH) return ROME_ERR_ARG;
    ctx->exec_mode = mode;
    return ROME_OK;
}

int rome_get_execution_stats(rome_ctx* ctx, long long* out) {
    if (!ctx || !out) return ROME_ERR_ARG;
    out[0] = ctx->last_exec;
    out[1] = ctx->graph_launches;
    out[2] = ctx->graph_captures;
    out[3] = ctx->dd_reruns;
    out[4] = (long long)(ctx->last_launch_ms * 1000.);
    out[5] = (long long)(ctx->last_wait_ms * 1000.);
    return ROME_OK;
}

int rome_create(int device, rome_ctx** out) {
    if (!out) return ROME_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return ROME_ERR_HIP;
    if (device < 0 || device >= ndev) return ROME_ERR_ARG;
    rome_ctx* ctx = new rome_ctx();
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return ROME_ERR_HIP;
    }
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && ncu > 0)
        ctx->n_cu = ncu;
    // the device address of the allocation-free entry point's staging array, once per device
    if (noalloc_init(device) != hipSuccess) {
        rome_destroy(ctx);
        return ROME_ERR_HIP;
    }
    // 5000-entry sin/cos table of shiftImageInFourierTransform (src/image.h:1486-1510)
    std::vector<double> st(5000), ct(5000);
    const double step = 2 * 3.14159265358979323846 / 5000.0;
    for (int i = 0; i < 5000; ++i) {
        st[i] = sin((double)i * step);
        ct[i] = cos((double)i * step);
    }
    if (ctx->sintab.upload(st.data(), 5000, ctx->stream) != hipSuccess ||
        ctx->costab.upload(ct.data(), 5000, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess) {
        rome_destroy(ctx);
        return ROME_ERR_HIP;
    }
    *out = ctx;
    return ROME_OK;
}

void rome_destroy(rome_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto& st : ctx->ev)
        for (auto& pr : st)
            for (hipEvent_t e : pr)
                if (e) (void)hipEventDestroy(e);
    for (auto& p : ctx->fft_plans) (void)hipfftDestroy(p.h);
    graphs_clear(ctx);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char* rome_last_error(const rome_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int rome_configure(rome_ctx* ctx, const rome_config* cfg) {
    if (ctx) ++ctx->state_epoch;  // kernels of captured pools read what this call changes
    if (!ctx || !cfg) return ROME_ERR_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    const rome_config& c = *cfg;
    ctx->class_lo = 0;
    ctx->class_hi = 1 << 30;
    if (c.ref_dim != 2 && c.ref_dim != 3) return fail(ctx, ROME_ERR_ARG, "ref_dim must be 2 or 3");
    if (c.adaptive_oversampling < 0 || c.adaptive_oversampling > 1)
        return fail(ctx, ROME_ERR_UNSUPPORTED, "adaptive_oversampling must be 0 or 1");
    if (c.current_size <= 0 || c.current_size > c.ori_size || (c.current_size & 1) || c.coarse_size <= 0 ||
        c.coarse_size > c.current_size || (c.coarse_size & 1))
        return fail(ctx, ROME_ERR_ARG, "bad image sizes");
    if (c.nr_classes <= 0 || c.nr_classes > kMaxK || c.nr_groups <= 0) return fail(ctx, ROME_ERR_ARG, "bad class/group count");
    if (c.nr_trans <= 0 || c.nr_trans > kMaxT || c.nr_classes * c.nr_trans > kMaxKT)
        return fail(ctx, ROME_ERR_UNSUPPORTED, "nr_trans %d outside the supported range", c.nr_trans);
    if (c.ori_size / 2 + 1 > kMaxShell) return fail(ctx, ROME_ERR_UNSUPPORTED, "ori_size too large");
    if (c.ref_dim == 2 && c.nr_dir != 1) return fail(ctx, ROME_ERR_ARG, "2D requires nr_dir == 1");
    if (c.precision < ROME_PREC_FP32 || c.precision > ROME_PREC_FP16)
        return fail(ctx, ROME_ERR_ARG, "precision must be 0 (FP32), 1 (BF16) or 2 (FP16)");
    if (c.precision != ROME_PREC_FP32 && c.adaptive_oversampling == 0)  // the coarse diff2 would be final
        return fail(ctx, ROME_ERR_UNSUPPORTED, "BF16/FP16 coarse pass needs adaptive_oversampling = 1");
    ctx->cfg = c;
    ctx->is3d = c.ref_dim == 3;
    ctx->K = c.nr_classes;
    ctx->D = c.nr_dir;
    ctx->PSI = c.nr_psi;
    ctx->T = c.nr_trans;
    ctx->R = c.nr_dir * c.nr_psi;
    ctx->Mtot = ctx->K * ctx->R;
    ctx->Mpad = (ctx->Mtot + 255) / 256 * 256;  // 256-row tiles of the BF16/FP16 cross term
    if ((double)ctx->Mtot * ctx->Mtot * ctx->T >= 281474976710656.0)  // 2^48: coarse index division
        return fail(ctx, ROME_ERR_UNSUPPORTED, "coarse grid too large (%d rotations x %d translations)", ctx->Mtot,
                    ctx->T);
    ctx->two_pass = c.adaptive_oversampling > 0;
    ctx->post_enabled = !(ctx->is3d && !ctx->two_pass);  // 3D stores no fine weights without oversampling
    ctx->NORf = ctx->two_pass ? (ctx->is3d ? 8 : 2) : 1;
    ctx->NOTf = ctx->two_pass ? 4 : 1;
    ctx->oriF = c.ori_size / 2 + 1;
    ctx->fs = c.current_size;
    ctx->cs = ctx->two_pass ? c.coarse_size : c.current_size;
    ctx->F2f = ctx->fs * (ctx->fs / 2 + 1);
    ctx->r_max = std::min(c.current_size / 2, c.ori_size / 2);
    ctx->pad = 2 * (2 * ctx->r_max + 1) + 1;
    ctx->padF = ctx->pad / 2 + 1;
    ctx->cgrid = (int)((float)ctx->pad / 2.0);
    ctx->grid_n = ctx->is3d ? (size_t)ctx->pad * ctx->pad * ctx->padF : (size_t)ctx->pad * ctx->padF;
    ctx->TW = (ctx->T + 31) / 32;
    // pixel sets: coarse = Mresol_coarse > 0 (Minv != 0), fine = Mresol_fine >= 0
    std::vector<int> cn, ci, cxy, fn, fi, fxy;
    {
        const int sz = ctx->cs, F = sz / 2 + 1;
        for (int i = 0; i < sz; ++i)
            for (int j = 0; j < F; ++j) {
                const int ires = mresol(sz, i, j);
                if (ires <= 0) continue;
                const int y = (i < F) ? i : i - sz;
                const int row_f = y >= 0 ? y : y + ctx->fs;
                cn.push_back(row_f * (ctx->fs / 2 + 1) + j);
                ci.push_back(ires);
                cxy.push_back(pack_xy(j, y));
            }
    }
    {
        const int sz = ctx->fs, F = sz / 2 + 1;
        for (int i = 0; i < sz; ++i)
            for (int j = 0; j < F; ++j) {
                const int ires = mresol(sz, i, j);
                if (ires < 0) continue;
                const int y = (i < F) ? i : i - sz;
                fn.push_back(i * F + j);
                fi.push_back(ires);
                fxy.push_back(pack_xy(j, y));
            }
    }
    ctx->Pc = (int)cn.size();
    ctx->Pc_pad = (ctx->Pc + 7) / 8 * 8;
    ctx->Kq = (ctx->Pc + 15) / 16 * 16;
    ctx->Pf = (int)fn.size();
    if (ctx->Pf > 16 * 1024) return fail(ctx, ROME_ERR_UNSUPPORTED, "current_size too large for the post kernel");
    ctx->f_nfine_h = fn;
    ctx->f_xy_h = fxy;
    hipStream_t s = ctx->stream;
    HIPCHK(ctx->c_nfine.upload(cn.data(), cn.size(), s));
    HIPCHK(ctx->c_ires.upload(ci.data(), ci.size(), s));
    HIPCHK(ctx->c_xy.upload(cxy.data(), cxy.size(), s));
    HIPCHK(ctx->f_nfine.upload(fn.data(), fn.size(), s));
    HIPCHK(ctx->f_ires.upload(fi.data(), fi.size(), s));
    HIPCHK(ctx->f_xy.upload(fxy.data(), fxy.size(), s));
    HIPCHK(ctx->grids.reserve(ctx->grid_n * ctx->K));
    HIPCHK(ctx->bp_data.reserve(2 * ctx->grid_n * ctx->K));
    HIPCHK(ctx->bp_weight.reserve(ctx->grid_n * ctx->K));
    HIPCHK(ctx->wsum_sigma2.reserve((size_t)c.nr_groups * ctx->oriF));
    HIPCHK(ctx->wsum_pdf_class.reserve(ctx->K));
    HIPCHK(ctx->wsum_pdf_dir.reserve((size_t)ctx->K * ctx->D));
    HIPCHK(ctx->wsum_sumw_group.reserve(c.nr_groups));
    HIPCHK(ctx->wsum_prior_x.reserve(ctx->K));
    HIPCHK(ctx->wsum_prior_y.reserve(ctx->K));
    HIPCHK(ctx->wsum_scalars.reserve(4));
    ctx->have_ref.assign(ctx->K, 0);
    ctx->configured = true;
    ctx->have_sampling = ctx->have_model = false;
    ctx->fct_dirty = true;
    HIPCHK(hipStreamSynchronize(s));
    return rome_reset_sums(ctx);
}

int rome_set_sampling(rome_ctx* ctx, const double* pose_coarse, const double* pose_fine, const double* shift_coarse_x,
                      const double* shift_coarse_y, const double* shift_fine_x, const double* shift_fine_y,
                      const double* trans_x, const double* trans_y) {
    if (ctx) ++ctx->state_epoch;  // kernels of captured pools read what this call changes
    if (!ctx) return ROME_ERR_ARG;
    if (!ctx->configured) return fail(ctx, ROME_ERR_ARG, "rome_configure has not been called");
    if (!pose_coarse || !shift_coarse_x || !shift_coarse_y || !trans_x || !trans_y)
        return fail(ctx, ROME_ERR_ARG, "null sampling table");
    if (ctx->two_pass && (!pose_fine || !shift_fine_x || !shift_fine_y))
        return fail(ctx, ROME_ERR_ARG, "fine sampling tables required with oversampling");
    HIPCHK(hipSetDevice(ctx->device));
    const int per = ctx->is3d ? 3 : 1;
    ctx->pose_c.assign(pose_coarse, pose_coarse + (size_t)ctx->R * per);
    if (ctx->two_pass) ctx->pose_f.assign(pose_fine, pose_fine + (size_t)ctx->R * ctx->NORf * per);
    else ctx->pose_f.clear();
    ctx->sxc.assign(shift_coarse_x, shift_coarse_x + ctx->T);
    ctx->syc.assign(shift_coarse_y, shift_coarse_y + ctx->T);
    if (ctx->two_pass) {
        ctx->sxf.assign(shift_fine_x, shift_fine_x + (size_t)ctx->T * 4);
        ctx->syf.assign(shift_fine_y, shift_fine_y + (size_t)ctx->T * 4);
    }
    ctx->tx.assign(trans_x, trans_x + ctx->T);
    ctx->ty.assign(trans_y, trans_y + ctx->T);
    int rc = upload_poses(ctx);
    if (rc) return rc;
    hipStream_t s = ctx->stream;
    HIPCHK(ctx->d_sxc.upload(ctx->sxc.data(), ctx->T, s));
    HIPCHK(ctx->d_syc.upload(ctx->syc.data(), ctx->T, s));
    HIPCHK(ctx->d_pose_c.upload(ctx->pose_c.data(), ctx->pose_c.size(), s));
    if (ctx->two_pass) HIPCHK(ctx->d_pose_f.upload(ctx->pose_f.data(), ctx->pose_f.size(), s));
    if (ctx->two_pass) {
        HIPCHK(ctx->d_sxf.upload(ctx->sxf.data(), ctx->sxf.size(), s));
        HIPCHK(ctx->d_syf.upload(ctx->syf.data(), ctx->syf.size(), s));
    }
    HIPCHK(ctx->d_tx.upload(ctx->tx.data(), ctx->T, s));
    HIPCHK(ctx->d_ty.upload(ctx->ty.data(), ctx->T, s));
    rc = build_phase_tables(ctx);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(s));
    ctx->have_sampling = true;
    ctx->fct_dirty = true;
    return ROME_OK;
}

int rome_set_model(rome_ctx* ctx, const double* sigma2_noise, const double* pdf_class, const double* pdf_direction,
                   const double* prior_offx, const double* prior_offy) {
    if (ctx) ++ctx->state_epoch;  // kernels of captured pools read what this call changes
    if (!ctx) return ROME_ERR_ARG;
    if (!ctx->configured) return fail(ctx, ROME_ERR_ARG, "rome_configure has not been called");
    if (!sigma2_noise || !pdf_class || !pdf_direction || !prior_offx || !prior_offy)
        return fail(ctx, ROME_ERR_ARG, "null model array");
    HIPCHK(hipSetDevice(ctx->device));
    const int G = ctx->cfg.nr_groups, K = ctx->K, D = ctx->D;
    ctx->sigma2_h.assign(sigma2_noise, sigma2_noise + (size_t)G * ctx->oriF);
    for (double v : ctx->sigma2_h)
        if (!(v > 0.) || !std::isfinite(v)) return fail(ctx, ROME_ERR_NUMERIC, "sigma2_noise must be positive and finite");
    ctx->pdf_class_h.assign(pdf_class, pdf_class + K);
    ctx->pdf_dir_h.assign(pdf_direction, pdf_direction + (size_t)K * D);
    ctx->prior_x_h.assign(prior_offx, prior_offx + K);
    ctx->prior_y_h.assign(prior_offy, prior_offy + K);
    return install_model(ctx);
}

// The device tables of the model the host copies hold, restricted to the classes of rome_set_class_range.
static int install_model(rome_ctx* ctx) {
    HIPCHK(hipSetDevice(ctx->device));
    const int G = ctx->cfg.nr_groups, K = ctx->K, D = ctx->D;
    const double* pdf_class = ctx->pdf_class_h.data();
    const double* pdf_direction = ctx->pdf_dir_h.data();
    std::vector<double> lp((size_t)K * D);
    std::vector<unsigned char> valid((size_t)ctx->Mpad, 0);
    std::vector<double> pm((size_t)ctx->Mtot), lpm((size_t)ctx->Mtot);
    std::vector<int> km((size_t)ctx->Mtot);
    std::vector<float2> mrow((size_t)ctx->Mpad, make_float2(-INFINITY, 0.f));  // padding rows: no prior
    for (int k = 0; k < K; ++k)
        for (int d = 0; d < D; ++d) {
            const double p = pdf_direction[(size_t)k * D + d];
            lp[(size_t)k * D + d] = p > 0 ? log(p) : -INFINITY;
            // the reference skips (class, direction) pairs with a non-positive prior
            // (src/map3d_optimizer.cpp:1204-1206): no diff2, no weight, no significance entry; and the pool
            // runs only classes exp_iclass_min .. exp_iclass_max (src/map3d_optimizer.cpp:823-840)
            const bool ok = pdf_class[k] > 0. && p > 0. && k >= ctx->class_lo && k <= ctx->class_hi;
            for (int s = 0; s < ctx->PSI; ++s) {
                const size_t m = (size_t)k * ctx->R + d * ctx->PSI + s;
                valid[m] = ok;
                pm[m] = ok ? p : 0.;
                lpm[m] = ok ? log(p) : -INFINITY;
                km[m] = k;
                const int kT = k * ctx->T;
                float kTf;
                std::memcpy(&kTf, &kT, sizeof kTf);
                mrow[m] = make_float2(ok ? (float)lpm[m] : -INFINITY, kTf);
            }
        }
    HIPCHK(ctx->pdf_m.upload(pm.data(), pm.size(), ctx->stream));
    HIPCHK(ctx->lpdf_m.upload(lpm.data(), lpm.size(), ctx->stream));
    HIPCHK(ctx->k_m.upload(km.data(), km.size(), ctx->stream));
    HIPCHK(ctx->mrow.upload(mrow.data(), mrow.size(), ctx->stream));
    std::vector<float> lpd_row(mrow.size());
    for (size_t m = 0; m < mrow.size(); ++m) lpd_row[m] = mrow[m].x;
    HIPCHK(ctx->lpd_row.upload(lpd_row.data(), lpd_row.size(), ctx->stream));
    ctx->lpdf_max = -INFINITY;
    for (double v : lpm) ctx->lpdf_max = std::max(ctx->lpdf_max, v);
    // logsigma2 per group (storeWeightedSums, src/map3d_optimizer.cpp:1998-2008): same loop order as the reference
    std::vector<double> ls(G, 0.);
    const int sz = ctx->fs, F = sz / 2 + 1;
    for (int g = 0; g < G; ++g) {
        double acc = 0.;
        for (int i = 0; i < sz; ++i)
            for (int j = 0; j < F; ++j) {
                const int ires = mresol(sz, i, j);
                if (ires > 0) acc += log(2. * 3.14159265358979323846 * ctx->sigma2_h[(size_t)g * ctx->oriF + ires]);
            }
        ls[g] = acc;
    }
    hipStream_t s = ctx->stream;
    HIPCHK(ctx->sigma2.upload(ctx->sigma2_h.data(), ctx->sigma2_h.size(), s));
    HIPCHK(ctx->pdf_dir.upload(ctx->pdf_dir_h.data(), ctx->pdf_dir_h.size(), s));
    HIPCHK(ctx->lpdf_dir.upload(lp.data(), lp.size(), s));
    HIPCHK(ctx->prior_x.upload(ctx->prior_x_h.data(), K, s));
    HIPCHK(ctx->prior_y.upload(ctx->prior_y_h.data(), K, s));
    HIPCHK(ctx->valid.upload(valid.data(), valid.size(), s));
    HIPCHK(ctx->logsigma2.upload(ls.data(), G, s));
    HIPCHK(hipStreamSynchronize(s));
    ctx->have_model = true;
    return ROME_OK;
}

int rome_set_class_range(rome_ctx* ctx, int iclass_min, int iclass_max) {
    if (ctx) ++ctx->state_epoch;  // kernels of captured pools read what this call changes
    if (!ctx) return ROME_ERR_ARG;
    if (!ctx->configured) return fail(ctx, ROME_ERR_ARG, "rome_configure has not been called");
    if (iclass_min < 0 && iclass_max < 0) {
        iclass_min = 0;
        iclass_max = ctx->K - 1;
    }
    if (iclass_min < 0 || iclass_max >= ctx->K || iclass_min > iclass_max)
        return fail(ctx, ROME_ERR_ARG, "class range %d .. %d outside 0 .. %d", iclass_min, iclass_max, ctx->K - 1);
    ctx->class_lo = iclass_min;
    ctx->class_hi = iclass_max;
    return ctx->have_model ? install_model(ctx) : ROME_OK;
}

int rome_set_reference(rome_ctx* ctx, int iclass, const double* grid) {
    if (ctx) ++ctx->state_epoch;  // kernels of captured pools read what this call changes
    if (!ctx || !grid) return ROME_ERR_ARG;
    if (!ctx->configured) return fail(ctx, ROME_ERR_ARG, "rome_configure has not been called");
    if (iclass < 0 || iclass >= ctx->K) return fail(ctx, ROME_ERR_ARG, "class index out of range");
    HIPCHK(hipSetDevice(ctx->device));
    std::vector<float2> h(ctx->grid_n);
    for (size_t q = 0; q < ctx->grid_n; ++q) h[q] = make_float2((float)grid[2 * q], (float)grid[2 * q + 1]);
    HIPCHK(hipMemcpyAsync(ctx->grids.p + ctx->grid_n * iclass, h.data(), ctx->grid_n * sizeof(float2),
                          hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->have_ref[iclass] = 1;
    ctx->fct_dirty = true;
    return ROME_OK;
}

int rome_reset_sums(rome_ctx* ctx) {
    if (!ctx) return ROME_ERR_ARG;
    if (!ctx->configured) return fail(ctx, ROME_ERR_ARG, "rome_configure has not been called");
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    HIPCHK(hipMemsetAsync(ctx->bp_data.p, 0, 2 * ctx->grid_n * ctx->K * sizeof(float), s));
    HIPCHK(hipMemsetAsync(ctx->bp_weight.p, 0, ctx->grid_n * ctx->K * sizeof(float), s));
    HIPCHK(hipMemsetAsync(ctx->wsum_sigma2.p, 0, (size_t)ctx->cfg.nr_groups * ctx->oriF * sizeof(double), s));
    HIPCHK(hipMemsetAsync(ctx->wsum_pdf_class.p, 0, ctx->K * sizeof(double), s));
    HIPCHK(hipMemsetAsync(ctx->wsum_pdf_dir.p, 0, (size_t)ctx->K * ctx->D * sizeof(double), s));
    HIPCHK(hipMemsetAsync(ctx->wsum_sumw_group.p, 0, ctx->cfg.nr_groups * sizeof(double), s));
    HIPCHK(hipMemsetAsync(ctx->wsum_prior_x.p, 0, ctx->K * sizeof(double), s));
    HIPCHK(hipMemsetAsync(ctx->wsum_prior_y.p, 0, ctx->K * sizeof(double), s));
    HIPCHK(hipMemsetAsync(ctx->wsum_scalars.p, 0, 4 * sizeof(double), s));
    HIPCHK(hipStreamSynchronize(s));
    return ROME_OK;
}

int rome_expectation(rome_ctx* ctx, const rome_images* im, rome_results* res) {
    if (!ctx || !im) return ROME_ERR_ARG;
    if (!ctx->configured) return fail(ctx, ROME_ERR_ARG, "rome_configure has not been called");
    const int n = im->n;
    if (n <= 0) return fail(ctx, ROME_ERR_ARG, "empty image pool");
    if (!im->fmask_re || !im->fmask_im || !im->fnomask_re || !im->fnomask_im || !im->fctf || !im->highres_Xi2 ||
        !im->power_img || !im->old_offx || !im->old_offy || !im->norm || !im->group)
        return fail(ctx, ROME_ERR_ARG, "null image array");
    HIPCHK(hipSetDevice(ctx->device));
    for (int i = 0; i < n; ++i)
        if (im->group[i] < 0 || im->group[i] >= ctx->cfg.nr_groups) return fail(ctx, ROME_ERR_ARG, "group out of range");
    const size_t F2 = ctx->F2f;
    std::vector<float2> fm((size_t)n * F2), fn((size_t)n * F2);
    std::vector<float> fc((size_t)n * F2);
    for (size_t q = 0; q < (size_t)n * F2; ++q) {
        if (!std::isfinite(im->fmask_re[q]) || !std::isfinite(im->fmask_im[q]))
            return fail(ctx, ROME_ERR_NUMERIC, "non-finite value in image %zu", q / F2);
        fm[q] = make_float2((float)im->fmask_re[q], (float)im->fmask_im[q]);
        fn[q] = make_float2((float)im->fnomask_re[q], (float)im->fnomask_im[q]);
        fc[q] = (float)im->fctf[q];
    }
    hipStream_t s = ctx->stream;
    HIPCHK(ctx->in_fmask.upload(fm.data(), fm.size(), s));
    HIPCHK(ctx->in_fnomask.upload(fn.data(), fn.size(), s));
    HIPCHK(ctx->in_fctf.upload(fc.data(), fc.size(), s));
    HIPCHK(ctx->in_hxi2.upload(im->highres_Xi2, n, s));
    HIPCHK(ctx->in_power.upload(im->power_img, (size_t)n * ctx->oriF, s));
    HIPCHK(ctx->in_offx.upload(im->old_offx, n, s));
    HIPCHK(ctx->in_offy.upload(im->old_offy, n, s));
    HIPCHK(ctx->in_norm.upload(im->norm, n, s));
    HIPCHK(ctx->in_group.upload(im->group, n, s));
    ctx->in = {ctx->in_fmask.p, ctx->in_fnomask.p, ctx->in_fctf.p, ctx->in_hxi2.p, ctx->in_power.p,
               ctx->in_offx.p, ctx->in_offy.p, ctx->in_norm.p, ctx->in_group.p};
    return run_estep(ctx, n, res);
}

int rome_expectation_device(rome_ctx* ctx, const rome_device_images* im, rome_results* res) {
    if (!ctx || !im) return ROME_ERR_ARG;
    if (!ctx->configured) return fail(ctx, ROME_ERR_ARG, "rome_configure has not been called");
    const int n = im->n;
    if (n <= 0) return fail(ctx, ROME_ERR_ARG, "empty image pool");
    HIPCHK(hipSetDevice(ctx->device));
    // read in place: no copy of the caller's arrays (each aligned to its element size)
    const void* ptrs[] = {im->fmask, im->fnomask, im->fctf, im->highres_Xi2, im->power_img,
                          im->old_offx, im->old_offy, im->norm, im->group};
    const size_t align[] = {8, 8, 4, 8, 8, 8, 8, 8, 4};
    for (int i = 0; i < 9; ++i)
        if (!ptrs[i] || (reinterpret_cast<uintptr_t>(ptrs[i]) % align[i]) != 0)
            return fail(ctx, ROME_ERR_ARG, "device image array %d is null or misaligned", i);
    ctx->in = {reinterpret_cast<const float2*>(im->fmask), reinterpret_cast<const float2*>(im->fnomask), im->fctf,
               im->highres_Xi2, im->power_img, im->old_offx, im->old_offy, im->norm, im->group};
    return run_estep(ctx, n, res);
}

int rome_expectation_fine_device(rome_ctx* ctx, const rome_device_images* im, const int* sig_keys,
                                 const int* sig_cnt, int cap, rome_results* res) {
    if (!ctx || !im || !sig_keys || !sig_cnt) return ROME_ERR_ARG;
    if (!ctx->configured) return fail(ctx, ROME_ERR_ARG, "rome_configure has not been called");
    if (!ctx->two_pass) return fail(ctx, ROME_ERR_ARG, "given coarse significance needs adaptive_oversampling 1");
    if (cap < 1 || cap > kSigCap) return fail(ctx, ROME_ERR_ARG, "key cap %d outside 1 .. %d", cap, kSigCap);
    if ((long long)ctx->Mtot * ctx->T > 0x7fffffffLL)
        return fail(ctx, ROME_ERR_UNSUPPORTED, "coarse samples per image exceed the 31-bit keys");
    if ((reinterpret_cast<uintptr_t>(sig_keys) | reinterpret_cast<uintptr_t>(sig_cnt)) % 4)
        return fail(ctx, ROME_ERR_ARG, "key arrays misaligned");
    // the pool runs with the caller's cap as its key stride (capacities of the device-driven path too)
    const int saved_cap = ctx->sig_cap, saved_left = ctx->sig_dense_left;
    ctx->sig_cap = cap;
    ctx->sig_dense_left = 0;
    ctx->given.keys = sig_keys;
    ctx->given.cnt = sig_cnt;
    ctx->given.cap = cap;
    const int rc = rome_expectation_device(ctx, im, res);
    ctx->given.keys = nullptr;
    ctx->given.cnt = nullptr;
    ctx->given.cap = 0;
    ctx->sig_cap = saved_cap;
    ctx->sig_dense_left = saved_left;
    return rc;
}

int rome_get_sums(rome_ctx* ctx, rome_sums* o) {
    if (!ctx || !o) return ROME_ERR_ARG;
    if (!ctx->configured) return fail(ctx, ROME_ERR_ARG, "rome_configure has not been called");
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    auto cp = [&](double* h, const double* d, size_t n) -> hipError_t {
        return h ? hipMemcpyAsync(h, d, n * sizeof(double), hipMemcpyDeviceToHost, s) : hipSuccess;
    };
    const int G = ctx->cfg.nr_groups;
    HIPCHK(cp(o->wsum_sigma2_noise, ctx->wsum_sigma2.p, (size_t)G * ctx->oriF));
    HIPCHK(cp(o->wsum_sumw_group, ctx->wsum_sumw_group.p, G));
    HIPCHK(cp(o->wsum_pdf_class, ctx->wsum_pdf_class.p, ctx->K));
    HIPCHK(cp(o->wsum_pdf_direction, ctx->wsum_pdf_dir.p, (size_t)ctx->K * ctx->D));
    HIPCHK(cp(o->wsum_prior_offx, ctx->wsum_prior_x.p, ctx->K));
    HIPCHK(cp(o->wsum_prior_offy, ctx->wsum_prior_y.p, ctx->K));
    HIPCHK(cp(o->wsum_scalars, ctx->wsum_scalars.p, 4));
    HIPCHK(hipStreamSynchronize(s));
    return ROME_OK;
}

int rome_get_backprojector(rome_ctx* ctx, int iclass, double* data, double* weight) {
    if (!ctx) return ROME_ERR_ARG;
    if (!ctx->configured) return fail(ctx, ROME_ERR_ARG, "rome_configure has not been called");
    if (iclass < 0 || iclass >= ctx->K) return fail(ctx, ROME_ERR_ARG, "class index out of range");
    HIPCHK(hipSetDevice(ctx->device));
    std::vector<float> d(2 * ctx->grid_n), w(ctx->grid_n);
    HIPCHK(hipMemcpyAsync(d.data(), ctx->bp_data.p + 2 * ctx->grid_n * iclass, d.size() * sizeof(float),
                          hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(w.data(), ctx->bp_weight.p + ctx->grid_n * iclass, w.size() * sizeof(float),
                          hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (data)
        for (size_t q = 0; q < d.size(); ++q) data[q] = d[q];
    if (weight)
        for (size_t q = 0; q < w.size(); ++q) weight[q] = w[q];
    return ROME_OK;
}

int rome_get_coarse_diff2(rome_ctx* ctx, double* out) {
    if (!ctx || !out) return ROME_ERR_ARG;
    const int n = ctx->last_n;
    if (n == 0) return fail(ctx, ROME_ERR_ARG, "no E-step has run");
    if (ctx->last_coarse_full == 0)
        return fail(ctx, ROME_ERR_ARG, "the last pool kept a coarse list: set ROME_COARSE_FULL first");
    HIPCHK(hipSetDevice(ctx->device));
    const size_t per = (size_t)ctx->Mtot * ctx->T;
    std::vector<float> e((size_t)n * per);
    std::vector<double> c2(n);
    HIPCHK(hipMemcpy(e.data(), ctx->E.p, e.size() * sizeof(float), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(c2.data(), ctx->C2c.p, n * sizeof(double), hipMemcpyDeviceToHost));
    std::vector<unsigned char> valid(ctx->Mtot);
    HIPCHK(hipMemcpy(valid.data(), ctx->valid.p, ctx->Mtot, hipMemcpyDeviceToHost));
    // device layout [img][t][m] -> [img][m][t]
    const int M = ctx->Mtot, T = ctx->T;
    for (int i = 0; i < n; ++i)
        for (int m = 0; m < M; ++m)
            for (int t = 0; t < T; ++t)
                out[(size_t)i * per + (size_t)m * T + t] =
                    valid[m] ? (double)e[(size_t)i * per + (size_t)t * M + m] + c2[i]
                             : std::numeric_limits<double>::quiet_NaN();
    return ROME_OK;
}

long long rome_get_fine_count(rome_ctx* ctx) {
    if (!ctx) return -1;
    return ctx->last_blocks * (long long)ctx->NORf * ctx->NOTf;
}

int rome_get_fine_list(rome_ctx* ctx, long long* keys, double* diff2) {
    if (!ctx) return ROME_ERR_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    const long long nb = ctx->last_blocks, nt = ctx->last_tasks;
    const int NE = ctx->NORf * ctx->NOTf;
    std::vector<int> bt(nb), btask(nb), tm(nt), ti(nt);
    std::vector<double> d((size_t)nb * NE);
    if (nb > 0) {
        HIPCHK(hipMemcpy(bt.data(), ctx->blk_t.p, nb * sizeof(int), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(btask.data(), ctx->blk_task.p, nb * sizeof(int), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(tm.data(), ctx->task_m.p, nt * sizeof(int), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(ti.data(), ctx->task_img.p, nt * sizeof(int), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(d.data(), ctx->D2.p, d.size() * sizeof(double), hipMemcpyDeviceToHost));
    }
    for (long long b = 0; b < nb; ++b) {
        const int q = btask[b];
        const long long base = (((long long)ti[q] * ctx->Mtot + tm[q]) * ctx->T + bt[b]) * NE;
        for (int e = 0; e < NE; ++e) {
            if (keys) keys[b * NE + e] = base + e;
            if (diff2) diff2[b * NE + e] = d[(size_t)b * NE + e];
        }
    }
    return ROME_OK;
}

int rome_project(rome_ctx* ctx, int iclass, int size, int nrot, const double* A9, double* out_re, double* out_im) {
    if (!ctx || !A9 || nrot < 0) return ROME_ERR_ARG;
    if (!ctx->configured) return fail(ctx, ROME_ERR_ARG, "rome_configure has not been called");
    if (iclass < 0 || iclass >= ctx->K || !ctx->have_ref[iclass]) return fail(ctx, ROME_ERR_ARG, "class not set");
    if (size <= 0 || (size & 1) || size / 2 > ctx->r_max) return fail(ctx, ROME_ERR_ARG, "bad projection size");
    HIPCHK(hipSetDevice(ctx->device));
    std::vector<double> ai(6 * (size_t)std::max(nrot, 1));
    for (int r = 0; r < nrot; ++r) rot6(A9 + 9 * (size_t)r, &ai[6 * r]);
    DevBuf<double> dai;
    DevBuf<float2> dout;
    const size_t F2 = (size_t)size * (size / 2 + 1);
    HIPCHK(dai.upload(ai.data(), ai.size(), ctx->stream));
    HIPCHK(dout.reserve(F2 * std::max(nrot, 1)));
    HIPCHK(launch_project_plane(ctx->stream, ctx->grids.p + ctx->grid_n * iclass, ctx->pad, ctx->padF, ctx->cgrid,
                                ctx->is3d, dai.p, nrot, size, std::min(ctx->r_max, size / 2), dout.p));
    std::vector<float2> h(F2 * nrot);
    HIPCHK(hipMemcpyAsync(h.data(), dout.p, h.size() * sizeof(float2), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    for (size_t q = 0; q < h.size(); ++q) {
        if (out_re) out_re[q] = h[q].x;
        if (out_im) out_im[q] = h[q].y;
    }
    return ROME_OK;
}

int rome_set_profiling(rome_ctx* ctx, int on) {
    if (!ctx) return ROME_ERR_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    if (on && !ctx->ev[0][0][0])
        for (auto& st : ctx->ev)
            for (auto& pr : st)
                for (hipEvent_t& e : pr) HIPCHK(hipEventCreate(&e));
    ctx->profiling = on != 0;
    for (int i = 0; i < rome_ctx::kStages; ++i) {
        ctx->stage_ms[i] = 0.;
        ctx->stage_calls[i] = 0;
        ctx->ev_used[i] = 0;
    }
    return ROME_OK;
}

int rome_autotune(rome_ctx* ctx, const rome_device_images* im, const int* sig_keys, const int* sig_cnt, int cap,
                  int* chosen, double* stage_ms) {
    if (!ctx || !im) return ROME_ERR_ARG;
    if (!ctx->configured) return fail(ctx, ROME_ERR_ARG, "rome_configure has not been called");
    HIPCHK(hipSetDevice(ctx->device));
    // per tuned kernel: its stage and the candidate grids in blocks per CU
    static const int stage_of[4] = {6, 5, 8, 10};
    static const int per_cu[4][kTuneCands] = {{1, 2, 3, 4}, {4, 8, 12, 16}, {1, 2, 3, 4}, {8, 16, 32, 64}};
    const bool prof = ctx->profiling;
    const int mode = ctx->exec_mode;
    int saved[4];
    std::copy(ctx->persist, ctx->persist + 4, saved);
    int rc = rome_set_profiling(ctx, 1);
    if (rc) return rc;
    ctx->exec_mode = ROME_EXEC_DEVICE;
    double t[4][kTuneCands];
    for (int c = 0; c < kTuneCands && !rc; ++c) {
        for (int i = 0; i < 4; ++i) ctx->persist[i] = ctx->n_cu * per_cu[i][c];
        for (int rep = 0; rep < 2 && !rc; ++rep) {  // the first run of a candidate is not timed
            for (int i = 0; i < rome_ctx::kStages; ++i) ctx->stage_ms[i] = 0.;
            rc = sig_keys ? rome_expectation_fine_device(ctx, im, sig_keys, sig_cnt, cap, nullptr)
                          : rome_expectation_device(ctx, im, nullptr);
        }
        if (!rc && ctx->last_exec != 1) rc = fail(ctx, ROME_ERR_UNSUPPORTED, "the pool does not run device-driven");
        for (int i = 0; i < 4; ++i) t[i][c] = ctx->stage_ms[stage_of[i]];
    }
    ctx->exec_mode = mode;
    (void)rome_set_profiling(ctx, prof ? 1 : 0);
    ++ctx->state_epoch;  // kernels of captured pools read what this call changes
    if (rc) {
        std::copy(saved, saved + 4, ctx->persist);
        return rc;
    }
    for (int i = 0; i < 4; ++i) {
        int best = 0;
        for (int c = 1; c < kTuneCands; ++c)
            if (t[i][c] < t[i][best]) best = c;
        ctx->persist[i] = ctx->n_cu * per_cu[i][best];
        if (chosen) chosen[i] = ctx->persist[i];
        if (stage_ms)
            for (int c = 0; c < kTuneCands; ++c) stage_ms[i * kTuneCands + c] = t[i][c];
    }
    return ROME_OK;
}

int rome_set_persistent_grids(rome_ctx* ctx, const int* blocks) {
    if (!ctx || !blocks) return ROME_ERR_ARG;
    for (int i = 0; i < 4; ++i)
        if (blocks[i] < 0) return fail(ctx, ROME_ERR_ARG, "negative grid size");
    std::copy(blocks, blocks + 4, ctx->persist);
    ctx->hist_tasks = ctx->hist_U = 0.;  // the next pool's first launches are sized from its pool size
    ++ctx->state_epoch;  // kernels of captured pools read what this call changes
    return ROME_OK;
}

int rome_get_profile(rome_ctx* ctx, double* stage_ms, long long* stage_calls, int nstages) {
    if (!ctx || nstages < 0) return ROME_ERR_ARG;
    for (int i = 0; i < nstages && i < rome_ctx::kStages; ++i) {
        if (stage_ms) stage_ms[i] = ctx->stage_ms[i];
        if (stage_calls) stage_calls[i] = ctx->stage_calls[i];
    }
    return rome_ctx::kStages;
}

int rome_set_coarse_mode(rome_ctx* ctx, int mode, int list_cap) {
    if (ctx) ++ctx->state_epoch;  // kernels of captured pools read what this call changes
    if (!ctx) return ROME_ERR_ARG;
    if (mode != ROME_COARSE_LIST && mode != ROME_COARSE_FULL) return fail(ctx, ROME_ERR_ARG, "unknown coarse mode %d", mode);
    ctx->coarse_mode = mode;
    ctx->list_cap = list_cap > 0 ? list_cap : 0;
    return ROME_OK;
}

int rome_set_significance_cap(rome_ctx* ctx, int cap) {
    if (ctx) ++ctx->state_epoch;  // kernels of captured pools read what this call changes
    if (!ctx) return ROME_ERR_ARG;
    if (cap > kSigCap) return fail(ctx, ROME_ERR_ARG, "significance cap %d above %d", cap, kSigCap);
    ctx->sig_cap = cap < 0 ? kSigCapDefault : cap;
    ctx->sig_dense_left = 0;
    return ROME_OK;
}

int rome_get_significance_stats(rome_ctx* ctx, long long* out) {
    if (!ctx || !out) return ROME_ERR_ARG;
    out[0] = ctx->last_sig_form;
    out[1] = ctx->last_sig_max;
    return ROME_OK;
}

int rome_get_coarse_stats(rome_ctx* ctx, long long* out) {
    if (!ctx || !out) return ROME_ERR_ARG;
    out[0] = ctx->last_coarse_full;
    out[1] = ctx->last_list_max;
    out[2] = ctx->last_list_total;
    return ROME_OK;
}

int rome_get_batch_stats(rome_ctx* ctx, long long* out) {
    if (!ctx || !out) return ROME_ERR_ARG;
    out[0] = ctx->last_n;
    out[1] = ctx->last_tasks;
    out[2] = ctx->last_blocks;
    out[3] = ctx->last_U;
    out[4] = ctx->Pc;
    out[5] = ctx->Pf;
    out[6] = ctx->Mtot;
    out[7] = ctx->T;
    return ROME_OK;
}

int rome_project_device(rome_ctx* ctx, int iclass, int size, int nrot, const double* A9, void* out) {
    if (!ctx || !A9 || !out || nrot < 0) return ROME_ERR_ARG;
    if (!ctx->configured) return fail(ctx, ROME_ERR_ARG, "rome_configure has not been called");
    if (iclass < 0 || iclass >= ctx->K || !ctx->have_ref[iclass]) return fail(ctx, ROME_ERR_ARG, "class not set");
    if (size <= 0 || (size & 1) || size / 2 > ctx->r_max) return fail(ctx, ROME_ERR_ARG, "bad projection size");
    if (nrot == 0) return ROME_OK;
    HIPCHK(hipSetDevice(ctx->device));
    std::vector<double> ai(6 * (size_t)nrot);
    for (int r = 0; r < nrot; ++r) rot6(A9 + 9 * (size_t)r, &ai[6 * r]);
    DevBuf<double> dai;
    HIPCHK(dai.upload(ai.data(), ai.size(), ctx->stream));
    HIPCHK(launch_project_plane(ctx->stream, ctx->grids.p + ctx->grid_n * iclass, ctx->pad, ctx->padF, ctx->cgrid,
                                ctx->is3d, dai.p, nrot, size, std::min(ctx->r_max, size / 2), (float2*)out));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return ROME_OK;
}

int rome_project_device_noalloc(rome_ctx* ctx, int iclass, int size, int nrot, const double* A9, void* out) {
    if (!ctx || !A9 || !out || nrot < 0) return ROME_ERR_ARG;
    if (!ctx->configured) return fail(ctx, ROME_ERR_ARG, "rome_configure has not been called");
    if (iclass < 0 || iclass >= ctx->K || !ctx->have_ref[iclass]) return fail(ctx, ROME_ERR_ARG, "class not set");
    if (size <= 0 || (size & 1) || size / 2 > ctx->r_max) return fail(ctx, ROME_ERR_ARG, "bad projection size");
    if (nrot > kNoallocMaxRot)
        return fail(ctx, ROME_ERR_ARG, "%d rotations exceed the %d the allocation-free projection stages", nrot,
                    kNoallocMaxRot);
    if (nrot == 0) return ROME_OK;
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(project_plane_noalloc(ctx->stream, ctx->device, A9, nrot, ctx->grids.p + ctx->grid_n * iclass, ctx->pad,
                                 ctx->padF, ctx->cgrid, ctx->is3d, size, std::min(ctx->r_max, size / 2), (float2*)out));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return ROME_OK;
}

int rome_sampling_pose3d(int order, int oversampling, double perturbation, double* out) {
    return out ? pose_table3d(order, oversampling, perturbation, out) : ROME_ERR_ARG;
}
int rome_sampling_pose2d(double psi_step, int oversampling, double perturbation, double* out) {
    return out ? pose_table2d(psi_step, oversampling, perturbation, out) : ROME_ERR_ARG;
}
int rome_sampling_translations(double range, double step, double* tx, double* ty) {
    return translations(range, step, tx, ty);
}
int rome_sampling_shifts(double range, double step, int oversampling, double perturbation, double* ox, double* oy) {
    return (ox && oy) ? shift_table(range, step, oversampling, perturbation, ox, oy) : ROME_ERR_ARG;
}
int rome_healpix_pix2ang(int order, int pix, double* z, double* phi) {
    if (!z || !phi || order < 0 || order > 13 || pix < 0 || pix >= (12 << (2 * order))) return ROME_ERR_ARG;
    healpix_pix2ang(order, pix, *z, *phi);
    return ROME_OK;
}

}  // extern "C"

int rome_update_model(rome_ctx* ctx, int install, rome_model* out) {
    if (ctx) ++ctx->state_epoch;  // kernels of captured pools read what this call changes
    if (!ctx) return ROME_ERR_ARG;
    if (!ctx->configured || !ctx->have_model) return fail(ctx, ROME_ERR_ARG, "no configured model to update");
    const rome_config& c = ctx->cfg;
    const int K = ctx->K, D = ctx->D, G = c.nr_groups, F = ctx->oriF;
    std::vector<double> ws((size_t)G * F), wg(G), wpc(K), wpd((size_t)K * D), wpx(K), wpy(K), wsc(4);
    rome_sums s{ws.data(), wg.data(), wpc.data(), wpd.data(), wpx.data(), wpy.data(), wsc.data()};
    int rc = rome_get_sums(ctx, &s);
    if (rc) return rc;
    std::vector<double> sig((size_t)G * F), pc(K), pd((size_t)K * D), px(K), py(K), sc(4);
    rome_model m{sig.data(), pc.data(), pd.data(), px.data(), py.data(), sc.data()};
    rc = rome_maximization_other(c.ref_dim, c.ori_size, K, G, D, c.do_norm_correction, c.avg_norm_correction,
                                 ctx->prior_x_h.data(), ctx->prior_y_h.data(), &s, &m);
    if (rc == ROME_ERR_NUMERIC) return fail(ctx, rc, "sum of the class weights is not positive");
    if (rc) return fail(ctx, rc, "rome_maximization_other failed");
    if (out) {
        auto put = [](double* dst, const std::vector<double>& v) {
            if (dst) std::copy(v.begin(), v.end(), dst);
        };
        put(out->sigma2_noise, sig);
        put(out->pdf_class, pc);
        put(out->pdf_direction, pd);
        put(out->prior_offx, px);
        put(out->prior_offy, py);
        put(out->scalars, sc);
    }
    if (install) {
        rc = rome_set_model(ctx, sig.data(), pc.data(), pd.data(), px.data(), py.data());
        if (rc) return rc;
        ctx->cfg.sigma2_offset = sc[0];
        ctx->cfg.avg_norm_correction = sc[1];
        if (c.do_ctf_correction) ctx->cfg.refs_are_ctf_corrected = 1;  // after the first iteration
    }
    return ROME_OK;
}

// cached hipFFT plan (type D2Z / Z2D, rank 2 or 3, n per axis) bound to the context's stream
static int fft_plan(rome_ctx* ctx, hipfftType type, int rank, int n, hipfftHandle* out, int batch = 1) {
    for (auto& p : ctx->fft_plans)
        if (p.type == (int)type && p.rank == rank && p.n == n && p.batch == batch) {
            *out = p.h;
            return ROME_OK;
        }
    hipfftHandle h = 0;
    int dims[3] = {n, n, n};
    const hipfftResult r = batch > 1 ? hipfftPlanMany(&h, rank, dims, nullptr, 1, 0, nullptr, 1, 0, type, batch)
                           : rank == 3 ? hipfftPlan3d(&h, n, n, n, type)
                                       : hipfftPlan2d(&h, n, n, type);
    if (r != HIPFFT_SUCCESS) return fail(ctx, ROME_ERR_HIP, "hipfftPlan (type %d, %d^%d) failed: %d", (int)type, n, rank, (int)r);
    if (hipfftSetStream(h, ctx->stream) != HIPFFT_SUCCESS) {
        (void)hipfftDestroy(h);
        return fail(ctx, ROME_ERR_HIP, "hipfftSetStream failed");
    }
    ctx->fft_plans.push_back({(int)type, rank, n, batch, h});
    *out = h;
    return ROME_OK;
}

int rome_fourier_map(rome_ctx* ctx, int iclass, const double* vol, int do_gridding, double* power_spectrum,
                     double* corrected_map, double* grid_out) {
    if (ctx) ++ctx->state_epoch;  // kernels of captured pools read what this call changes
    if (!ctx || !vol) return ROME_ERR_ARG;
    if (!ctx->configured) return fail(ctx, ROME_ERR_ARG, "rome_configure has not been called");
    if (iclass < 0 || iclass >= ctx->K) return fail(ctx, ROME_ERR_ARG, "class index out of range");
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    const int pf = 2, ori = ctx->cfg.ori_size, padori = pf * ori, dim3 = ctx->is3d ? 1 : 0;
    const size_t nvol = dim3 ? (size_t)ori * ori * ori : (size_t)ori * ori;
    const size_t npad = dim3 ? (size_t)padori * padori * padori : (size_t)padori * padori;
    const size_t nF = (dim3 ? (size_t)padori : 1) * padori * (padori / 2 + 1);
    const int oriF = ctx->oriF;
    hipfftHandle plan;
    int rc = fft_plan(ctx, HIPFFT_D2Z, dim3 ? 3 : 2, padori, &plan);
    if (rc) return rc;
    HIPCHK(ctx->fm_vol.upload(vol, nvol, s));
    HIPCHK(ctx->fm_Mc.reserve(npad));
    HIPCHK(hipMemsetAsync(ctx->fm_Mc.p, 0, npad * sizeof(double), s));
    if (corrected_map) HIPCHK(ctx->fm_corr.reserve(nvol));
    HIPCHK(launch_fm_pad(s, ctx->fm_vol.p, ori, padori, dim3, do_gridding ? 1 : 0, pf, ctx->fm_Mc.p,
                         corrected_map ? ctx->fm_corr.p : nullptr));
    HIPCHK(ctx->fm_F.reserve(nF));
    const hipfftResult fr = hipfftExecD2Z(plan, ctx->fm_Mc.p, reinterpret_cast<hipfftDoubleComplex*>(ctx->fm_F.p));
    if (fr != HIPFFT_SUCCESS) return fail(ctx, ROME_ERR_HIP, "hipfftExecD2Z failed: %d", (int)fr);
    HIPCHK(ctx->fm_grid64.reserve(ctx->grid_n));
    HIPCHK(hipMemsetAsync(ctx->fm_grid64.p, 0, ctx->grid_n * sizeof(double2), s));
    float2* g32 = ctx->grids.p + ctx->grid_n * iclass;
    HIPCHK(hipMemsetAsync(g32, 0, ctx->grid_n * sizeof(float2), s));
    HIPCHK(ctx->fm_pw.reserve(oriF));
    HIPCHK(ctx->fm_cnt.reserve(oriF));
    HIPCHK(hipMemsetAsync(ctx->fm_pw.p, 0, oriF * sizeof(double), s));
    HIPCHK(hipMemsetAsync(ctx->fm_cnt.p, 0, oriF * sizeof(double), s));
    const double normfft = dim3 ? (double)(pf * pf * pf * ori) : (double)(pf * pf);
    HIPCHK(launch_fm_window(s, ctx->fm_F.p, padori, dim3, pf, ctx->r_max, ctx->pad, oriF, normfft, ctx->fm_grid64.p,
                            g32, ctx->fm_pw.p, ctx->fm_cnt.p));
    std::vector<double> pw(oriF), cnt(oriF);
    HIPCHK(hipMemcpyAsync(pw.data(), ctx->fm_pw.p, oriF * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(cnt.data(), ctx->fm_cnt.p, oriF * sizeof(double), hipMemcpyDeviceToHost, s));
    if (corrected_map)
        HIPCHK(hipMemcpyAsync(corrected_map, ctx->fm_corr.p, nvol * sizeof(double), hipMemcpyDeviceToHost, s));
    if (grid_out)
        HIPCHK(hipMemcpyAsync(grid_out, ctx->fm_grid64.p, ctx->grid_n * sizeof(double2), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (power_spectrum)  // radial average (src/reconstruct.cpp:213-221)
        for (int i = 0; i < oriF; ++i) power_spectrum[i] = cnt[i] < 1. ? 0. : pw[i] / cnt[i];
    ctx->have_ref[iclass] = 1;
    ctx->fct_dirty = true;
    return ROME_OK;
}

// Fourier transform of the order-0 Kaiser-Bessel blob, tabulated as TabFtBlob::init does
// (src/reconstruct.h:615-627, kaiser_Fourier_value src/bessel.cpp:26-50): 10000 samples over [0, 1/2).
// Only ratios to the value at 0 are used, so the constant factor (2 pi)^{3/2} a^3 / I0(alpha) only has
// to be the same for every entry; I0 by its power series.
static std::vector<double> kaiser_ft_table(double a, double alpha, int nr_elem, double* sampling) {
    *sampling = 0.5 / (double)nr_elem;
    double i0 = 1., term = 1.;
    for (int k = 1; term > 1e-17 * i0; ++k) {
        term *= (alpha / 2.) * (alpha / 2.) / ((double)k * k);
        i0 += term;
    }
    const double pi = 3.14159265358979323846;
    const double cst = std::pow(2. * pi, 1.5) * a * a * a;
    std::vector<double> t(nr_elem);
    for (int i = 0; i < nr_elem; ++i) {
        const double w = (double)i * *sampling;
        const double x = 2. * pi * a * w;
        const double sig = std::sqrt(std::fabs(alpha * alpha - x * x));
        double b;
        if (x > alpha)  // spherical-order J_{3/2}
            b = std::sqrt(2. / (pi * sig)) * (std::sin(sig) / sig - std::cos(sig));
        else            // I_{3/2}
            b = sig == 0. ? 0. : std::sqrt(2. / (pi * sig)) * (std::cosh(sig) - std::sinh(sig) / sig);
        t[i] = cst * b / (i0 * std::pow(sig, 1.5));
    }
    return t;
}

int rome_reconstruct(rome_ctx* ctx, int iclass, const rome_recon_params* prm, const double* data_in,
                     const double* weight_in, double* tau2, const double* fsc, double* sigma2,
                     double* data_vs_prior, double* vol_out) {
    if (!ctx || !prm || !tau2 || !sigma2 || !data_vs_prior || !vol_out) return ROME_ERR_ARG;
    if (!ctx->configured) return fail(ctx, ROME_ERR_ARG, "rome_configure has not been called");
    if (iclass < 0 || iclass >= ctx->K) return fail(ctx, ROME_ERR_ARG, "class index out of range");
    if ((data_in == nullptr) != (weight_in == nullptr)) return fail(ctx, ROME_ERR_ARG, "give both grids or neither");
    if (prm->update_tau2_with_fsc && !fsc) return fail(ctx, ROME_ERR_ARG, "update_tau2_with_fsc needs fsc");
    if (prm->max_iter_preweight < 0 || !(prm->normalise != 0.)) return fail(ctx, ROME_ERR_ARG, "bad parameters");
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    const int pf = 2, ori = ctx->cfg.ori_size, oriF = ctx->oriF, r_max = ctx->r_max;
    const int P = ctx->pad, PF = ctx->padF, Dz = ctx->is3d ? P : 1, rank = ctx->is3d ? 3 : 2;
    const int N = pf * ori, NF = N / 2 + 1, NZ = ctx->is3d ? N : 1, OZ = ctx->is3d ? ori : 1;
    const int max_r2 = r_max * r_max * pf * pf;
    const size_t ng = (size_t)Dz * P * PF, nM = (size_t)Dz * P * P, nFt = (size_t)NZ * N * NF;
    const size_t nMo = (size_t)NZ * N * N, nv = (size_t)OZ * ori * ori;
    const double osc = ctx->is3d ? (double)(pf * pf * pf) : (double)(pf * pf);
    hipfftHandle planPi, planPf, planNi;
    int rc = fft_plan(ctx, HIPFFT_Z2D, rank, P, &planPi);
    if (!rc) rc = fft_plan(ctx, HIPFFT_D2Z, rank, P, &planPf);
    if (!rc) rc = fft_plan(ctx, HIPFFT_Z2D, rank, N, &planNi);
    if (rc) return rc;
    HIPCHK(ctx->rc_data.reserve(ng));
    HIPCHK(ctx->rc_weight.reserve(ng));
    HIPCHK(ctx->rc_Fw.reserve(ng));
    HIPCHK(ctx->rc_Fnw.reserve(ng));
    // rc_Fc and rc_M also take the transform of the output map (tau2 from the FSC)
    HIPCHK(ctx->rc_Fc.reserve(std::max(ng, (size_t)OZ * ori * (ori / 2 + 1))));
    HIPCHK(ctx->rc_M.reserve(std::max(nM, nv)));
    HIPCHK(ctx->rc_Ft.reserve(nFt));
    HIPCHK(ctx->rc_Mout.reserve(nMo));
    HIPCHK(ctx->rc_vol.reserve(nv));
    HIPCHK(ctx->rc_sums.reserve(4 * (size_t)std::max(oriF, ori)));
    HIPCHK(ctx->rc_tau2.reserve(oriF));
    HIPCHK(ctx->rc_bad.reserve(1));
    if (data_in) {
        HIPCHK(hipMemcpyAsync(ctx->rc_data.p, data_in, ng * sizeof(double2), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(ctx->rc_weight.p, weight_in, ng * sizeof(double), hipMemcpyHostToDevice, s));
    } else {
        HIPCHK(rc_from_float(s, ctx->bp_data.p ? reinterpret_cast<const float2*>(ctx->bp_data.p) + ctx->grid_n * iclass
                                               : nullptr,
                             ctx->bp_weight.p + ctx->grid_n * iclass, ctx->rc_data.p, ctx->rc_weight.p, ng));
    }
    // Hermitian fold of the x = 0 plane, then the weights in FFTW order
    HIPCHK(rc_fold(s, ctx->rc_data.p, ctx->rc_weight.p, Dz, P, PF));
    HIPCHK(rc_decenter_real(s, ctx->rc_weight.p, ctx->rc_Fw.p, Dz, P, PF, max_r2));
    // radial average of the inverse noise power (:2019-2060)
    double* s0 = ctx->rc_sums.p;
    double* s1 = s0 + oriF;
    HIPCHK(hipMemsetAsync(s0, 0, 2 * oriF * sizeof(double), s));
    HIPCHK(rc_spectrum(s, ctx->rc_Fw.p, Dz, P, PF, max_r2, pf, osc, oriF, s0, s1));
    std::vector<double> h0(oriF), h1(oriF);
    HIPCHK(hipMemcpyAsync(h0.data(), s0, oriF * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(h1.data(), s1, oriF * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    for (int i = 0; i < oriF; ++i) {
        if (h0[i] > 1e-10) sigma2[i] = h1[i] / h0[i];
        else if (h0[i] == 0.) sigma2[i] = 0.;
        else return fail(ctx, ROME_ERR_NUMERIC, "unexpectedly small, yet non-zero sigma2 value (shell %d)", i);
    }
    if (prm->update_tau2_with_fsc) {  // tau2 from the FSC (:2062-2080)
        for (int i = 0; i < oriF; ++i) {
            double myfsc = std::max(0.001, fsc[i]);
            if (prm->is_whole_instead_of_half) myfsc = std::sqrt(2. * myfsc / (myfsc + 1.));
            myfsc = std::min(0.999, myfsc);
            const double myssnr = myfsc / (1. - myfsc);
            tau2[i] = myssnr * sigma2[i];
            data_vs_prior[i] = myssnr;
        }
    }
    if (prm->do_map) {
        HIPCHK(hipMemcpyAsync(ctx->rc_tau2.p, tau2, oriF * sizeof(double), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemsetAsync(s0, 0, 2 * oriF * sizeof(double), s));
        HIPCHK(hipMemsetAsync(ctx->rc_bad.p, 0, sizeof(int), s));
        HIPCHK(rc_map(s, ctx->rc_Fw.p, Dz, P, PF, max_r2, pf, osc, prm->tau2_fudge, ctx->rc_tau2.p, prm->minres_map,
                      oriF, s0, s1, ctx->rc_bad.p));
        int bad = 0;
        HIPCHK(hipMemcpyAsync(h0.data(), s0, oriF * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(h1.data(), s1, oriF * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(&bad, ctx->rc_bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        if (bad) return fail(ctx, ROME_ERR_NUMERIC, "negative values in the tau2 spectrum");
        if (!prm->update_tau2_with_fsc)
            for (int i = 0; i < oriF; ++i)
                data_vs_prior[i] = i > r_max ? 0. : (h1[i] < 0.001 ? 999. : h0[i] / h1[i]);
    }
    HIPCHK(rc_normalise(s, ctx->rc_Fw.p, ctx->rc_data.p, ng, prm->normalise));
    // iterative gridding weights (Pipe & Menon), :2196-2274
    if (ctx->rc_tab_h.empty()) {
        double sampling = 0.;
        ctx->rc_tab_h = kaiser_ft_table(1.9 * pf, 15.0, 10000, &sampling);
        HIPCHK(ctx->rc_tab.upload(ctx->rc_tab_h.data(), ctx->rc_tab_h.size(), s));
    }
    const double sampling = 0.5 / 10000.;
    HIPCHK(rc_newweight_init(s, ctx->rc_Fnw.p, Dz, P, PF, max_r2));
    auto Z = [](double2* p) { return reinterpret_cast<hipfftDoubleComplex*>(p); };
    for (int it = 0; it < prm->max_iter_preweight; ++it) {
        HIPCHK(rc_fconv(s, ctx->rc_Fnw.p, ctx->rc_Fw.p, ctx->rc_Fc.p, ng));
        if (hipfftExecZ2D(planPi, Z(ctx->rc_Fc.p), ctx->rc_M.p) != HIPFFT_SUCCESS)
            return fail(ctx, ROME_ERR_HIP, "hipfftExecZ2D failed");
        HIPCHK(rc_blob(s, ctx->rc_M.p, Dz, P, ori, pf, ctx->rc_tab.p, 10000, sampling, ctx->rc_tab_h[0]));
        if (hipfftExecD2Z(planPf, ctx->rc_M.p, Z(ctx->rc_Fc.p)) != HIPFFT_SUCCESS)
            return fail(ctx, ROME_ERR_HIP, "hipfftExecD2Z failed");
        HIPCHK(rc_newweight(s, ctx->rc_Fc.p, ctx->rc_Fnw.p, Dz, P, PF, max_r2, (double)nM));
    }
    // weighted data -> window to 2 ori -> real space -> window to ori, soft mask, gridding correction
    HIPCHK(rc_final_fconv(s, ctx->rc_data.p, ctx->rc_Fnw.p, ctx->rc_Fc.p, Dz, P, PF, max_r2));
    HIPCHK(rc_window_ft(s, ctx->rc_Fc.p, ctx->rc_Ft.p, Dz, P, PF, NZ, N, NF));
    if (hipfftExecZ2D(planNi, Z(ctx->rc_Ft.p), ctx->rc_Mout.p) != HIPFFT_SUCCESS)
        return fail(ctx, ROME_ERR_HIP, "hipfftExecZ2D failed");
    const double normfft = ctx->is3d ? (double)(pf * pf * pf * ori) : (double)(pf * pf);
    HIPCHK(rc_center_window(s, ctx->rc_Mout.p, ctx->rc_vol.p, NZ, N, OZ, ori, normfft));
    HIPCHK(hipMemsetAsync(s0, 0, 2 * sizeof(double), s));
    HIPCHK(rc_bgsum(s, ctx->rc_vol.p, OZ, ori, (double)ori / 2., 3., s0));
    HIPCHK(rc_mask_grid(s, ctx->rc_vol.p, OZ, ori, pf, (double)ori / 2., 3., s0));
    if (prm->update_tau2_with_fsc) {  // new tau2 = power spectrum of the new map (:2473-2510)
        hipfftHandle planO;
        rc = fft_plan(ctx, HIPFFT_D2Z, rank, ori, &planO);
        if (rc) return rc;
        // the transform input must stay intact (it is the output map): transform a copy
        HIPCHK(hipMemcpyAsync(ctx->rc_M.p, ctx->rc_vol.p, nv * sizeof(double), hipMemcpyDeviceToDevice, s));
        if (hipfftExecD2Z(planO, ctx->rc_M.p, Z(ctx->rc_Fc.p)) != HIPFFT_SUCCESS)
            return fail(ctx, ROME_ERR_HIP, "hipfftExecD2Z failed");
        HIPCHK(hipMemsetAsync(s0, 0, 2 * ori * sizeof(double), s));
        HIPCHK(rc_map_spectrum(s, ctx->rc_Fc.p, OZ, ori, (double)nv, s0, s0 + ori));
        std::vector<double> sp(2 * ori);
        HIPCHK(hipMemcpyAsync(sp.data(), s0, 2 * ori * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        const double nf = ctx->is3d ? (double)ori * ori : 1.;
        for (int i = 0; i < oriF; ++i) tau2[i] = prm->tau2_fudge * (sp[i] / sp[ori + i] * nf / 2.);
    }
    HIPCHK(hipMemcpyAsync(vol_out, ctx->rc_vol.p, nv * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return ROME_OK;
}

int rome_prepare_images(rome_ctx* ctx, const rome_raw_images* in, const rome_prep_params* p, rome_prepared* ho,
                        rome_prepared_device* dout) {
    if (!ctx || !in || !p || in->n < 0 || (in->n > 0 && (!in->images || !in->xoff || !in->yoff)))
        return ROME_ERR_ARG;
    if (!ctx->configured) return fail(ctx, ROME_ERR_ARG, "rome_configure has not been called");
    const rome_config& c = ctx->cfg;
    if (c.ref_dim != 2 && c.ref_dim != 3) return fail(ctx, ROME_ERR_ARG, "bad ref_dim");
    if (c.do_norm_correction && !in->norm) return fail(ctx, ROME_ERR_ARG, "do_norm_correction needs NORM");
    if (c.do_ctf_correction && !in->ctf) return fail(ctx, ROME_ERR_ARG, "do_ctf_correction needs CTF parameters");
    if (!(p->pixel_size > 0.) || !(p->width_mask_edge > 0.)) return fail(ctx, ROME_ERR_ARG, "bad mask parameters");
    if (in->n == 0) return ROME_OK;
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    const int ori = c.ori_size, cur = c.current_size, oriF = ori / 2 + 1, curF = cur / 2 + 1;
    const size_t n2 = (size_t)ori * ori, nF = (size_t)ori * oriF, cF = (size_t)cur * curF;
    const int chunk = 1024;  // images per device pass
    for (int i0 = 0; i0 < in->n; i0 += chunk) {
        const int n = std::min(chunk, in->n - i0);
        HIPCHK(ctx->ip_img.upload(in->images + (size_t)i0 * n2, (size_t)n * n2, s));
        // scalars: norm, xoff, yoff, old_offx, old_offy, hxi2 (6 x n)
        HIPCHK(ctx->ip_scal.reserve((size_t)6 * chunk));
        double* sc = ctx->ip_scal.p;
        if (in->norm) HIPCHK(hipMemcpyAsync(sc, in->norm + i0, n * sizeof(double), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(sc + chunk, in->xoff + i0, n * sizeof(double), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(sc + 2 * chunk, in->yoff + i0, n * sizeof(double), hipMemcpyHostToDevice, s));
        if (in->noise) HIPCHK(ctx->ip_noise.upload(in->noise + (size_t)i0 * n2, (size_t)n * n2, s));
        if (c.do_ctf_correction) HIPCHK(ctx->ip_ctf.upload(in->ctf + (size_t)7 * i0, (size_t)7 * n, s));
        HIPCHK(ctx->ip_timg.reserve((size_t)n * n2));
        HIPCHK(ctx->ip_cbuf.reserve((size_t)n * n2));
        HIPCHK(ctx->ip_F1.reserve((size_t)n * nF));
        HIPCHK(ctx->ip_F2.reserve((size_t)n * nF));
        HIPCHK(ctx->ip_fn.reserve((size_t)n * cF));
        HIPCHK(ctx->ip_fm.reserve((size_t)n * cF));
        HIPCHK(ctx->ip_fctf.reserve((size_t)n * cF));
        HIPCHK(ctx->ip_power.reserve((size_t)n * oriF));
        ImagePrepArgs a{};
        a.ori = ori;
        a.cur = cur;
        a.images = ctx->ip_img.p;
        a.norm = sc;
        a.avg_norm = c.avg_norm_correction;
        a.do_norm = c.do_norm_correction ? 1 : 0;
        a.xoff = sc + chunk;
        a.yoff = sc + 2 * chunk;
        a.old_offx = sc + 3 * chunk;
        a.old_offy = sc + 4 * chunk;
        a.timg = ctx->ip_timg.p;
        a.cbuf = ctx->ip_cbuf.p;
        a.noise = in->noise ? ctx->ip_noise.p : nullptr;
        a.radius = p->particle_diameter / (2. * p->pixel_size);
        a.cosine_width = p->width_mask_edge;
        a.F1 = ctx->ip_F1.p;
        a.F2 = ctx->ip_F2.p;
        a.fnomask = ctx->ip_fn.p;
        a.fmask = ctx->ip_fm.p;
        a.fnomask32 = dout && dout->fnomask ? (float2*)dout->fnomask + (size_t)i0 * cF : nullptr;
        a.fmask32 = dout && dout->fmask ? (float2*)dout->fmask + (size_t)i0 * cF : nullptr;
        a.power = ctx->ip_power.p;
        a.hxi2 = sc + 5 * chunk;
        a.ctf = c.do_ctf_correction ? ctx->ip_ctf.p : nullptr;
        a.angpix = p->pixel_size;
        a.ctf_abs = p->ctf_phase_flipped ? 1 : 0;
        a.ctf_flip = p->only_flip_phases ? 1 : 0;
        a.intact_first_peak = p->intact_ctf_first_peak ? 1 : 0;
        a.fctf = ctx->ip_fctf.p;
        a.fctf32 = dout && dout->fctf ? dout->fctf + (size_t)i0 * cF : nullptr;
        hipfftHandle plan;
        int rc = fft_plan(ctx, HIPFFT_D2Z, 2, ori, &plan, n);
        if (rc) return rc;
        HIPCHK(launch_imageprep_pre(s, n, a));
        hipfftResult fr = hipfftExecD2Z(plan, a.cbuf, reinterpret_cast<hipfftDoubleComplex*>(ctx->ip_F1.p));
        if (fr != HIPFFT_SUCCESS) return fail(ctx, ROME_ERR_HIP, "hipfftExecD2Z failed: %d", (int)fr);
        HIPCHK(launch_imageprep_mask(s, n, a));
        fr = hipfftExecD2Z(plan, a.cbuf, reinterpret_cast<hipfftDoubleComplex*>(ctx->ip_F2.p));
        if (fr != HIPFFT_SUCCESS) return fail(ctx, ROME_ERR_HIP, "hipfftExecD2Z failed: %d", (int)fr);
        if (cur >= ori) {
            HIPCHK(hipMemsetAsync(a.power, 0, (size_t)n * oriF * sizeof(double), s));
            HIPCHK(hipMemsetAsync(a.hxi2, 0, (size_t)n * sizeof(double), s));
        }
        HIPCHK(launch_imageprep_post(s, n, a));
        if (dout) {
            auto d2d = [&](double* dst, const double* src, size_t cnt) -> hipError_t {
                return dst ? hipMemcpyAsync(dst + (size_t)i0 * (cnt / n), src, cnt * sizeof(double),
                                            hipMemcpyDeviceToDevice, s)
                           : hipSuccess;
            };
            HIPCHK(d2d(dout->highres_Xi2, a.hxi2, n));
            HIPCHK(d2d(dout->power_img, a.power, (size_t)n * oriF));
            HIPCHK(d2d(dout->old_offx, a.old_offx, n));
            HIPCHK(d2d(dout->old_offy, a.old_offy, n));
        }
        if (ho) {
            std::vector<double2> fn, fm;
            if (ho->fnomask_re || ho->fnomask_im) {
                fn.resize((size_t)n * cF);
                HIPCHK(hipMemcpyAsync(fn.data(), a.fnomask, fn.size() * sizeof(double2), hipMemcpyDeviceToHost, s));
            }
            if (ho->fmask_re || ho->fmask_im) {
                fm.resize((size_t)n * cF);
                HIPCHK(hipMemcpyAsync(fm.data(), a.fmask, fm.size() * sizeof(double2), hipMemcpyDeviceToHost, s));
            }
            auto d2h = [&](double* dst, const double* src, size_t cnt) -> hipError_t {
                return dst ? hipMemcpyAsync(dst + (size_t)i0 * (cnt / n), src, cnt * sizeof(double),
                                            hipMemcpyDeviceToHost, s)
                           : hipSuccess;
            };
            HIPCHK(d2h(ho->fctf, a.fctf, (size_t)n * cF));
            HIPCHK(d2h(ho->highres_Xi2, a.hxi2, n));
            HIPCHK(d2h(ho->power_img, a.power, (size_t)n * oriF));
            HIPCHK(d2h(ho->old_offx, a.old_offx, n));
            HIPCHK(d2h(ho->old_offy, a.old_offy, n));
            HIPCHK(hipStreamSynchronize(s));
            for (size_t q = 0; q < fn.size(); ++q) {
                if (ho->fnomask_re) ho->fnomask_re[(size_t)i0 * cF + q] = fn[q].x;
                if (ho->fnomask_im) ho->fnomask_im[(size_t)i0 * cF + q] = fn[q].y;
            }
            for (size_t q = 0; q < fm.size(); ++q) {
                if (ho->fmask_re) ho->fmask_re[(size_t)i0 * cF + q] = fm[q].x;
                if (ho->fmask_im) ho->fmask_im[(size_t)i0 * cF + q] = fm[q].y;
            }
        }
        HIPCHK(hipStreamSynchronize(s));  // the staging buffers are reused by the next chunk
    }
    return ROME_OK;
}

// ------------------------------------------------------------------ learned pose proposal (C-ABI)
int rome_proposal_configure(rome_ctx* ctx, const rome_proposal_config* cfg) {
    if (ctx) ++ctx->state_epoch;  // kernels of captured pools read what this call changes
    if (!ctx || !cfg) return ROME_ERR_ARG;
    if (!ctx->configured || !ctx->have_sampling)
        return fail(ctx, ROME_ERR_ARG, "rome_proposal_configure needs rome_configure and rome_set_sampling");
    if (!ctx->is3d || !ctx->two_pass)
        return fail(ctx, ROME_ERR_UNSUPPORTED, "the learned proposal needs 3D with adaptive oversampling");
    const int nb = cfg->nr_psi_bins;
    if (cfg->hidden <= 0 || cfg->hidden % 32 || cfg->nr_coef <= 0 || nb <= 0 || ctx->PSI % nb || cfg->top_k <= 0 ||
        cfg->target_cells <= 0 || cfg->target_cells > 1024 || cfg->max_train < 0 || cfg->audit_fraction < 0 ||
        cfg->audit_fraction > 1)
        return fail(ctx, ROME_ERR_ARG, "bad proposal configuration");
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ProposalState& P = ctx->prop;
    P.ready = false;
    P.cfg = *cfg;
    P.mode = ROME_PROPOSAL_OFF;
    P.K = ctx->K;
    P.D = ctx->D;
    P.PSI = ctx->PSI;
    P.cur = ctx->cfg.current_size;
    P.nbins = nb;
    P.C = 4 * ctx->D * nb;  // directions one HEALPix order finer (4 NEST children per coarse direction)
    P.ncell = ctx->K * P.C;
    P.O = P.ncell + 2;
    P.O_pad = (P.O + 31) / 32 * 32;
    P.ncoef = cfg->nr_coef;
    P.Fin = 2 * P.ncoef;
    P.Fin_pad = (P.Fin + 31) / 32 * 32;
    P.H = cfg->hidden;
    P.gwords = (ctx->K * ctx->D * nb + 31) / 32;
    if (P.gwords > kMaxGroupWords || (size_t)P.ncell * sizeof(float) > 64 * 1024)
        return fail(ctx, ROME_ERR_UNSUPPORTED, "too many proposal cells (%d)", P.ncell);
    // input coefficients: the lowest frequencies of the stored half-transform (x >= 0, the x = 0 column
    // for y >= 0 only: its y < 0 half holds their conjugates), by |k| then by index
    const int cur = ctx->cfg.current_size, Fs = cur / 2 + 1;
    std::vector<std::pair<long long, int>> px;
    for (int i = 0; i < cur; ++i)
        for (int j = 0; j < Fs; ++j) {
            const int y = i < Fs ? i : i - cur;
            if (j == 0 && y < 0) continue;
            px.push_back({(long long)(j * j + y * y) * (1LL << 32) + i * Fs + j, i * Fs + j});
        }
    if ((int)px.size() < P.ncoef) return fail(ctx, ROME_ERR_ARG, "nr_coef exceeds the image's coefficients");
    std::sort(px.begin(), px.end());
    std::vector<int> cn(P.ncoef), ci(P.ncoef);
    for (int q = 0; q < P.ncoef; ++q) {
        cn[q] = px[q].second;
        const int i = cn[q] / Fs, j = cn[q] % Fs, y = i < Fs ? i : i - cur;
        ci[q] = std::min((int)std::lround(std::sqrt((double)(j * j + y * y))), ctx->oriF - 1);
    }
    HIPCHK(P.coef_n.upload(cn.data(), cn.size(), ctx->stream));
    HIPCHK(P.coef_ires.upload(ci.data(), ci.size(), ctx->stream));
    // the 8 nearest coarse directions of each (angles between the HEALPix pixel centres)
    int order = 0;
    while ((12 << (2 * order)) < ctx->D) ++order;
    if ((12 << (2 * order)) != ctx->D) return fail(ctx, ROME_ERR_UNSUPPORTED, "nr_dir is not a HEALPix grid");
    std::vector<double> ux(ctx->D), uy(ctx->D), uz(ctx->D);
    for (int d = 0; d < ctx->D; ++d) {
        double z, phi;
        healpix_pix2ang(order, d, z, phi);
        const double st = std::sqrt(std::max(0.0, 1.0 - z * z));
        ux[d] = st * std::cos(phi);
        uy[d] = st * std::sin(phi);
        uz[d] = z;
    }
    std::vector<int> nbr((size_t)ctx->D * 8, -1);
    for (int d = 0; d < ctx->D; ++d) {
        std::vector<std::pair<double, int>> c;
        for (int e = 0; e < ctx->D; ++e)
            if (e != d) c.push_back({-(ux[d] * ux[e] + uy[d] * uy[e] + uz[d] * uz[e]), e});
        std::partial_sort(c.begin(), c.begin() + std::min<size_t>(8, c.size()), c.end());
        for (int q = 0; q < 8 && q < (int)c.size(); ++q) nbr[(size_t)d * 8 + q] = c[q].second;
    }
    HIPCHK(P.nbr.upload(nbr.data(), nbr.size(), ctx->stream));
    // parameters: W ~ N(0, 1 / fan_in) (the output layer at a tenth, for near-uniform first proposals), b = 0
    const int rows[4] = {P.H, P.H, P.H, P.O}, cols[4] = {P.Fin_pad, P.H, P.H, P.H}, fan[4] = {P.Fin, P.H, P.H, P.H};
    P.rng = cfg->seed ^ 0x5DEECE66DULL;
    for (int l = 0; l < 4; ++l) {
        P.rows[l] = rows[l];
        P.cols[l] = cols[l];
        const size_t nw = (size_t)rows[l] * cols[l];
        std::vector<float> h(nw, 0.f);
        const double sd = (l == 3 ? 0.1 : 1.0) / std::sqrt((double)fan[l]);
        for (int r = 0; r < rows[l]; ++r)
            for (int q = 0; q < fan[l]; ++q) h[(size_t)r * cols[l] + q] = (float)(sd * gauss(P.rng));
        HIPCHK(P.w[2 * l].upload(h.data(), nw, ctx->stream));
        std::vector<float> b(rows[l], 0.f);
        HIPCHK(P.w[2 * l + 1].upload(b.data(), b.size(), ctx->stream));
        for (int q = 2 * l; q <= 2 * l + 1; ++q) {
            const size_t nq = (q & 1) ? (size_t)rows[l] : nw;
            HIPCHK(P.m[q].reserve(nq));
            HIPCHK(P.v[q].reserve(nq));
            HIPCHK(P.g[q].reserve(nq));
            HIPCHK(hipMemsetAsync(P.m[q].p, 0, nq * sizeof(float), ctx->stream));
            HIPCHK(hipMemsetAsync(P.v[q].p, 0, nq * sizeof(float), ctx->stream));
        }
        HIPCHK(P.wq[l].reserve(nw));
        if (l > 0) {
            const size_t nt = (size_t)cols[l] * (l == 3 ? P.O_pad : rows[l]);
            HIPCHK(P.wqT[l].reserve(nt));
            HIPCHK(hipMemsetAsync(P.wqT[l].p, 0, nt * 2, ctx->stream));
        }
    }
    P.adam_t = 0;
    const size_t mt = (size_t)std::max(cfg->max_train, 1), tc = (size_t)cfg->target_cells;
    HIPCHK(P.tx.reserve(mt * P.Fin_pad));
    HIPCHK(P.tcell.reserve(mt * tc));
    HIPCHK(P.tprob.reserve(mt * tc));
    HIPCHK(P.tcnt.reserve(mt));
    HIPCHK(P.tshift.reserve(2 * mt));
    HIPCHK(P.loss.reserve(1));
    P.ntrain = 0;
    int r = prop_refresh(ctx);
    if (r) return r;
    P.audit_pos = 0;
    P.st_pools = P.st_images = P.st_all = P.st_audited = P.st_fallback = P.st_uncertain = 0;
    P.st_cand = 0.;
    P.st_min_mass = 1.;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    P.ready = true;
    return ROME_OK;
}

int rome_proposal_mode(rome_ctx* ctx, int mode) {
    if (ctx) ++ctx->state_epoch;  // kernels of captured pools read what this call changes
    if (!ctx || mode < ROME_PROPOSAL_OFF || mode > ROME_PROPOSAL_USE) return ROME_ERR_ARG;
    if (mode != ROME_PROPOSAL_OFF && !ctx->prop.ready)
        return fail(ctx, ROME_ERR_ARG, "rome_proposal_configure has not been called");
    if (mode == ROME_PROPOSAL_USE && ctx->cfg.precision == ROME_PREC_FP32)
        return fail(ctx, ROME_ERR_UNSUPPORTED, "the learned proposal needs precision BF16 or FP16");
    ctx->prop.mode = mode;
    return ROME_OK;
}

int rome_proposal_set_thresholds(rome_ctx* ctx, int top_k, double min_mass, double min_audit_mass) {
    if (!ctx || top_k < 1 || !(min_mass >= 0. && min_mass <= 1.) || !(min_audit_mass >= 0. && min_audit_mass <= 1.))
        return ROME_ERR_ARG;
    if (!ctx->prop.ready) return fail(ctx, ROME_ERR_ARG, "rome_proposal_configure has not been called");
    ctx->prop.cfg.top_k = top_k;
    ctx->prop.cfg.min_mass = min_mass;
    ctx->prop.cfg.min_audit_mass = min_audit_mass;
    return ROME_OK;
}

int rome_proposal_train(rome_ctx* ctx, int epochs, int batch, double lr, double* loss_per_epoch) {
    if (!ctx || epochs < 0 || batch <= 0 || batch % 32 || !(lr > 0)) return ROME_ERR_ARG;
    ProposalState& P = ctx->prop;
    if (!P.ready) return fail(ctx, ROME_ERR_ARG, "rome_proposal_configure has not been called");
    if (P.ntrain < batch) return fail(ctx, ROME_ERR_ARG, "training set (%d images) smaller than a batch", P.ntrain);
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    int rc = prop_acts(ctx, batch);
    if (rc) return rc;
    HIPCHK(P.bidx.reserve(P.ntrain));
    std::vector<int> perm(P.ntrain);
    const int nb = P.ntrain / batch;
    for (int e = 0; e < epochs; ++e) {
        for (int i = 0; i < P.ntrain; ++i) perm[i] = i;
        for (int i = P.ntrain - 1; i > 0; --i) std::swap(perm[i], perm[(int)(splitmix(P.rng) % (unsigned long long)(i + 1))]);
        HIPCHK(hipMemcpyAsync(P.bidx.p, perm.data(), perm.size() * sizeof(int), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemsetAsync(P.loss.p, 0, sizeof(float), s));
        for (int b = 0; b < nb; ++b) {
            const int* idx = P.bidx.p + (size_t)b * batch;
            HIPCHK(mlp_gather_rows(s, P.tx.p, idx, batch, P.Fin_pad, P.x.p));
            HIPCHK(mlp_gather_targets(s, idx, batch, P.cfg.target_cells, P.tcell.p, P.tprob.p, P.tcnt.p, P.tshift.p,
                                      P.bcell.p, P.bprob.p, P.bcnt.p, P.bshift.p));
            if ((rc = prop_forward(ctx, batch)) || (rc = prop_backward(ctx, batch)) || (rc = prop_adam(ctx, lr))) return rc;
        }
        float l = 0.f;
        HIPCHK(hipMemcpyAsync(&l, P.loss.p, sizeof(float), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));  // also keeps the host permutation alive until its copy is done
        if (loss_per_epoch) loss_per_epoch[e] = nb ? l / nb : 0.0;
    }
    return ROME_OK;
}

int rome_proposal_stats(rome_ctx* ctx, double* out) {
    if (!ctx || !out) return ROME_ERR_ARG;
    const ProposalState& P = ctx->prop;
    out[0] = (double)P.st_pools;
    out[1] = (double)P.st_images;
    out[2] = P.st_images ? P.st_cand / (double)P.st_images : 0.;
    out[3] = (double)P.st_all;
    out[4] = (double)P.st_audited;
    out[5] = P.st_min_mass;
    out[6] = (double)P.st_fallback;
    out[7] = (double)P.ntrain;
    return ROME_OK;
}

int rome_proposal_get_params(rome_ctx* ctx, rome_proposal_params* o) {
    if (!ctx || !o) return ROME_ERR_ARG;
    ProposalState& P = ctx->prop;
    if (!P.ready) return fail(ctx, ROME_ERR_ARG, "rome_proposal_configure has not been called");
    HIPCHK(hipSetDevice(ctx->device));
    float* dst[8] = {o->W1, o->b1, o->W2, o->b2, o->W3, o->b3, o->W4, o->b4};
    HIPCHK(hipStreamSynchronize(ctx->stream));
    for (int q = 0; q < 8; ++q) {
        if (!dst[q]) continue;
        const int l = q / 2;
        if (q & 1) {
            HIPCHK(hipMemcpy(dst[q], P.w[q].p, (size_t)P.rows[l] * sizeof(float), hipMemcpyDeviceToHost));
        } else {
            const int c = l == 0 ? P.Fin : P.cols[l];
            HIPCHK(hipMemcpy2D(dst[q], c * sizeof(float), P.w[q].p, P.cols[l] * sizeof(float), c * sizeof(float),
                               P.rows[l], hipMemcpyDeviceToHost));
        }
    }
    return ROME_OK;
}

int rome_proposal_set_params(rome_ctx* ctx, const rome_proposal_params* in) {
    if (!ctx || !in) return ROME_ERR_ARG;
    ProposalState& P = ctx->prop;
    if (!P.ready) return fail(ctx, ROME_ERR_ARG, "rome_proposal_configure has not been called");
    HIPCHK(hipSetDevice(ctx->device));
    const float* src[8] = {in->W1, in->b1, in->W2, in->b2, in->W3, in->b3, in->W4, in->b4};
    HIPCHK(hipStreamSynchronize(ctx->stream));
    for (int q = 0; q < 8; ++q) {
        if (!src[q]) return fail(ctx, ROME_ERR_ARG, "parameter %d missing", q);
        const int l = q / 2;
        if (q & 1) {
            HIPCHK(hipMemcpy(P.w[q].p, src[q], (size_t)P.rows[l] * sizeof(float), hipMemcpyHostToDevice));
        } else {
            const int c = l == 0 ? P.Fin : P.cols[l];
            if (l == 0) HIPCHK(hipMemset(P.w[q].p, 0, (size_t)P.rows[l] * P.cols[l] * sizeof(float)));
            HIPCHK(hipMemcpy2D(P.w[q].p, P.cols[l] * sizeof(float), src[q], c * sizeof(float), c * sizeof(float),
                               P.rows[l], hipMemcpyHostToDevice));
        }
    }
    int r = prop_refresh(ctx);
    if (r) return r;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return ROME_OK;
}

namespace {
// host features [n][Fin] -> P.x (BF16, zero-padded rows)
int prop_upload_x(rome_ctx* ctx, int n, const float* x) {
    ProposalState& P = ctx->prop;
    int r = prop_acts(ctx, n);
    if (r) return r;
    std::vector<float> h((size_t)n * P.Fin_pad, 0.f);
    for (int i = 0; i < n; ++i) std::memcpy(&h[(size_t)i * P.Fin_pad], x + (size_t)i * P.Fin, P.Fin * sizeof(float));
    DevBuf<float> tmp;
    HIPCHK(tmp.upload(h.data(), h.size(), ctx->stream));
    HIPCHK(mlp_to_bf16(ctx->stream, tmp.p, P.x.p, h.size()));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return ROME_OK;
}
}  // namespace

int rome_proposal_forward(rome_ctx* ctx, int n, const float* x, float* z) {
    if (!ctx || n <= 0 || !x || !z) return ROME_ERR_ARG;
    ProposalState& P = ctx->prop;
    if (!P.ready) return fail(ctx, ROME_ERR_ARG, "rome_proposal_configure has not been called");
    HIPCHK(hipSetDevice(ctx->device));
    int r = prop_upload_x(ctx, n, x);
    if (r || (r = prop_forward(ctx, n))) return r;
    HIPCHK(hipMemcpy2DAsync(z, P.O * sizeof(float), P.z.p, P.O_pad * sizeof(float), P.O * sizeof(float), n,
                            hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return ROME_OK;
}

int rome_proposal_gradients(rome_ctx* ctx, int n, const float* x, const float* target, const float* shift,
                            double* loss, rome_proposal_params* grads) {
    if (!ctx || n <= 0 || n % 32 || !x || !target || !shift) return ROME_ERR_ARG;
    ProposalState& P = ctx->prop;
    if (!P.ready) return fail(ctx, ROME_ERR_ARG, "rome_proposal_configure has not been called");
    HIPCHK(hipSetDevice(ctx->device));
    const int tc = P.cfg.target_cells;
    std::vector<int> cell((size_t)n * tc, 0), cnt(n, 0);
    std::vector<float> prob((size_t)n * tc, 0.f);
    for (int i = 0; i < n; ++i)
        for (int c = 0; c < P.ncell; ++c) {
            const float p = target[(size_t)i * P.ncell + c];
            if (p == 0.f) continue;
            if (cnt[i] == tc) return fail(ctx, ROME_ERR_ARG, "target row %d has more than %d cells", i, tc);
            cell[(size_t)i * tc + cnt[i]] = c;
            prob[(size_t)i * tc + cnt[i]++] = p;
        }
    int r = prop_upload_x(ctx, n, x);
    if (r) return r;
    hipStream_t s = ctx->stream;
    HIPCHK(hipMemcpyAsync(P.bcell.p, cell.data(), cell.size() * sizeof(int), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(P.bprob.p, prob.data(), prob.size() * sizeof(float), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(P.bcnt.p, cnt.data(), n * sizeof(int), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(P.bshift.p, shift, 2 * (size_t)n * sizeof(float), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemsetAsync(P.loss.p, 0, sizeof(float), s));
    if ((r = prop_forward(ctx, n)) || (r = prop_backward(ctx, n))) return r;
    float l = 0.f;
    HIPCHK(hipMemcpyAsync(&l, P.loss.p, sizeof(float), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (loss) *loss = l;
    if (grads) {
        float* dst[8] = {grads->W1, grads->b1, grads->W2, grads->b2, grads->W3, grads->b3, grads->W4, grads->b4};
        for (int q = 0; q < 8; ++q) {
            if (!dst[q]) continue;
            const int lq = q / 2;
            if (q & 1) {
                HIPCHK(hipMemcpy(dst[q], P.g[q].p, (size_t)P.rows[lq] * sizeof(float), hipMemcpyDeviceToHost));
            } else {
                const int c = lq == 0 ? P.Fin : P.cols[lq];
                HIPCHK(hipMemcpy2D(dst[q], c * sizeof(float), P.g[q].p, P.cols[lq] * sizeof(float), c * sizeof(float),
                                   P.rows[lq], hipMemcpyDeviceToHost));
            }
        }
    }
    return ROME_OK;
}

int rome_proposal_training_rows(rome_ctx* ctx, int row0, int n, float* x, float* target, float* shift) {
    if (!ctx || n < 0 || row0 < 0) return ROME_ERR_ARG;
    ProposalState& P = ctx->prop;
    if (!P.ready) return fail(ctx, ROME_ERR_ARG, "rome_proposal_configure has not been called");
    if (row0 + n > P.ntrain) return fail(ctx, ROME_ERR_ARG, "rows past the training set (%d)", P.ntrain);
    if (n == 0) return ROME_OK;
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    const int tc = P.cfg.target_cells;
    std::vector<uint16_t> hx((size_t)n * P.Fin_pad);
    std::vector<int> cell((size_t)n * tc), cnt(n);
    std::vector<float> prob((size_t)n * tc), sh(2 * (size_t)n);
    HIPCHK(hipMemcpy(hx.data(), P.tx.p + (size_t)row0 * P.Fin_pad, hx.size() * 2, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(cell.data(), P.tcell.p + (size_t)row0 * tc, cell.size() * sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(prob.data(), P.tprob.p + (size_t)row0 * tc, prob.size() * sizeof(float), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(cnt.data(), P.tcnt.p + row0, n * sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(sh.data(), P.tshift.p + 2 * (size_t)row0, sh.size() * sizeof(float), hipMemcpyDeviceToHost));
    for (int i = 0; i < n; ++i) {
        if (x)
            for (int q = 0; q < P.Fin; ++q) {
                const uint32_t b = (uint32_t)hx[(size_t)i * P.Fin_pad + q] << 16;
                float f;
                std::memcpy(&f, &b, 4);
                x[(size_t)i * P.Fin + q] = f;
            }
        if (target) {
            std::fill(target + (size_t)i * P.ncell, target + (size_t)(i + 1) * P.ncell, 0.f);
            for (int q = 0; q < cnt[i]; ++q) target[(size_t)i * P.ncell + cell[(size_t)i * tc + q]] += prob[(size_t)i * tc + q];
        }
        if (shift) {
            shift[2 * i] = sh[2 * i];
            shift[2 * i + 1] = sh[2 * i + 1];
        }
    }
    return ROME_OK;
}

// ------------------------------------------------------------------ offline reconstruction (backprojection)
int rome_backproject_images(rome_ctx* ctx, int iclass, int n, const double* fre, const double* fim, const double* fctf,
                            const double* euler, const double* shift) {
    if (!ctx || n < 0 || (n > 0 && (!fre || !fim || !euler || !shift))) return ROME_ERR_ARG;
    if (!ctx->configured) return fail(ctx, ROME_ERR_ARG, "rome_configure has not been called");
    if (iclass < 0 || iclass >= ctx->K) return fail(ctx, ROME_ERR_ARG, "class index out of range");
    if (n == 0) return ROME_OK;
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    const int F2 = ctx->F2f, Pf = ctx->Pf;
    const int chunk = 2048;
    DevBuf<float2> X, S;
    DevBuf<float> C;
    DevBuf<double> Ai, sx, sy;
    std::vector<float2> hx((size_t)std::min(n, chunk) * F2);
    std::vector<float> hc((size_t)std::min(n, chunk) * F2);
    std::vector<double> ha(6 * (size_t)std::min(n, chunk)), hsx(std::min(n, chunk)), hsy(std::min(n, chunk));
    for (int i0 = 0; i0 < n; i0 += chunk) {
        const int m = std::min(chunk, n - i0);
        for (size_t q = 0; q < (size_t)m * F2; ++q) {
            hx[q] = make_float2((float)fre[(size_t)i0 * F2 + q], (float)fim[(size_t)i0 * F2 + q]);
            hc[q] = fctf && ctx->cfg.do_ctf_correction ? (float)fctf[(size_t)i0 * F2 + q] : 1.f;
        }
        for (int i = 0; i < m; ++i) {
            double A[9];
            const double* e = euler + 3 * (size_t)(i0 + i);
            if (ctx->is3d) euler_matrix(e[0], e[1], e[2], A);
            else euler_matrix(0, 0, e[2], A);
            rot6(A, &ha[6 * (size_t)i]);
            hsx[i] = shift[2 * (size_t)(i0 + i)];
            hsy[i] = shift[2 * (size_t)(i0 + i) + 1];
        }
        HIPCHK(X.upload(hx.data(), (size_t)m * F2, s));
        HIPCHK(C.upload(hc.data(), (size_t)m * F2, s));
        HIPCHK(Ai.upload(ha.data(), 6 * (size_t)m, s));
        HIPCHK(sx.upload(hsx.data(), m, s));
        HIPCHK(sy.upload(hsy.data(), m, s));
        HIPCHK(S.reserve((size_t)m * Pf));
        HIPCHK(launch_phase_table(s, sx.p, sy.p, m, ctx->f_xy.p, Pf, ctx->cfg.ori_size, ctx->sintab.p, ctx->costab.p,
                                  S.p));
        BpImageArgs ba{X.p,     C.p,       S.p,     Ai.p,       F2,        Pf,      ctx->f_nfine.p,
                       ctx->f_xy.p, ctx->is3d, ctx->pad, ctx->padF, ctx->cgrid, ctx->r_max,
                       ctx->bp_data.p + 2 * ctx->grid_n * iclass, ctx->bp_weight.p + ctx->grid_n * iclass};
        HIPCHK(launch_bp_images(s, m, ba));
        HIPCHK(hipStreamSynchronize(s));  // the host staging is reused by the next chunk
    }
    return ROME_OK;
}
