// diff2 -> weight conversion and significance (K3):
//   convertSquaredDifferencesToWeights  src/map3d_optimizer.cpp:1610-1810 (2D :1077-1224)
//   findAllSignificantPoints            src/map3d_optimizer.cpp:1812-1931 (2D :1226-1311)
//
// One 1024-thread block per image.  With u = d - ln(pdf_direction * pdf_offset) the weight of
// an entry is w = exp(dmin - u); wherever a weight is used this kernel evaluates the reference's
// own expression pdf_direction * pdf_offset * exp(-(diff2 - min_diff2)) (0 beyond 700).
//
//   scan   : one pass over the image's entries.  Entries within kWin nats of the lower bound
//            u0 <= min u get their exact weight: sum, max / argmax, a 1/32-nat log histogram and
//            a per-image candidate list in HBM.  Entries further away weigh < exp(-48) of the
//            maximum (the window is re-anchored when the priors spread u by more than 16 nats),
//            below the FP64 resolution of the sum for any grid of up to 1e6 entries.
//   select : the bucket where the descending cumulative mass crosses adaptive_fraction * sum,
//            refined while it holds more candidates than the LDS sort; exact weights of that
//            bucket sorted and accumulated from the top as the reference's sorted loop does
//            -> threshold and NR_SIGN.
//   mark   : (coarse) the samples with w >= threshold: their keys m*T + t per image, or bits of
//            the significance mask.
// When an image has more candidates than the list holds, select / mark rescan the entries.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_util.h"
#include "kernels.h"

namespace rome {

namespace {

constexpr double kWin = kWeightWindow;  // histogram window (2048 bins x 1/32 nat)
constexpr double kAnchorSlack = 16.0;

struct WeightOut {
    double* dmin;
    double* sumw;
    double* maxw;
    long long* amax;
    double* thr;
    int* nsig;
    int* status;
    double* tail;
};

struct Cand {
    double* w;
    double* du;
    long long* e;
    int cap;
};

struct WShared {
    double poff[kMaxKT];
    double lpoff[kMaxKT];
    float lpoff_f[kMaxKT];
    double hsum[kBinCount];
    unsigned int hcnt[kBinCount];
    double cw[kSortCap];
    long long ce[kSortCap];
    double red_d[16];
    long long red_l[16];
    int ncand_sort;
    int ncand;
    int nmark;
    int done;
    double lo, width, mbefore;
    long long cbefore;
    double thr;
    long long nsig;
};

// Pre(e0, bound, lpoff_f) -> bit mask over the Pre::VW entries e0 .. e0 + VW - 1 (those < count): a
// cheap FP32 test whose bit is clear only for entries whose exact FP64 du = u - u0 is >= kWin when
// bound = (float)(u0 + kWin) + margin (the margin covers the FP32 rounding).
// Key(e, m, t): the entry's position in the reference's loop order; the max-posterior entry is the
// first maximum in that order (ties between equal weights).
template <typename Get, typename Pre, typename Key>
__device__ void weight_pass(int img, size_t count, double cimg, const WeightParams& wp, Get get, Pre pre, Key key,
                            WeightOut wo, WShared& sh, SigArgs sig, Cand cand, const int* emin_key,
                            TailArgs tail = TailArgs{nullptr, nullptr}) {
    constexpr int NT = 1024;
    const int tid = threadIdx.x;
    const int T = wp.T;
    double lpo_max = -INFINITY;
    for (int q = tid; q < wp.K * T; q += NT) {
        const int k = q / T, t = q % T;
        const double po = pdf_offset(wp.trans_x[t], wp.trans_y[t], wp.old_offx[img], wp.old_offy[img], wp.prior_x[k],
                                     wp.prior_y[k], wp.sigma2_offset);
        sh.poff[q] = po;
        sh.lpoff[q] = po > 0 ? log(po) : -INFINITY;
        sh.lpoff_f[q] = (float)sh.lpoff[q];
        lpo_max = fmax(lpo_max, sh.lpoff[q]);
    }
    lpo_max = block_reduce(lpo_max, [](double a, double b) { return fmax(a, b); }, sh.red_d);
    auto u_of = [&](double d, int m, int t) { return d - (wp.lpdf_m[m] + sh.lpoff[wp.k_m[m] * T + t]); };
    auto w_of = [&](double d, int m, int t, double dmin) {
        const double dd = (d + cimg) - dmin;
        const double w = wp.pdf_m[m] * sh.poff[wp.k_m[m] * T + t];
        return dd > 700. ? 0. : w * exp(-dd);
    };
    // ---- minimum of d (the coarse pass gets it from the cross-term epilogue)
    double mn;
    if (emin_key) {
        mn = (double)float_from_order_key(emin_key[img]);
    } else {
        mn = INFINITY;
        for (size_t e = tid; e < count; e += NT) {
            double d;
            int m, t;
            if (get(e, d, m, t)) mn = fmin(mn, d);
        }
        mn = block_reduce(mn, [](double a, double b) { return fmin(a, b); }, sh.red_d);
    }
    const double dmin = mn + cimg;  // (d + c) is monotone in d
    double u0 = mn - (wp.lpdf_max + lpo_max);
    // FP32 pre-test bound: rounding of d - lpdf - lpoff and of u0 + kWin stays far below the margin
    // (relative 2^-22 of magnitudes that the 1e-4 term covers; log priors are > -746)
    auto pre_bound = [&](double a0) {
        return (float)(a0 + kWin) + 1.0f + 1e-4f * (float)(fabs(a0) + fabs(mn) + kWin + fabs(wp.lpdf_max) + fabs(lpo_max));
    };
    float bound = pre_bound(u0);
    // every entry that may lie inside the window, each thread in increasing e (four groups of VW in
    // flight per step): fn(e)
    auto for_window = [&](float bnd, auto&& fn) {
        constexpr size_t VW = Pre::VW, STEP = (size_t)NT * VW;
        for (size_t e0 = (size_t)tid * VW; e0 < count; e0 += 4 * STEP) {
            unsigned m[4];
#pragma unroll
            for (int g = 0; g < 4; ++g) m[g] = pre(e0 + g * STEP, bnd, sh.lpoff_f);
#pragma unroll
            for (int g = 0; g < 4; ++g)
                while (m[g]) {
                    const int u = __builtin_ctz(m[g]);
                    m[g] &= m[g] - 1;
                    fn(e0 + g * STEP + u);
                }
        }
    };
    // ---- scan
    double sum = 0., mx = -1.;
    long long am = -1;
    for (int attempt = 0; attempt < 2; ++attempt) {
        for (int b = tid; b < kBinCount; b += NT) {
            sh.hsum[b] = 0.;
            sh.hcnt[b] = 0;
        }
        if (tid == 0) sh.ncand = 0;
        __syncthreads();
        sum = 0.;
        mx = -1.;
        am = -1;
        double umin = INFINITY;
        for_window(bound, [&](size_t e) {
            double d;
            int m, t;
            if (!get(e, d, m, t)) return;
            const double u = u_of(d, m, t);
            umin = fmin(umin, u);  // over the entries that pass the pre-test
            const double du = u - u0;
            if (!(du < kWin)) return;
            const double w = w_of(d, m, t, dmin);
            sum += w;
            const long long ke = key(e, m, t);
            if (w > mx || (w == mx && ke < am)) {
                mx = w;
                am = ke;
            }
            if (w > 0.) {
                const int b = du > 0. ? (int)(du * kBinScale) : 0;
                atomicAdd(&sh.hsum[b], w);
                atomicAdd(&sh.hcnt[b], 1u);
                const int slot = atomicAdd(&sh.ncand, 1);
                if (slot < cand.cap) {
                    const size_t o = (size_t)img * cand.cap + slot;
                    cand.w[o] = w;
                    cand.du[o] = du;
                    cand.e[o] = (long long)e;
                }
            }
        });
        umin = block_reduce(umin, [](double a, double b) { return fmin(a, b); }, sh.red_d);
        // every entry with du < kWin passed the pre-test, so umin is the exact minimum of u whenever
        // umin - u0 < kWin; in particular umin - u0 <= kAnchorSlack holds iff it holds for all entries
        if (!(umin - u0 > kAnchorSlack)) break;
        if (!(umin - u0 < kWin)) {  // nothing inside the window: exact minimum from a full pass
            umin = INFINITY;
            for (size_t e = tid; e < count; e += NT) {
                double d;
                int m, t;
                if (get(e, d, m, t)) umin = fmin(umin, u_of(d, m, t));
            }
            umin = block_reduce(umin, [](double a, double b) { return fmin(a, b); }, sh.red_d);
        }
        u0 = umin;  // priors spread u widely: anchor the window at the true minimum and rescan
        bound = pre_bound(u0);
    }
    sum = block_reduce(sum, [](double a, double b) { return a + b; }, sh.red_d);
    // mass of the previous pass's unmarked entries, re-anchored from its minimum to this one:
    // pdf exp(-(d - dmin)) = pdf exp(-(d - dmin_prev)) exp(dmin - dmin_prev)
    if (tail.mass) sum += tail.mass[img] * exp(dmin - tail.dmin[img]);
    {
        const double m2 = block_reduce(mx, [](double a, double b) { return fmax(a, b); }, sh.red_d);
        const long long c = (mx == m2 && am >= 0) ? am : (long long)0x7fffffffffffffffLL;
        am = block_reduce(c, [](long long a, long long b) { return a < b ? a : b; }, sh.red_l);
        mx = m2;
    }
    const double target = wp.adaptive_fraction * sum;
    const int ncand = sh.ncand;
    const bool listed = ncand <= cand.cap;
    // visit(lo, hi, fn): every entry with weight > 0 and lo <= du < hi, as (w, du, e)
    auto visit = [&](double lo, double hi, auto&& fn) {
        if (listed) {
            for (int q = tid; q < ncand; q += NT) {
                const size_t o = (size_t)img * cand.cap + q;
                const double du = cand.du[o];
                if (du >= lo && du < hi) fn(cand.w[o], du, cand.e[o]);
            }
        } else {
            for_window(bound, [&](size_t e) {
                double d;
                int m, t;
                if (!get(e, d, m, t)) return;
                const double du = u_of(d, m, t) - u0;
                if (!(du >= lo && du < hi && du < kWin)) return;
                const double w = w_of(d, m, t, dmin);
                if (w > 0.) fn(w, du, (long long)e);
            });
        }
    };
    // ---- select
    if (tid == 0) {
        sh.lo = 0.;
        sh.width = 1.0 / kBinScale;
        sh.mbefore = 0.;
        sh.cbefore = 0;
        sh.done = 0;
    }
    __syncthreads();
    int status = 0;
    for (int level = 0;; ++level) {
        if (tid == 0) {
            double cum = sh.mbefore;
            long long cnt = sh.cbefore;
            int bstar = -1, last = -1;
            for (int b = 0; b < kBinCount; ++b) {
                if (sh.hcnt[b] == 0) continue;
                last = b;
                if (cum + sh.hsum[b] > target) {
                    bstar = b;
                    break;
                }
                cum += sh.hsum[b];
                cnt += sh.hcnt[b];
            }
            if (bstar < 0 && last >= 0) {  // rounding left the crossing unreached: last bucket
                bstar = last;
                cum -= sh.hsum[last];
                cnt -= sh.hcnt[last];
            }
            if (bstar < 0) {
                sh.done = 2;
            } else {
                sh.mbefore = cum;
                sh.cbefore = cnt;
                sh.lo = sh.lo + bstar * sh.width;
                sh.done = (sh.hcnt[bstar] <= (unsigned)kSortCap || level >= 2) ? 1 : 0;
                sh.width = sh.width / (double)kBinCount;
            }
            sh.ncand_sort = 0;
        }
        __syncthreads();
        if (sh.done == 2) {
            status = 1;
            break;
        }
        const double lo = sh.lo, hi = lo + sh.width * kBinCount;
        if (sh.done == 1) {
            visit(lo, hi, [&](double w, double, long long e) {
                const int slot = atomicAdd(&sh.ncand_sort, 1);
                if (slot < kSortCap) {
                    sh.cw[slot] = w;
                    sh.ce[slot] = e;
                }
            });
            __syncthreads();
            const int nc = sh.ncand_sort < kSortCap ? sh.ncand_sort : kSortCap;
            if (sh.ncand_sort > kSortCap) status = 2;
            int np2 = 1;
            while (np2 < nc) np2 <<= 1;
            for (int q = nc + tid; q < np2; q += NT) {
                sh.cw[q] = -1.0;
                sh.ce[q] = -1;
            }
            __syncthreads();
            for (int ksz = 2; ksz <= np2; ksz <<= 1)
                for (int j = ksz >> 1; j > 0; j >>= 1) {
                    for (int q = tid; q < np2; q += NT) {
                        const int ixj = q ^ j;
                        if (ixj > q) {
                            const bool desc = ((q & ksz) == 0);
                            const double a1 = sh.cw[q], a2 = sh.cw[ixj];
                            if (desc ? (a1 < a2) : (a1 > a2)) {
                                sh.cw[q] = a2;
                                sh.cw[ixj] = a1;
                                const long long t1 = sh.ce[q];
                                sh.ce[q] = sh.ce[ixj];
                                sh.ce[ixj] = t1;
                            }
                        }
                    }
                    __syncthreads();
                }
            if (tid == 0) {
                double frac = sh.mbefore;
                long long cnt = sh.cbefore;
                double my = 0.;
                for (int q = 0; q < nc; ++q) {
                    ++cnt;
                    my = sh.cw[q];
                    frac += my;
                    if (frac > target) break;
                }
                sh.thr = my;
                sh.nsig = cnt;
            }
            __syncthreads();
            break;
        }
        // refine the crossing bucket
        for (int b = tid; b < kBinCount; b += NT) {
            sh.hsum[b] = 0.;
            sh.hcnt[b] = 0;
        }
        __syncthreads();
        const double sw = sh.width;
        visit(lo, hi, [&](double w, double du, long long) {
            int b = (int)((du - lo) / sw);
            b = b < 0 ? 0 : (b >= kBinCount ? kBinCount - 1 : b);
            atomicAdd(&sh.hsum[b], w);
            atomicAdd(&sh.hcnt[b], 1u);
        });
        __syncthreads();
    }
    const double thr = status == 1 ? 0. : sh.thr;
    if (tid == 0) {
        wo.dmin[img] = dmin;
        wo.sumw[img] = sum;
        wo.maxw[img] = mx;
        wo.amax[img] = am;
        wo.thr[img] = thr;
        if (wo.nsig) wo.nsig[img] = status == 1 ? 0 : (int)sh.nsig;
        wo.status[img] = status | (listed ? 0 : 4);
    }
    // ---- mark: w >= threshold, with the weight the reference compares (or the superset within
    // mark_margin nats of it); the mass left unmarked goes to wo.tail
    const double thr_mark = wp.mark_margin > 0. ? thr * exp(-wp.mark_margin) : thr;
    double marked = 0.;
    if (sig.keys) {
        if (tid == 0) sh.nmark = 0;
        __syncthreads();
        if (status != 1)
            visit(0.0, kWin, [&](double w, double, long long e) {
                if (!(w >= thr_mark)) return;
                double d;
                int m, t;
                get((size_t)e, d, m, t);
                marked += w;
                const int slot = atomicAdd(&sh.nmark, 1);
                if (slot < sig.cap) sig.keys[(size_t)img * sig.cap + slot] = m * T + t;
            });
        __syncthreads();
        if (tid == 0) sig.cnt[img] = sh.nmark;
    } else if (sig.mask) {
        const int TW = sig.TW;
        uint32_t* msk = sig.mask + (size_t)img * wp.Mtot * TW;
        for (size_t q = tid; q < (size_t)wp.Mtot * TW; q += NT) msk[q] = 0u;
        __syncthreads();
        if (status != 1)
            visit(0.0, kWin, [&](double w, double, long long e) {
                if (!(w >= thr_mark)) return;
                double d;
                int m, t;
                get((size_t)e, d, m, t);
                marked += w;
                atomicOr(&msk[(size_t)m * TW + (t >> 5)], 1u << (t & 31));
            });
    }
    if (wo.tail) {
        marked = block_reduce(marked, [](double a, double b) { return a + b; }, sh.red_d);
        if (tid == 0) wo.tail[img] = fmax(sum - marked, 0.);
    }
}

// Coarse entries e = t * Mtot + m (E rows are rotation-contiguous).
DEV void coarse_mt(size_t e, unsigned long long magic, int Mtot, int& m, int& t) {
    t = (int)(((unsigned long long)e * magic) >> 48);
    m = (int)e - t * Mtot;
}

// Coarse pre-test: d - lpdf_m - lpoff in FP32 from the per-rotation row {lpdf_m or -inf, k*T}, four
// consecutive entries per call (one 16-byte load of E and one translation when Mtot % 4 == 0).
struct CoarsePre {
    static constexpr int VW = 4;
    const float* Ei;
    size_t count;
    unsigned long long magic;
    int Mtot;
    const float2* mrow;
    bool vec;
    DEV unsigned operator()(size_t e0, float bound, const float* lpoff_f) const {
        if (e0 >= count) return 0u;
        unsigned msk = 0u;
        if (vec) {
            typedef float f4v __attribute__((ext_vector_type(4)));
            const f4v q = __builtin_nontemporal_load(reinterpret_cast<const f4v*>(Ei + e0));  // read once
            const float v[4] = {q[0], q[1], q[2], q[3]};
            int m0, t;
            coarse_mt(e0, magic, Mtot, m0, t);
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const float2 mr = mrow[m0 + u];
                if (v[u] - mr.x - lpoff_f[__float_as_int(mr.y) + t] < bound) msk |= 1u << u;
            }
        } else {
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const size_t e = e0 + u;
                if (e >= count) break;
                int m, t;
                coarse_mt(e, magic, Mtot, m, t);
                const float2 mr = mrow[m];
                if (Ei[e] - mr.x - lpoff_f[__float_as_int(mr.y) + t] < bound) msk |= 1u << u;
            }
        }
        return msk;
    }
};

// Fine pass: every entry goes to the exact test.
struct AllPre {
    static constexpr int VW = 1;
    size_t count;
    DEV unsigned operator()(size_t e0, float, const float*) const { return e0 < count ? 1u : 0u; }
};

__global__ __launch_bounds__(1024) void k_weights_coarse(WeightParams wp, const float* __restrict__ E,
                                                         const double* __restrict__ C2c,
                                                         const int* __restrict__ emin_key, WeightOut wo,
                                                         SigArgs sig, Cand cand) {
    __shared__ WShared sh;
    const int img = blockIdx.x;
    const int T = wp.T;
    const unsigned long long magic = wp.magicM;
    const int Mtot = wp.Mtot;
    const float* Ei = E + (size_t)img * Mtot * T;
    auto get = [&](size_t e, double& d, int& m, int& t) -> bool {
        coarse_mt(e, magic, Mtot, m, t);
        if (!(wp.pdf_m[m] > 0.)) return false;
        d = (double)Ei[e];
        return true;
    };
    const CoarsePre pre{Ei, (size_t)Mtot * T, magic, Mtot, wp.mrow, (Mtot & 3) == 0};
    auto key = [&](size_t, int m, int t) { return (long long)m * T + t; };  // (class, dir, psi, t) order
    weight_pass(img, (size_t)Mtot * T, C2c[img], wp, get, pre, key, wo, sh, sig, cand, emin_key);
}

// Coarse weights from the list the cross-term GEMM kept (every entry within kWin of the minimum
// u, and more): entries i = 0 .. count-1 of the image's list, argmax keys in the reference's order.
// 8 waves per SIMD: two 1024-thread blocks per CU (64 VGPRs; the few spills cost less than the
// second block gains on this latency-bound pass)
__global__ __launch_bounds__(1024, 8) void k_weights_list(WeightParams wp, ListArgs L, const double* __restrict__ C2c,
                                                       const int* __restrict__ emin_key, WeightOut wo,
                                                       SigArgs sig, Cand cand) {
    __shared__ WShared sh;
    const int img = blockIdx.x;
    const int cnt = L.count[img];
    if (cnt > L.cap) {  // the list overflowed: the caller reruns the batch with the full pose term
        if (threadIdx.x == 0) {
            wo.status[img] = 8;
            if (sig.keys) sig.cnt[img] = 0;
        }
        return;
    }
    const int T = wp.T;
    const unsigned long long magic = wp.magicM;
    const int Mtot = wp.Mtot;
    const unsigned* ent = L.ent + (size_t)img * L.cap;
    const float* val = L.val + (size_t)img * L.cap;
    auto get = [&](size_t i, double& d, int& m, int& t) -> bool {
        coarse_mt(ent[i], magic, Mtot, m, t);
        if (!(wp.pdf_m[m] > 0.)) return false;
        d = (double)val[i];
        return true;
    };
    const AllPre pre{(size_t)cnt};
    auto key = [&](size_t, int m, int t) { return (long long)m * T + t; };
    weight_pass(img, (size_t)cnt, C2c[img], wp, get, pre, key, wo, sh, sig, cand, emin_key);
}

// reach[img] = kWin + (max - min of ln pdf_offset over the image's (class, translation) pairs) + 1
// nat of slack (+inf when a pdf_offset is zero: the list then takes every entry and overflows).
__global__ void k_list_reach(WeightParams wp, float* __restrict__ reach) {
    __shared__ double red[16];
    const int img = blockIdx.x;
    double lo = INFINITY, hi = -INFINITY;
    for (int q = threadIdx.x; q < wp.K * wp.T; q += blockDim.x) {
        const int k = q / wp.T, t = q - k * wp.T;
        const double po = pdf_offset(wp.trans_x[t], wp.trans_y[t], wp.old_offx[img], wp.old_offy[img], wp.prior_x[k],
                                     wp.prior_y[k], wp.sigma2_offset);
        const double l = po > 0 ? log(po) : -INFINITY;
        lo = fmin(lo, l);
        hi = fmax(hi, l);
    }
    lo = block_reduce(lo, [](double a, double b) { return fmin(a, b); }, red);
    hi = block_reduce(hi, [](double a, double b) { return fmax(a, b); }, red);
    if (threadIdx.x == 0) reach[img] = lo > -INFINITY ? (float)(kWin + (hi - lo) + 1.0) : INFINITY;
}

// two blocks per CU, as k_weights_list
__global__ __launch_bounds__(1024, 8) void k_weights_fine(WeightParams wp, const double* __restrict__ D2,
                                                       const int* __restrict__ blk_task, const int* __restrict__ blk_t,
                                                       const int* __restrict__ task_m,
                                                       const long long* __restrict__ img_blk, int log2NE, WeightOut wo,
                                                       Cand cand, TailArgs tail, SigArgs sig) {
    __shared__ WShared sh;
    const int img = blockIdx.x;
    const long long b0 = img_blk[img];
    const size_t count = (size_t)(img_blk[img + 1] - b0) << log2NE;
    const double* Di = D2 + ((size_t)b0 << log2NE);
    auto get = [&](size_t e, double& d, int& m, int& t) -> bool {
        const long long b = b0 + (long long)(e >> log2NE);
        m = task_m[blk_task[b]];
        t = blk_t[b];
        d = Di[e];
        return true;
    };
    const AllPre pre{count};
    auto key = [](size_t e, int, int) { return (long long)e; };
    weight_pass(img, count, 0.0, wp, get, pre, key, wo, sh, sig, cand, nullptr, tail);
}

}  // namespace

hipError_t launch_weights_coarse(hipStream_t s, int n, const WeightParams& wp, const float* E, const double* C2c,
                                 const int* emin_key, const WeightOutArgs& wo, const SigArgs& sig,
                                 const CandArgs& c) {
    if (sig.keys && ((long long)wp.Mtot * wp.T > 0x7fffffffLL || sig.cap > kSigCap)) return hipErrorInvalidValue;
    WeightOut o{wo.dmin, wo.sumw, wo.maxw, wo.amax, wo.thr, wo.nsig, wo.status, wo.tail};
    k_weights_coarse<<<n, 1024, 0, s>>>(wp, E, C2c, emin_key, o, sig, Cand{c.w, c.du, c.e, c.cap});
    return hipGetLastError();
}

hipError_t launch_weights_list(hipStream_t s, int n, const WeightParams& wp, const ListArgs& list, const double* C2c,
                               const int* emin_key, const WeightOutArgs& wo, const SigArgs& sig,
                               const CandArgs& c) {
    if (wp.K * wp.T > kMaxKT) return hipErrorInvalidValue;
    if (sig.keys && ((long long)wp.Mtot * wp.T > 0x7fffffffLL || sig.cap > kSigCap)) return hipErrorInvalidValue;
    WeightOut o{wo.dmin, wo.sumw, wo.maxw, wo.amax, wo.thr, wo.nsig, wo.status, wo.tail};
    k_weights_list<<<n, 1024, 0, s>>>(wp, list, C2c, emin_key, o, sig, Cand{c.w, c.du, c.e, c.cap});
    return hipGetLastError();
}

hipError_t launch_list_reach(hipStream_t s, int n, const WeightParams& wp, float* reach) {
    if (n == 0) return hipSuccess;
    k_list_reach<<<n, 256, 0, s>>>(wp, reach);
    return hipGetLastError();
}

hipError_t launch_weights_fine(hipStream_t s, int n, const WeightParams& wp, const double* D2, const int* blk_task,
                               const int* blk_t, const int* task_m, const long long* img_blk, int NE,
                               const WeightOutArgs& wo, const CandArgs& c, TailArgs tail, SigArgs sig) {
    int log2NE = 0;
    while ((1 << log2NE) < NE) ++log2NE;
    if ((1 << log2NE) != NE) return hipErrorInvalidValue;
    if ((sig.keys || sig.mask) && NE != 1) return hipErrorInvalidValue;  // marking keys m*T + t: one entry per block
    if (sig.keys && ((long long)wp.Mtot * wp.T > 0x7fffffffLL || sig.cap > kSigCap)) return hipErrorInvalidValue;
    WeightOut o{wo.dmin, wo.sumw, wo.maxw, wo.amax, wo.thr, wo.nsig, wo.status, wo.tail};
    k_weights_fine<<<n, 1024, 0, s>>>(wp, D2, blk_task, blk_t, task_m, img_blk, log2NE, o, Cand{c.w, c.du, c.e, c.cap},
                                      tail, sig);
    return hipGetLastError();
}

}  // namespace rome
