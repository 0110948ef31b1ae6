// Binned 3D backprojection of the significant slices: the scatter of
// MyBackProjector::backproject (src/reconstruct.cpp:977-1164) reorganised so that the trilinear
// corner updates land in LDS instead of the HBM grids.
//
//   k_post  : per slice s = task * NOR + o, the weighted sum of the shifted unmasked images
//             V_s(p) = sum_t w_t/sumw * Minv CTF X_nomask(p) S_t(p) (float2 [S][Pf]) and
//             W_s = sum_t w_t/sumw (double [S]; 0: the slice has no significant sample).
//   count   : every (slice, pixel) with a positive weight, inside r_max -> trilinear base voxel ->
//             brick (8x8x8 base voxels of one class grid); counts per brick, first gathered per
//             block in an LDS hash table (one global atomic per distinct brick of the block).
//   scan    : exclusive prefix sums of the counts and of the chunks (kChunk records) per brick.
//   fill    : records (slice << 32 | pixel) grouped by brick; ranges reserved per distinct brick
//             of the block as in count.
//   brick   : one block per chunk of a brick's records (bricks near the origin, crossed by nearly
//             every slice, hold hundreds of thousands): the records' 8 corners accumulate in a
//             9x9x9 LDS box (base voxels plus the +1 corner halo), then each non-zero voxel goes to
//             the grids with one atomic per component.  Corner weights and values are computed
//             exactly as the direct scatter computes them; only the order of the FP32 additions
//             differs.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_util.h"
#include "kernels.h"

namespace rome {

namespace {

constexpr int kBrick = 8, kBox = kBrick + 1, kBox3 = kBox * kBox * kBox;
constexpr unsigned kChunk = 1024;  // records per brick block (41 KB of LDS: three blocks per CU)
constexpr int kHash = 512;         // LDS hash slots for the <= 256 bricks of one count / fill block

// Slot of brick b in the block's hash table (keys b + 1, 0 = empty; linear probing).  At most 256
// keys go into 512 slots, so a free slot always exists.
DEV int hash_slot(unsigned* keys, unsigned b) {
    int h = (int)((b * 2654435761u) >> 23);  // 9 bits
    for (;;) {
        const unsigned k = keys[h];
        if (k == b + 1) return h;
        if (k == 0) {
            const unsigned old = atomicCAS(&keys[h], 0u, b + 1);
            if (old == 0 || old == b + 1) return h;
        }
        h = (h + 1) & (kHash - 1);
    }
}

struct Sample {
    Coords c;
    int k;
    float wgt;
};

// The pixel's contribution geometry, or false when the direct scatter would skip it.
DEV bool bp_sample(const BpArgs& a, long long s, int p, Sample& out) {
    const double wsl = a.slice_w[s];
    if (!(wsl > 0.)) return false;
    const long long q = s / a.NOR;
    const int o = (int)(s - q * a.NOR);
    const int img = a.task_img[q], m = a.task_m[q];
    const int xy = a.f_xy[p];
    const int x = px_x(xy), y = px_y(xy);
    if (x * x + y * y > a.r_max * a.r_max) return false;
    const float wgt = (float)((double)a.MC2bp[(size_t)img * a.Pf + p] * wsl);
    if (!(wgt > 0.f)) return false;
    const int r = m % a.R;
    out.k = m / a.R;
    out.c = map_coords(a.Ai_fine + 6 * ((size_t)r * a.NOR + o), x, y, a.c, true);
    out.wgt = wgt;
    return true;
}

// brick of the base voxel: class-major, then z, y, x in bricks of 8 (grid coordinates are >= 0)
DEV int brick_of(const BpArgs& a, const Sample& sm) {
    const int bx = sm.c.x0 / kBrick, by = sm.c.y0 / kBrick, bz = sm.c.z0 / kBrick;
    return sm.k * a.nb_class + (bz * a.nby + by) * a.nbx + bx;
}

// Slices of the pool: the host's count, or (device-driven pools) the device task total x NOR; 0 when
// the pool is aborted or its slices exceed the slice buffer (the direct scatter took it).
DEV long long bp_slices(const BpArgs& a, long long S_host) {
    if (!a.ntask_dev) return S_host;
    if (*a.abort) return 0;
    const long long S = *a.ntask_dev * a.NOR;
    return S > a.slice_cap ? 0 : S;
}

DEV void bp_count_one(const BpArgs& a, long long s, int pb, unsigned* keys, unsigned* cnt) {
    if (!(a.slice_w[s] > 0.)) return;  // a slice without weight has no record (whole block)
    for (int h = threadIdx.x; h < kHash; h += blockDim.x) {
        keys[h] = 0u;
        cnt[h] = 0u;
    }
    __syncthreads();
    const int p = pb * blockDim.x + threadIdx.x;
    Sample sm;
    if (p < a.Pf && bp_sample(a, s, p, sm)) atomicAdd(&cnt[hash_slot(keys, (unsigned)brick_of(a, sm))], 1u);
    __syncthreads();
    for (int h = threadIdx.x; h < kHash; h += blockDim.x)
        if (keys[h]) atomicAdd(&a.count[keys[h] - 1], cnt[h]);
}

// blockIdx.x = slice, blockIdx.y = pixel block; device-driven (Q): a fixed grid of the same shape whose
// blocks stride over the slices of the slice total on the device.  No work counter: a device-scope
// atomic on one address serialises about a million items a pool.
template <bool Q>
__global__ __launch_bounds__(256) void k_bp_count(BpArgs a, long long S_host) {
    __shared__ unsigned keys[kHash], cnt[kHash];
    if constexpr (!Q) {
        bp_count_one(a, blockIdx.x, blockIdx.y, keys, cnt);
        return;
    }
    const long long S = bp_slices(a, S_host);
    for (long long sl = blockIdx.x; sl < S; sl += gridDim.x) {  // slices strided over the grid
        bp_count_one(a, sl, (int)blockIdx.y, keys, cnt);
        __syncthreads();  // the hash table of the next slice
    }
}

// Exclusive scans of count[0..nb) into off[0..nb] and of the chunks per brick into coff[0..nb];
// cursor cleared.  One block of 1024 threads over tiles of 4096 bricks: four consecutive bricks per
// thread, wave scans by shuffles, the 16 wave totals through LDS, a running carry between tiles.
__global__ __launch_bounds__(1024) void k_bp_scan(const unsigned* __restrict__ count, int nb,
                                                  unsigned* __restrict__ off, unsigned* __restrict__ coff,
                                                  unsigned* __restrict__ cursor) {
    __shared__ unsigned wa[16], wc[16];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    unsigned carry = 0u, ccarry = 0u;
    for (int t0 = 0; t0 < nb; t0 += 4096) {
        const int b = t0 + 4 * tid;
        unsigned c[4], cc[4], loc = 0u, cloc = 0u;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            c[j] = b + j < nb ? count[b + j] : 0u;
            cc[j] = (c[j] + kChunk - 1) / kChunk;
            loc += c[j];
            cloc += cc[j];
        }
        unsigned in = loc, cin = cloc;
        for (int o = 1; o < 64; o <<= 1) {
            const unsigned v = __shfl_up(in, o, 64), cv = __shfl_up(cin, o, 64);
            if (lane >= o) {
                in += v;
                cin += cv;
            }
        }
        if (lane == 63) {
            wa[wid] = in;
            wc[wid] = cin;
        }
        __syncthreads();
        unsigned pre = 0u, cpre = 0u, tot = 0u, ctot = 0u;
        for (int w = 0; w < 16; ++w) {
            if (w < wid) {
                pre += wa[w];
                cpre += wc[w];
            }
            tot += wa[w];
            ctot += wc[w];
        }
        unsigned run = carry + pre + in - loc, crun = ccarry + cpre + cin - cloc;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (b + j < nb) {
                off[b + j] = run;
                coff[b + j] = crun;
                cursor[b + j] = 0u;
            }
            run += c[j];
            crun += cc[j];
        }
        carry += tot;
        ccarry += ctot;
        __syncthreads();  // the wave totals are rewritten by the next tile
    }
    if (tid == 0) {
        off[nb] = carry;
        coff[nb] = ccarry;
    }
}

DEV void bp_fill_one(const BpArgs& a, long long s, int pb, unsigned* keys, unsigned* cnt, unsigned* base) {
    if (!(a.slice_w[s] > 0.)) return;  // a slice without weight has no record (whole block)
    for (int h = threadIdx.x; h < kHash; h += blockDim.x) {
        keys[h] = 0u;
        cnt[h] = 0u;
    }
    __syncthreads();
    const int p = pb * blockDim.x + threadIdx.x;
    Sample sm;
    int h = -1;
    unsigned rank = 0;
    if (p < a.Pf && bp_sample(a, s, p, sm)) {
        h = hash_slot(keys, (unsigned)brick_of(a, sm));
        rank = atomicAdd(&cnt[h], 1u);
    }
    __syncthreads();
    for (int q = threadIdx.x; q < kHash; q += blockDim.x)
        if (keys[q]) base[q] = a.off[keys[q] - 1] + atomicAdd(&a.cursor[keys[q] - 1], cnt[q]);
    __syncthreads();
    if (h >= 0) a.rec[base[h] + rank] = ((unsigned long long)s << 32) | (unsigned)p;
}

template <bool Q>
__global__ __launch_bounds__(256) void k_bp_fill(BpArgs a, long long S_host) {
    __shared__ unsigned keys[kHash], cnt[kHash], base[kHash];
    if constexpr (!Q) {
        bp_fill_one(a, blockIdx.x, blockIdx.y, keys, cnt, base);
        return;
    }
    const long long S = bp_slices(a, S_host);
    for (long long sl = blockIdx.x; sl < S; sl += gridDim.x) {  // slices strided over the grid
        bp_fill_one(a, sl, (int)blockIdx.y, keys, cnt, base);
        __syncthreads();  // the hash table of the next slice
    }
}

// One chunk of a brick's records without floating-point atomics: the records are counting-sorted by
// base voxel in LDS, each thread sums the 8 corner contributions of the records of its base voxels in
// registers, then the corners are added to the box in 8 passes, one corner offset per pass, in which
// distinct base voxels reach distinct box voxels.
DEV void bp_brick_one(const BpArgs& a, int nb, unsigned ck) {
    constexpr int kVox = kBrick * kBrick * kBrick;  // base voxels of a brick
    __shared__ float box[3][kBox3];
    __shared__ unsigned vcnt[kVox], voff[kVox + 1];
    __shared__ unsigned short order[kChunk], vox[kChunk];
    __shared__ float rec_f[6][kChunk];  // fx, fy, fz, value re, value im (mirror applied), weight
    const int tid = threadIdx.x;
    // brick b with coff[b] <= ck < coff[b + 1] (bricks without records hold no chunk)
    int lo = 0, hi = nb - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (a.coff[mid] <= ck) lo = mid;
        else hi = mid - 1;
    }
    const int b = lo;
    const unsigned r0 = a.off[b] + (ck - a.coff[b]) * kChunk, r1 = min(a.off[b + 1], r0 + kChunk);
    if (r0 >= r1) return;
    const int n = (int)(r1 - r0);
    const int k = b / a.nb_class;
    const int rem = b - k * a.nb_class;
    const int bx = rem % a.nbx, by = (rem / a.nbx) % a.nby, bz = rem / (a.nbx * a.nby);
    for (int v = tid; v < 3 * kBox3; v += blockDim.x) (&box[0][0])[v] = 0.f;
    for (int v = tid; v < kVox; v += blockDim.x) vcnt[v] = 0u;
    __syncthreads();
    // records -> LDS (geometry and values), base voxel histogram
    for (int i = tid; i < n; i += blockDim.x) {
        const unsigned long long rc = a.rec[r0 + i];
        const long long s = (long long)(rc >> 32);
        const int p = (int)(rc & 0xffffffffu);
        Sample sm;
        bp_sample(a, s, p, sm);  // true: the record was made from the same test
        const Coords& c = sm.c;
        const float2 v = a.slice_val[(size_t)s * a.Pf + p];
        const int lx = c.x0 - bx * kBrick, ly = c.y0 - by * kBrick, lz = c.z0 - bz * kBrick;
        const int lv = (lz * kBrick + ly) * kBrick + lx;
        vox[i] = (unsigned short)lv;
        rec_f[0][i] = c.fx;
        rec_f[1][i] = c.fy;
        rec_f[2][i] = c.fz;
        rec_f[3][i] = v.x;
        rec_f[4][i] = c.neg ? -v.y : v.y;
        rec_f[5][i] = sm.wgt;
        atomicAdd(&vcnt[lv], 1u);
    }
    __syncthreads();
    if (tid < 64) {  // exclusive scan of the 512 counts by one wave, 8 per lane
        unsigned loc[8], sum = 0;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            loc[q] = vcnt[tid * 8 + q];
            sum += loc[q];
        }
        unsigned incl = sum;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const unsigned o = __shfl_up(incl, d, 64);
            if (tid >= d) incl += o;
        }
        unsigned run = incl - sum;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            voff[tid * 8 + q] = run;
            vcnt[tid * 8 + q] = run;  // becomes the scatter cursor
            run += loc[q];
        }
        if (tid == 63) voff[kVox] = run;
    }
    __syncthreads();
    for (int i = tid; i < n; i += blockDim.x) order[atomicAdd(&vcnt[vox[i]], 1u)] = (unsigned short)i;
    __syncthreads();
    // per base voxel (this thread: tid and tid + 256): sums of the 8 corner contributions
    float acc[2][8][3];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int lv = tid + 256 * h;
#pragma unroll
        for (int cc = 0; cc < 8; ++cc) acc[h][cc][0] = acc[h][cc][1] = acc[h][cc][2] = 0.f;
        for (unsigned j = voff[lv]; j < voff[lv + 1]; ++j) {
            const int i = order[j];
            const float fx = rec_f[0][i], fy = rec_f[1][i], fz = rec_f[2][i];
            const float vr = rec_f[3][i], vi = rec_f[4][i], w = rec_f[5][i];
            const float mfx = 1.f - fx, mfy = 1.f - fy, mfz = 1.f - fz;
            const float dd[8] = {mfz * mfy * mfx, mfz * mfy * fx, mfz * fy * mfx, mfz * fy * fx,
                                 fz * mfy * mfx, fz * mfy * fx, fz * fy * mfx, fz * fy * fx};
#pragma unroll
            for (int cc = 0; cc < 8; ++cc) {
                acc[h][cc][0] += dd[cc] * vr;
                acc[h][cc][1] += dd[cc] * vi;
                acc[h][cc][2] += dd[cc] * w;
            }
        }
    }
#pragma unroll
    for (int cc = 0; cc < 8; ++cc) {
        const int dz = cc >> 2, dy = (cc >> 1) & 1, dx = cc & 1;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int lv = tid + 256 * h;
            const int lx = lv & 7, ly = (lv >> 3) & 7, lz = lv >> 6;
            const int idx = ((lz + dz) * kBox + (ly + dy)) * kBox + (lx + dx);
            box[0][idx] += acc[h][cc][0];
            box[1][idx] += acc[h][cc][1];
            box[2][idx] += acc[h][cc][2];
        }
        __syncthreads();
    }
    float* gd = a.bp_data + a.grid_stride * 2 * k;
    float* gw = a.bp_weight + a.grid_stride * k;
    for (int v = tid; v < kBox3; v += blockDim.x) {
        const float d0 = box[0][v], d1 = box[1][v], w = box[2][v];
        if (d0 == 0.f && d1 == 0.f && w == 0.f) continue;
        const int lx = v % kBox, ly = (v / kBox) % kBox, lz = v / (kBox * kBox);
        const int gx = bx * kBrick + lx, gy = by * kBrick + ly, gz = bz * kBrick + lz;
        if (gx >= a.padF || gy >= a.pad || gz >= a.pad) continue;  // nothing lands there (zero)
        const size_t idx = ((size_t)gz * a.pad + gy) * a.padF + gx;
        if (d0 != 0.f) atomicAdd(&gd[2 * idx], d0);
        if (d1 != 0.f) atomicAdd(&gd[2 * idx + 1], d1);
        if (w != 0.f) atomicAdd(&gw[idx], w);
    }
}

// One block per chunk (the grid sized for the most chunks there can be), or (device-driven pools) a
// fixed grid striding over the chunks, up to the total the scan left in coff[nb].
template <bool Q>
__global__ __launch_bounds__(256) void k_bp_brick(BpArgs a, int nb) {
    if constexpr (!Q) {
        if (blockIdx.x < a.coff[nb]) bp_brick_one(a, nb, blockIdx.x);
        return;
    }
    if (*a.abort) return;
    const unsigned nck = a.coff[nb];
    for (unsigned ck = blockIdx.x; ck < nck; ck += gridDim.x) {  // chunks strided over the grid
        bp_brick_one(a, nb, ck);
        __syncthreads();  // the box of the next chunk
    }
}

}  // namespace

hipError_t launch_backproject_binned(hipStream_t s, long long S, const BpArgs& a, int persist_blocks) {
    const int nb = a.nb_class * a.K;
    hipError_t e;
    if (a.ntask_dev) {  // device-driven: fixed grids striding over the device counts
        if (!a.abort || persist_blocks <= 0) return hipErrorInvalidValue;
        if ((unsigned long long)a.slice_cap * a.Pf >= (1ULL << 32)) return hipErrorInvalidValue;
        if ((e = launch_fill32(s, a.count, 0u, (size_t)nb)) != hipSuccess) return e;
        const int nPB = (a.Pf + 255) / 256;
        const dim3 g2((unsigned)std::max(1, persist_blocks / nPB), (unsigned)nPB);
        k_bp_count<true><<<g2, 256, 0, s>>>(a, 0);
        k_bp_scan<<<1, 1024, 0, s>>>(a.count, nb, a.off, a.coff, a.cursor);
        k_bp_fill<true><<<g2, 256, 0, s>>>(a, 0);
        k_bp_brick<true><<<persist_blocks, 256, 0, s>>>(a, nb);
        return hipGetLastError();
    }
    if (S <= 0) return hipSuccess;
    if (S > 0x7fffffffLL || (unsigned long long)S * a.Pf >= (1ULL << 32)) return hipErrorInvalidValue;
    if ((e = launch_fill32(s, a.count, 0u, (size_t)nb)) != hipSuccess) return e;
    const dim3 grid((unsigned)S, (a.Pf + 255) / 256);
    k_bp_count<false><<<grid, 256, 0, s>>>(a, S);
    k_bp_scan<<<1, 1024, 0, s>>>(a.count, nb, a.off, a.coff, a.cursor);
    k_bp_fill<false><<<grid, 256, 0, s>>>(a, S);
    // at most one chunk per brick plus one per kChunk records
    const unsigned long long chunks = (unsigned long long)nb + ((unsigned long long)S * a.Pf + kChunk - 1) / kChunk;
    k_bp_brick<false><<<(unsigned)chunks, 256, 0, s>>>(a, nb);
    return hipGetLastError();
}

}  // namespace rome
