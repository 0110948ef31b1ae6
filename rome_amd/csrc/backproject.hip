// Binned 3D backprojection of the significant slices: the scatter of
// MyBackProjector::backproject (src/reconstruct.cpp:977-1164) reorganised so that the trilinear
// corner updates land in LDS instead of the HBM grids.
//
//   k_post  : per slice s = task * NOR + o, the weighted sum of the shifted unmasked images
//             V_s(p) = sum_t w_t/sumw * Minv CTF X_nomask(p) S_t(p) (float2 [S][Pf]) and
//             W_s = sum_t w_t/sumw (double [S]; 0: the slice has no significant sample).
//   count   : every (slice, pixel) with a positive weight, inside r_max -> trilinear base voxel ->
//             brick (8x8x8 base voxels of one class grid); counts per brick.
//   scan    : exclusive prefix sum of the counts (one block).
//   fill    : records (slice << 32 | pixel) grouped by brick.
//   brick   : one block per brick: the records' 8 corners accumulate in a 9x9x9 LDS box (base voxels
//             plus the +1 corner halo), then each non-zero voxel goes to the grids with one atomic
//             per component.  Corner weights and values are computed exactly as the direct scatter
//             computes them; only the order of the FP32 additions differs.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_util.h"
#include "kernels.h"

namespace rome {

namespace {

constexpr int kBrick = 8, kBox = kBrick + 1, kBox3 = kBox * kBox * kBox;

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

__global__ __launch_bounds__(256) void k_bp_count(BpArgs a) {
    const long long s = blockIdx.x;
    const int p = blockIdx.y * blockDim.x + threadIdx.x;
    if (p >= a.Pf) return;
    Sample sm;
    if (!bp_sample(a, s, p, sm)) return;
    atomicAdd(&a.count[brick_of(a, sm)], 1u);
}

// Exclusive scan of count[0..nb) into off[0..nb], cursor cleared.  One block of 1024 threads.
__global__ __launch_bounds__(1024) void k_bp_scan(const unsigned* __restrict__ count, int nb,
                                                  unsigned* __restrict__ off, unsigned* __restrict__ cursor) {
    __shared__ unsigned part[1024];
    const int tid = threadIdx.x;
    const int chunk = (nb + 1023) / 1024;
    const int b0 = min(nb, tid * chunk), b1 = min(nb, b0 + chunk);
    unsigned loc = 0;
    for (int b = b0; b < b1; ++b) loc += count[b];
    part[tid] = loc;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {  // inclusive Hillis-Steele scan of the per-thread sums
        const unsigned v = tid >= d ? part[tid - d] : 0u;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    unsigned run = part[tid] - loc;
    for (int b = b0; b < b1; ++b) {
        off[b] = run;
        cursor[b] = 0u;
        run += count[b];
    }
    if (tid == 1023) off[nb] = part[1023];
}

__global__ __launch_bounds__(256) void k_bp_fill(BpArgs a) {
    const long long s = blockIdx.x;
    const int p = blockIdx.y * blockDim.x + threadIdx.x;
    if (p >= a.Pf) return;
    Sample sm;
    if (!bp_sample(a, s, p, sm)) return;
    const int b = brick_of(a, sm);
    const unsigned slot = a.off[b] + atomicAdd(&a.cursor[b], 1u);
    a.rec[slot] = ((unsigned long long)s << 32) | (unsigned)p;
}

__global__ __launch_bounds__(256) void k_bp_brick(BpArgs a) {
    __shared__ float box[3][kBox3];
    const int b = blockIdx.x, tid = threadIdx.x;
    const unsigned r0 = a.off[b], r1 = a.off[b + 1];
    if (r0 == r1) return;
    const int k = b / a.nb_class;
    const int rem = b - k * a.nb_class;
    const int bx = rem % a.nbx, by = (rem / a.nbx) % a.nby, bz = rem / (a.nbx * a.nby);
    for (int v = tid; v < 3 * kBox3; v += blockDim.x) (&box[0][0])[v] = 0.f;
    __syncthreads();
    for (unsigned i = r0 + tid; i < r1; i += blockDim.x) {
        const unsigned long long rc = a.rec[i];
        const long long s = (long long)(rc >> 32);
        const int p = (int)(rc & 0xffffffffu);
        Sample sm;
        if (!bp_sample(a, s, p, sm)) continue;  // cannot happen: the record was made from the same test
        const Coords& c = sm.c;
        const float2 v = a.slice_val[(size_t)s * a.Pf + p];
        const float vr = v.x, vi = c.neg ? -v.y : v.y;
        const float mfx = 1.f - c.fx, mfy = 1.f - c.fy, mfz = 1.f - c.fz;
        const float dd[8] = {mfz * mfy * mfx, mfz * mfy * c.fx, mfz * c.fy * mfx, mfz * c.fy * c.fx,
                             c.fz * mfy * mfx, c.fz * mfy * c.fx, c.fz * c.fy * mfx, c.fz * c.fy * c.fx};
        const int lx = c.x0 - bx * kBrick, ly = c.y0 - by * kBrick, lz = c.z0 - bz * kBrick;
#pragma unroll
        for (int cc = 0; cc < 8; ++cc) {
            const int dz = cc >> 2, dy = (cc >> 1) & 1, dx = cc & 1;
            const int idx = ((lz + dz) * kBox + (ly + dy)) * kBox + (lx + dx);
            atomicAdd(&box[0][idx], dd[cc] * vr);
            atomicAdd(&box[1][idx], dd[cc] * vi);
            atomicAdd(&box[2][idx], dd[cc] * sm.wgt);
        }
    }
    __syncthreads();
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

}  // namespace

hipError_t launch_backproject_binned(hipStream_t s, long long S, const BpArgs& a) {
    if (S <= 0) return hipSuccess;
    if (S > 0x7fffffffLL || (unsigned long long)S * a.Pf >= (1ULL << 32)) return hipErrorInvalidValue;
    const int nb = a.nb_class * a.K;
    hipError_t e = hipMemsetAsync(a.count, 0, (size_t)nb * sizeof(unsigned), s);
    if (e != hipSuccess) return e;
    const dim3 grid((unsigned)S, (a.Pf + 255) / 256);
    k_bp_count<<<grid, 256, 0, s>>>(a);
    k_bp_scan<<<1, 1024, 0, s>>>(a.count, nb, a.off, a.cursor);
    k_bp_fill<<<grid, 256, 0, s>>>(a);
    k_bp_brick<<<nb, 256, 0, s>>>(a);
    return hipGetLastError();
}

}  // namespace rome
