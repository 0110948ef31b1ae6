// Device helpers shared by the E-step kernels (kernels.hip, weights.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rome {

#define DEV __device__ __forceinline__

// Monotone map float -> int (signed order), for atomicMin on floats.
DEV int float_order_key(float f) {
    const int i = __float_as_int(f);
    return i >= 0 ? i : i ^ 0x7fffffff;
}
DEV float float_from_order_key(int k) { return __int_as_float(k >= 0 ? k : k ^ 0x7fffffff); }

// ------------------------------------------------------------------ helpers
// Separately rounded FP64 multiply / add: the reference's host arithmetic has no FMA,
// and table indices / floor() decisions depend on the last bit.
DEV double dmul(double a, double b) {
#pragma clang fp contract(off)
    return a * b;
}
DEV double dadd(double a, double b) {
#pragma clang fp contract(off)
    return a + b;
}

DEV float linf(float a, float l, float h) { return l + (h - l) * a; }

struct Coords {
    int x0, y0, z0;
    float fx, fy, fz;
    bool neg;
};

// Logical -> grid coordinates exactly as MyProjector::projectOneTile computes them
// (src/reconstruct.cpp:517-560), FP64 without contraction so that floor() agrees.
DEV Coords map_coords(const double* Ai, int x, int y, int c, bool is3d) {
#pragma clang fp contract(off)
    double xp = dadd(dmul(Ai[0], (double)x), dmul(Ai[1], (double)y));
    double yp = dadd(dmul(Ai[2], (double)x), dmul(Ai[3], (double)y));
    double zp = is3d ? dadd(dmul(Ai[4], (double)x), dmul(Ai[5], (double)y)) : 0.0;
    Coords r;
    r.neg = xp < 0;
    if (r.neg) {
        xp = -xp;
        yp = -yp;
        zp = -zp;
    }
    double fx0 = floor(xp), fy0 = floor(yp), fz0 = floor(zp);
    r.x0 = (int)fx0;
    r.y0 = (int)fy0 + c;
    r.z0 = (int)fz0 + c;
    r.fx = (float)(xp - fx0);
    r.fy = (float)(yp - fy0);
    r.fz = (float)(zp - fz0);
    return r;
}

// Trilinear (3D) / bilinear (2D) gather of the half-transform with Friedel mirror.
DEV float2 gather(const float2* __restrict__ g, int pad, int padF, const Coords& q, bool is3d) {
    float2 out;
    if (is3d) {
        const float2* p00 = g + ((size_t)q.z0 * pad + q.y0) * padF + q.x0;
        const float2* p01 = p00 + padF;                  // (z0, y1)
        const float2* p10 = p00 + (size_t)pad * padF;    // (z1, y0)
        const float2* p11 = p10 + padF;                  // (z1, y1)
        float2 d000 = p00[0], d001 = p00[1], d010 = p01[0], d011 = p01[1];
        float2 d100 = p10[0], d101 = p10[1], d110 = p11[0], d111 = p11[1];
        float dx00r = linf(q.fx, d000.x, d001.x), dx00i = linf(q.fx, d000.y, d001.y);
        float dx01r = linf(q.fx, d100.x, d101.x), dx01i = linf(q.fx, d100.y, d101.y);
        float dx10r = linf(q.fx, d010.x, d011.x), dx10i = linf(q.fx, d010.y, d011.y);
        float dx11r = linf(q.fx, d110.x, d111.x), dx11i = linf(q.fx, d110.y, d111.y);
        float dxy0r = linf(q.fy, dx00r, dx10r), dxy0i = linf(q.fy, dx00i, dx10i);
        float dxy1r = linf(q.fy, dx01r, dx11r), dxy1i = linf(q.fy, dx01i, dx11i);
        out.x = linf(q.fz, dxy0r, dxy1r);
        out.y = linf(q.fz, dxy0i, dxy1i);
    } else {
        const float2* p0 = g + (size_t)q.y0 * padF + q.x0;
        const float2* p1 = p0 + padF;
        float2 d00 = p0[0], d01 = p0[1], d10 = p1[0], d11 = p1[1];
        float dx0r = linf(q.fx, d00.x, d01.x), dx0i = linf(q.fx, d00.y, d01.y);
        float dx1r = linf(q.fx, d10.x, d11.x), dx1i = linf(q.fx, d10.y, d11.y);
        out.x = linf(q.fy, dx0r, dx1r);
        out.y = linf(q.fy, dx0i, dx1i);
    }
    if (q.neg) out.y = -out.y;
    return out;
}

// The same trilinear interpolation from the neighbourhood grid: record i holds the eight values
// the flat gather reads for base voxel i (x0, x0+1 of rows (y0,z0), (y1,z0), (y0,z1), (y1,z1)),
// so one sample reads 64 contiguous bytes instead of touching four 128-byte lines.
// Neighbourhood grid layout: 8 x 8 x 8 bricks of records (x fastest inside a brick, bricks x fastest),
// nbx = ceil(padF / 8), nby = ceil(pad / 8) bricks along x and y.  Records of nearby base voxels share a
// 32 KB brick (and a 2 MB page), where the plain z-major order puts each z plane of records 2 MB apart and
// a slice through the grid touches a new page at almost every sample.
DEV size_t cube_index(int x, int y, int z, int nbx, int nby) {
    return ((((size_t)(z >> 3) * nby + (y >> 3)) * nbx + (x >> 3)) << 9) + ((z & 7) << 6) + ((y & 7) << 3) + (x & 7);
}

// record of base voxel q in the neighbourhood grid
DEV const float4* cube_record(const float4* __restrict__ cube, int nbx, int nby, const Coords& q) {
    return cube + 4 * cube_index(q.x0, q.y0, q.z0, nbx, nby);
}

// trilinear interpolation from a record: a = (v000, v001), b = (v010, v011), c = (v100, v101),
// d = (v110, v111) (value index z y x), the same arithmetic as gather
DEV float2 cube_interp(const float4& a, const float4& b, const float4& c, const float4& d, const Coords& q) {
    float dx00r = linf(q.fx, a.x, a.z), dx00i = linf(q.fx, a.y, a.w);
    float dx01r = linf(q.fx, c.x, c.z), dx01i = linf(q.fx, c.y, c.w);
    float dx10r = linf(q.fx, b.x, b.z), dx10i = linf(q.fx, b.y, b.w);
    float dx11r = linf(q.fx, d.x, d.z), dx11i = linf(q.fx, d.y, d.w);
    float dxy0r = linf(q.fy, dx00r, dx10r), dxy0i = linf(q.fy, dx00i, dx10i);
    float dxy1r = linf(q.fy, dx01r, dx11r), dxy1i = linf(q.fy, dx01i, dx11i);
    float2 out;
    out.x = linf(q.fz, dxy0r, dxy1r);
    out.y = linf(q.fz, dxy0i, dxy1i);
    if (q.neg) out.y = -out.y;
    return out;
}


DEV int px_x(int xy) { return xy & 0xffff; }
DEV int px_y(int xy) { return (int)(short)(xy >> 16); }

DEV float2 cmul(float2 a, float2 b) { return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }

template <typename T>
DEV T wave_sum(T v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
DEV double wave_min(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}

// block-wide reduction over blockDim.x threads (multiple of 64), result broadcast
template <typename T, typename Op>
DEV T block_reduce(T v, Op op, T* scratch) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (int o = 32; o > 0; o >>= 1) v = op(v, __shfl_xor(v, o, 64));
    __syncthreads();
    if (lane == 0) scratch[wid] = v;
    __syncthreads();
    T r = scratch[0];
    for (int w = 1; w < nw; ++w) r = op(r, scratch[w]);
    __syncthreads();
    return r;
}

DEV double pdf_offset(double tx, double ty, double ox, double oy, double px, double py, double s2off) {
    const double offx = tx + ox, offy = ty + oy;
    const double td = (offx - px) * (offx - px) + (offy - py) * (offy - py);
    if (s2off < 0.0001) return td > 0. ? 0. : 1.;
    return exp(td / (-2. * s2off)) / (2. * 3.14159265358979323846 * s2off);
}

}  // namespace rome
