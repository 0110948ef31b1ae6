// CDNA4 (gfx950) kernels of the MAP E-step (ROME src/map3d_optimizer.cpp:883-2504,
// src/map2d_optimizer.cpp:802-1873) and the Fourier projector/backprojector
// (src/reconstruct.cpp:412-1164, 823-958, 1353-1499).
//
// Data layout in HBM (all per E-step batch unless noted):
//   grid    [K][pad][pad][pad/2+1] float2   projector reference (per iteration)
//   FcT     [2*Pc_pad][M_pad] float          coarse projections, K-major GEMM operand
//   WcT     [2*Pc_pad][N_pad] float          CTF*Minv weighted shifted images, N = img*T + t
//   E       [n][M][T] float                  coarse pose term  1/2*sum Minv|CF|^2 - sum Minv Re(CF conj(XS))
//   Pfine   [U][NOR][Pf] float2              fine projections of the rotations any image needs
//   D2      [blocks][NOR*NOT] double         fine squared differences, block = (task, significant t)
// diff2 = E + 1/2 (sum Minv |X|^2 + highres_Xi2): the per-image constant is added in FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdint.h>

#include "device_util.h"
#include "gemm_common.h"
#include "kernels.h"

namespace rome {

// ------------------------------------------------------------------ phase tables
// Tabulated shift phase of shiftImageInFourierTransform (src/image.h:1478-1574):
// index = (int)(|2 pi (x xs + y ys)| / (2 pi / 5000)) % 5000, xs = shift_x / (-ori).
__global__ void k_phase_table(const double* __restrict__ sx, const double* __restrict__ sy, int nshift,
                              const int* __restrict__ xy, int P, int ori, const double* __restrict__ sintab,
                              const double* __restrict__ costab, float2* __restrict__ out) {
#pragma clang fp contract(off)
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)nshift * P) return;
    const int s = (int)(idx / P), p = (int)(idx % P);
    const double xs = __ddiv_rn(sx[s], (double)(-ori));
    const double ys = __ddiv_rn(sy[s], (double)(-ori));
    float2 r;
    if (fabs(xs) < 1e-6 && fabs(ys) < 1e-6) {
        r = make_float2(1.f, 0.f);
    } else {
        const double twopi = 2.0 * 3.14159265358979323846;
        const double sampling = twopi / 5000.0;
        const double x = (double)px_x(xy[p]), y = (double)px_y(xy[p]);
        const double dotp = dmul(twopi, dadd(dmul(x, xs), dmul(y, ys)));
        const int it = ((int)__ddiv_rn(fabs(dotp), sampling)) % 5000;
        double a = costab[it], b = sintab[it];
        if (dotp < 0) b = -b;
        r = make_float2((float)a, (float)b);
    }
    out[idx] = r;
}

// ------------------------------------------------------------------ projections
// Coarse projection table (image independent): FcT[p][m] = Re, FcT[Pc_pad+p][m] = Im
// for rotation m = k*R + r.  Pixels outside r <= size/2 stay zero (projectOneTile).
__global__ void k_project_coarse(const float2* __restrict__ grids, size_t grid_stride, int pad, int padF, int c,
                                 int is3d, const double* __restrict__ Ai, int R, int Mtot, const int* __restrict__ xy,
                                 int Pc, int Pc_pad, int my_r_max, int Mpad, float* __restrict__ FcT) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    const int p = blockIdx.y;
    if (m >= Mtot) return;
    const int k = m / R, r = m % R;
    const int x = px_x(xy[p]), y = px_y(xy[p]);
    float2 v = make_float2(0.f, 0.f);
    if (x * x + y * y <= my_r_max * my_r_max) {
        Coords q = map_coords(Ai + 6 * (size_t)r, x, y, c, is3d != 0);
        v = gather(grids + grid_stride * k, pad, padF, q, is3d != 0);
    }
    FcT[(size_t)p * Mpad + m] = v.x;
    FcT[(size_t)(Pc_pad + p) * Mpad + m] = v.y;
}

// Neighbourhood grid of a 3D class grid: the record of base voxel i = the eight values the trilinear
// gather reads (flat indices i, i+1, i+padF, i+padF+1, then the same one z plane on), zero past the end
// of the class grid.  64 bytes per voxel: 8x the grid's HBM for one 64-byte read per fine-projection
// sample instead of four scattered 128-byte lines.  Records in 8 x 8 x 8 bricks (cube_index); one
// thread per record of the brick layout, blockIdx.y = class.
__global__ void k_build_cube(const float2* __restrict__ grids, size_t n, int pad, int padF, float4* __restrict__ cube,
                             size_t cube_n) {
    const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;  // record in the brick layout
    if (j >= cube_n) return;
    const int nbx = (padF + 7) / 8, nby = (pad + 7) / 8;
    const size_t b = j >> 9;
    const int x = (int)(b % nbx) * 8 + (int)(j & 7), y = (int)((b / nbx) % nby) * 8 + (int)((j >> 3) & 7),
              z = (int)(b / ((size_t)nbx * nby)) * 8 + (int)((j >> 6) & 7);
    float4* o = cube + 4 * (cube_n * blockIdx.y + j);
    if (x >= padF || y >= pad || z >= pad) {  // brick padding: never read
        o[0] = o[1] = o[2] = o[3] = make_float4(0.f, 0.f, 0.f, 0.f);
        return;
    }
    const size_t i = ((size_t)z * pad + y) * padF + x;
    const float2* g = grids + n * blockIdx.y;
    const size_t sy = (size_t)padF, sz = (size_t)pad * padF;
    auto at = [&](size_t q) { return q < n ? g[q] : make_float2(0.f, 0.f); };
    const float2 v0 = at(i), v1 = at(i + 1), v2 = at(i + sy), v3 = at(i + sy + 1);
    const float2 v4 = at(i + sz), v5 = at(i + sz + 1), v6 = at(i + sz + sy), v7 = at(i + sz + sy + 1);
    o[0] = make_float4(v0.x, v0.y, v1.x, v1.y);
    o[1] = make_float4(v2.x, v2.y, v3.x, v3.y);
    o[2] = make_float4(v4.x, v4.y, v5.x, v5.y);
    o[3] = make_float4(v6.x, v6.y, v7.x, v7.y);
}

// Fine projections for the unique rotations the batch needs: Pf[u][o][p].  3D classes read the
// neighbourhood grid (cube, same values and arithmetic as the flat gather); 2D the flat grid.
// U_dev (device-driven pools): the count of unique rotations is read from the device (at most U_host, the
// capacity of Pfine; a pool with more is aborted); the first launch covers the items of U_guess rotations,
// a fixed tail grid the rest.  A work item is (rotation u, block of 256 pixels) with all NOR oversampled
// orientations: each thread gathers NOR independent samples of its pixel (the loads of one item are in
// flight together, and the chain rotation -> matrices -> pixel is paid once per item, not per sample).
template <int NOR>
__global__ __launch_bounds__(256) void k_project_fine(const float2* __restrict__ grids, const float4* __restrict__ cube,
                                                      size_t grid_stride, int pad, int padF, int c, int is3d,
                                                      const double* __restrict__ Ai_fine, int R,
                                                      const int* __restrict__ ulist, int U_host,
                                                      const int* __restrict__ U_dev, unsigned long long* __restrict__ xq,
                                                      long long jcap, const int* __restrict__ xy, int Pf, int r_max,
                                                      float2* __restrict__ Pfine) {
    // Workgroups go to the 8 XCDs round-robin by launch index; each XCD takes one contiguous range of
    // items w = u * nPB + pixel block, so the pixel blocks of a rotation share one L2.
    const int U = U_dev ? min(*U_dev, U_host) : U_host;  // device-driven: U_host is the buffer's capacity
    const int nPB = (Pf + 255) / 256;
    const long long G = (long long)U * nPB, per = (G + 7) / 8;
    const size_t cube_stride = cube ? cube_records(pad, padF) : 0;
    auto item = [&](long long w) {
        if (w >= G) return;
        const int u = (int)(w / nPB);
        const int p = (int)(w % nPB) * 256 + threadIdx.x;
        if (p >= Pf) return;
        const int m = __builtin_amdgcn_readfirstlane(ulist[u]);  // block-uniform: matrices by scalar loads
        const int k = m / R, r = m % R;
        const int x = px_x(xy[p]), y = px_y(xy[p]);
        float2* out = Pfine + (size_t)u * NOR * Pf + p;
        if (x * x + y * y > r_max * r_max) {
#pragma unroll
            for (int o = 0; o < NOR; ++o) out[(size_t)o * Pf] = make_float2(0.f, 0.f);
            return;
        }
        const double* A = Ai_fine + 6 * (size_t)r * NOR;
        Coords q[NOR];
#pragma unroll
        for (int o = 0; o < NOR; ++o) q[o] = map_coords(A + 6 * o, x, y, c, is3d != 0);
        float2 v[NOR];
        if (cube) {  // every record load of the item issued before the first is used
            const float4* cb = cube + 4 * cube_stride * k;
            const int nbx = (padF + 7) / 8, nby = (pad + 7) / 8;
            float4 rec[NOR][4];
#pragma unroll
            for (int o = 0; o < NOR; ++o) {
                const float4* rp = cube_record(cb, nbx, nby, q[o]);
#pragma unroll
                for (int e = 0; e < 4; ++e) rec[o][e] = rp[e];
            }
            __builtin_amdgcn_sched_barrier(0);  // keep the loads ahead of their first use
#pragma unroll
            for (int o = 0; o < NOR; ++o) v[o] = cube_interp(rec[o][0], rec[o][1], rec[o][2], rec[o][3], q[o]);
        } else {
#pragma unroll
            for (int o = 0; o < NOR; ++o) v[o] = gather(grids + grid_stride * k, pad, padF, q[o], false);
        }
#pragma unroll
        for (int o = 0; o < NOR; ++o) out[(size_t)o * Pf] = v[o];
    };
    if (!xq) {  // one work item per block: item j = L >> 3 of range L & 7 (a device count: j < jcap)
        const long long L = (long long)blockIdx.y * gridDim.x + blockIdx.x;  // gridDim.x is a multiple of 8
        if ((L >> 3) < per && (jcap == 0 || (L >> 3) < jcap)) item((L & 7) * per + (L >> 3));
        return;
    }
    // device-driven tail: a fixed grid (a multiple of 8 blocks); the blocks of XCD x take the items j >= jcap
    // of range x in order from that range's counter xq[x]
    __shared__ long long s_i;
    const int xcd = blockIdx.x & 7;
    for (;;) {
        if (threadIdx.x == 0) s_i = jcap + (long long)atomicAdd(xq + xcd, 1ull);
        __syncthreads();
        const long long i = s_i;
        __syncthreads();
        if (i >= per) break;
        item(xcd * per + i);
    }
}

// Standalone projection of arbitrary rotations into full half-planes (parity tests and
// synthetic data): out[i][n] for rotation i, n over size*(size/2+1).
__global__ void k_project_plane(const float2* __restrict__ grid, int pad, int padF, int c, int is3d,
                                const double* __restrict__ Ai, int size, int my_r_max, float2* __restrict__ out) {
    const int Fsize = size / 2 + 1;
    const int n = blockIdx.y * blockDim.x + threadIdx.x;
    const int i = blockIdx.x;
    if (n >= size * Fsize) return;
    const int row = n / Fsize, x = n % Fsize;
    int y;
    float2 v = make_float2(0.f, 0.f);
    bool ok = true;
    if (row <= my_r_max) y = row;
    else if (row >= size - my_r_max) y = row - size;
    else ok = false;
    if (ok && x <= my_r_max && x * x + y * y <= my_r_max * my_r_max) {
        Coords q = map_coords(Ai + 6 * (size_t)i, x, y, c, is3d != 0);
        v = gather(grid, pad, padF, q, is3d != 0);
    }
    out[(size_t)i * size * Fsize + n] = v;
}

// ------------------------------------------------------------------ image preparation (K6)
// One block per image.  Windows the masked image to the coarse size, builds the
// CTF-Minv weighted shifted images (GEMM operand), the coarse norm weights and the
// FP64 image norm; gathers the fine pixel set (masked, unmasked*Minv*CTF for the
// backprojector).  src/exp_model.cpp:935-1103, src/map3d_optimizer.cpp:2246-2308.
__global__ void k_prep_images(PrepArgs a) {
    const int img = blockIdx.x;
    const int g = a.group[img];
    const double* s2 = a.sigma2 + (size_t)g * a.ori_Fsize;
    const float2* fm = a.fmask + (size_t)img * a.F2f;
    const float2* fn = a.fnomask + (size_t)img * a.F2f;
    const float* fc = a.fctf + (size_t)img * a.F2f;
    __shared__ double red[16];
    // coarse
    double x2 = 0.0;
    for (int p = threadIdx.x; p < a.Pc; p += blockDim.x) {
        const int nf = a.c_nfine[p];
        const float2 X = fm[nf];
        const double minv = 1.0 / (a.sigma2_fudge * s2[a.c_ires[p]]);
        const float ctf = a.ctf_in_diff ? fc[nf] : 1.f;
        const float mc = (float)minv * ctf;
        a.MC2c[(size_t)p * a.mc2_ld + img] = mc * ctf;
        x2 += minv * ((double)X.x * X.x + (double)X.y * X.y);
        if (!a.WcT) continue;  // the mixed-precision arm writes its own operand
        for (int t = 0; t < a.T; ++t) {
            const float2 S = a.Sc[(size_t)t * a.Pc + p];
            const float2 XS = cmul(X, S);
            const size_t col = (size_t)img * a.T + t;
            a.WcT[(size_t)p * a.Npad + col] = mc * XS.x;
            a.WcT[(size_t)(a.Pc_pad + p) * a.Npad + col] = mc * XS.y;
        }
    }
    double s = block_reduce(x2, [](double u, double v) { return u + v; }, red);
    if (threadIdx.x == 0) a.C2c[img] = 0.5 * (s + a.hxi2[img]);
    if (!a.have_fine) return;
    // fine
    double x2f = 0.0;
    for (int p = threadIdx.x; p < a.Pf; p += blockDim.x) {
        const int nf = a.f_nfine[p];
        const int ires = a.f_ires[p];
        const float2 X = fm[nf];
        const double minv = ires > 0 ? 1.0 / (a.sigma2_fudge * s2[ires]) : 0.0;
        const float ctf_d = a.ctf_in_diff ? fc[nf] : 1.f;
        const size_t o = (size_t)img * a.Pf + p;
        a.Xf[o] = X;
        a.Minvf[o] = (float)minv;
        a.Ctff[o] = ctf_d;
        x2f += minv * ((double)X.x * X.x + (double)X.y * X.y);
        // backprojection inputs: Minv with the origin restored, do_map / do_ctf rules
        double mbp = (ires > 0) ? minv : (ires == 0 ? 1.0 / (a.sigma2_fudge * s2[0]) : 0.0);
        if (!a.do_map) mbp = 1.0;
        const float ctf_b = a.do_ctf ? fc[nf] : 1.f;
        const float mc = (float)mbp * ctf_b;
        const float2 N = fn[nf];
        a.NMf[o] = make_float2(N.x * mc, N.y * mc);
        a.MC2bp[o] = mc * ctf_b;
    }
    s = block_reduce(x2f, [](double u, double v) { return u + v; }, red);
    if (threadIdx.x == 0) a.C2f[img] = 0.5 * (s + a.hxi2[img]);
}

// Squared magnitudes of the coarse projections, the A-term's K-major operand: Qc[p][m] = Re^2 + Im^2 of
// FcT for p < Pc (zero for p in [Pc, Kq)), once per reference change.
__global__ void k_fct_norm(const float* __restrict__ FcT, int Pc, int Pc_pad, int Mpad, int Kq, float* __restrict__ Qc) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)Kq * Mpad) return;
    const int p = (int)(i / Mpad);
    const size_t m = i - (size_t)p * Mpad;
    float q = 0.f;
    if (p < Pc) {
        const float fr = FcT[(size_t)p * Mpad + m], fi = FcT[(size_t)(Pc_pad + p) * Mpad + m];
        q = fr * fr + fi * fi;
    }
    Qc[i] = q;
}

// A-term as an FP32 MFMA GEMM: At[img][m] = 0.5 * sum_p Qc[p][m] * MC2cT[p][img] (rows padded to Mpad,
// zero beyond Mtot).  128 rotations x 128 images per block of 4 waves (2 x 2, each 2 x 2 blocks of
// v_mfma_f32_32x32x2f32), BK = 16, two LDS stages: the rotation table is read once per 128 images.
__global__ __launch_bounds__(256) void k_coarse_aterm(const float* __restrict__ Qc, int Kq, int Mtot, int Mpad,
                                                      const float* __restrict__ MC2cT, int ldn, int n,
                                                      float* __restrict__ At) {
    constexpr int BM = 128, BN = 128, BK = 16;
    __shared__ __attribute__((aligned(16))) float As[2][BK][BM];
    __shared__ __attribute__((aligned(16))) float Bs[2][BK][BN];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
    const int lr = tid >> 5, lc = (tid & 31) * 4;  // rows lr and lr + 8 of a 16 x 128 chunk
    f16v acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;
    const float* Ag = Qc + (size_t)lr * Mpad + m0 + lc;
    const float* Bg = MC2cT + (size_t)lr * ldn + n0 + lc;
    const size_t a8 = (size_t)8 * Mpad, b8 = (size_t)8 * ldn;
    float4 ra0 = *reinterpret_cast<const float4*>(Ag), ra1 = *reinterpret_cast<const float4*>(Ag + a8);
    float4 rb0 = *reinterpret_cast<const float4*>(Bg), rb1 = *reinterpret_cast<const float4*>(Bg + b8);
    *reinterpret_cast<float4*>(&As[0][lr][lc]) = ra0;
    *reinterpret_cast<float4*>(&As[0][lr + 8][lc]) = ra1;
    *reinterpret_cast<float4*>(&Bs[0][lr][lc]) = rb0;
    *reinterpret_cast<float4*>(&Bs[0][lr + 8][lc]) = rb1;
    __syncthreads();
    const int nk = Kq / BK;
    for (int kc = 0; kc < nk; ++kc) {
        const int cur = kc & 1;
        if (kc + 1 < nk) {
            const float* an = Ag + (size_t)(kc + 1) * BK * Mpad;
            const float* bn = Bg + (size_t)(kc + 1) * BK * ldn;
            ra0 = *reinterpret_cast<const float4*>(an);
            ra1 = *reinterpret_cast<const float4*>(an + a8);
            rb0 = *reinterpret_cast<const float4*>(bn);
            rb1 = *reinterpret_cast<const float4*>(bn + b8);
        }
#pragma unroll
        for (int kk = 0; kk < BK; kk += 2) {
            float av[2], bv[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) av[i] = As[cur][kk + (lane >> 5)][wm * 64 + i * 32 + (lane & 31)];
#pragma unroll
            for (int j = 0; j < 2; ++j) bv[j] = Bs[cur][kk + (lane >> 5)][wn * 64 + j * 32 + (lane & 31)];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[i], bv[j], acc[i][j], 0, 0, 0);
        }
        if (kc + 1 < nk) {
            *reinterpret_cast<float4*>(&As[cur ^ 1][lr][lc]) = ra0;
            *reinterpret_cast<float4*>(&As[cur ^ 1][lr + 8][lc]) = ra1;
            *reinterpret_cast<float4*>(&Bs[cur ^ 1][lr][lc]) = rb0;
            *reinterpret_cast<float4*>(&Bs[cur ^ 1][lr + 8][lc]) = rb1;
        }
        __syncthreads();
    }
    // lane: column img = n0 + wn*64 + j*32 + lane%32, rows m0 + wm*64 + i*32 + 8q + 4(lane/32) + u
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int img = n0 + wn * 64 + j * 32 + (lane & 31);
        if (img >= n) continue;
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int m = m0 + wm * 64 + i * 32 + 8 * q + 4 * (lane >> 5);
                float4 v;
                v.x = m < Mtot ? 0.5f * acc[i][j][4 * q] : 0.f;
                v.y = m + 1 < Mtot ? 0.5f * acc[i][j][4 * q + 1] : 0.f;
                v.z = m + 2 < Mtot ? 0.5f * acc[i][j][4 * q + 2] : 0.f;
                v.w = m + 3 < Mtot ? 0.5f * acc[i][j][4 * q + 3] : 0.f;
                *reinterpret_cast<float4*>(&At[(size_t)img * Mpad + m]) = v;
            }
    }
}

// ------------------------------------------------------------------ coarse cross term (K2), FP32 MFMA
// C[m][j] = sum_kk FcT[kk][m] * WcT[kk][j];  E[img][m][t] = At[img][m] - C[m][img*T+t].
// 256 threads, 128x128 tile, BK = 16, v_mfma_f32_32x32x2f32, LDS double buffer.
template <bool LIST>
__global__ __launch_bounds__(256) void k_coarse_gemm(GemmArgs a) {
    constexpr int BM = 128, BN = 128, BK = 16;
    constexpr int kOps = 2 * 2 * BK * BM;  // operand stages
    __shared__ __attribute__((aligned(16))) float smem[kOps];
    typedef float Stage[BK][BM];
    Stage* As = reinterpret_cast<Stage*>(smem);
    Stage* Bs = reinterpret_cast<Stage*>(smem + 2 * BK * BM);
    __shared__ int slot_min[epi_slots<BN>()];
    __shared__ int slot_vb[LIST ? epi_slots<BN>() : 1];
    __shared__ float slot_lim[LIST ? epi_slots<BN>() : 1];
    __shared__ int slot_cnt[LIST ? epi_slots<BN>() : 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    int bm, bn;
    gemm_tile(blockIdx.x, a.Mpad / BM, a.Npad / BN, a.sband, bm, bn);
    const int m0 = bm * BM, n0 = bn * BN;
    if (n0 >= a.Ncols) return;  // a tile of padding columns only (block-uniform, before any barrier)
    // global load mapping: 16 rows x 128 cols = 2048 floats = 512 float4 -> 2 per thread per operand
    const int lr = tid >> 5, lc = (tid & 31) * 4;  // rows lr and lr+8
    f16v acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;

    const int nk = a.K / BK;
    // staging in named registers (arrays of them end up in scratch)
    const float* Ag = a.A + (size_t)lr * a.Mpad + m0 + lc;
    const float* Bg = a.B + (size_t)lr * a.Npad + n0 + lc;
    const size_t a8 = (size_t)8 * a.Mpad, b8 = (size_t)8 * a.Npad;
    const size_t astep = (size_t)BK * a.Mpad, bstep = (size_t)BK * a.Npad;
    float4 ra0 = *reinterpret_cast<const float4*>(Ag), ra1 = *reinterpret_cast<const float4*>(Ag + a8);
    float4 rb0 = *reinterpret_cast<const float4*>(Bg), rb1 = *reinterpret_cast<const float4*>(Bg + b8);
    *reinterpret_cast<float4*>(&As[0][lr][lc]) = ra0;
    *reinterpret_cast<float4*>(&As[0][lr + 8][lc]) = ra1;
    *reinterpret_cast<float4*>(&Bs[0][lr][lc]) = rb0;
    *reinterpret_cast<float4*>(&Bs[0][lr + 8][lc]) = rb1;
    __syncthreads();
    for (int kc = 0; kc < nk; ++kc) {
        const int cur = kc & 1;
        if (kc + 1 < nk) {
            const float* an = Ag + (size_t)(kc + 1) * astep;
            const float* bnp = Bg + (size_t)(kc + 1) * bstep;
            ra0 = *reinterpret_cast<const float4*>(an);
            ra1 = *reinterpret_cast<const float4*>(an + a8);
            rb0 = *reinterpret_cast<const float4*>(bnp);
            rb1 = *reinterpret_cast<const float4*>(bnp + b8);
        }
#pragma unroll
        for (int kk = 0; kk < BK; kk += 2) {
            float av[2], bv[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) av[i] = As[cur][kk + (lane >> 5)][wm * 64 + i * 32 + (lane & 31)];
#pragma unroll
            for (int j = 0; j < 2; ++j) bv[j] = Bs[cur][kk + (lane >> 5)][wn * 64 + j * 32 + (lane & 31)];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[i], bv[j], acc[i][j], 0, 0, 0);
        }
        if (kc + 1 < nk) {
            *reinterpret_cast<float4*>(&As[cur ^ 1][lr][lc]) = ra0;
            *reinterpret_cast<float4*>(&As[cur ^ 1][lr + 8][lc]) = ra1;
            *reinterpret_cast<float4*>(&Bs[cur ^ 1][lr][lc]) = rb0;
            *reinterpret_cast<float4*>(&Bs[cur ^ 1][lr + 8][lc]) = rb1;
        }
        __syncthreads();
    }
    if constexpr (LIST)
        gemm_epilogue_list<false, 2, 2, BN>(acc, ContigRows{m0 + wm * 64, lane >> 5}, n0 + wn * 64, n0, lane, tid, a.At, nullptr,
                                            a.emin_key, a.Mtot, a.Mpad, a.Ncols, a.T, a.list, slot_min, slot_vb,
                                            slot_lim, slot_cnt);
    else
        gemm_epilogue<false, 2, 2, BN>(acc, m0 + wm * 64, n0 + wn * 64, n0, lane, tid, a.At, nullptr, a.E, a.valid,
                                       a.emin_key, a.Mtot, a.Mpad, a.Ncols, a.T, slot_min);
}

// ------------------------------------------------------------------ task building
// Per image: count rotations (m) with any significant translation and the number of
// significant (m, t) pairs; flag the rotations in used[].  One block per image.
__global__ void k_task_count(const uint32_t* __restrict__ sigmask, int M, int TW, int* __restrict__ ntask,
                             int* __restrict__ nblk, unsigned char* __restrict__ used) {
    __shared__ long long red[16];
    const int img = blockIdx.x;
    const uint32_t* msk = sigmask + (size_t)img * M * TW;
    long long nt = 0, nb = 0;
    for (int m = threadIdx.x; m < M; m += blockDim.x) {
        int c = 0;
        for (int w = 0; w < TW; ++w) c += __popc(msk[(size_t)m * TW + w]);
        nt += c > 0;
        nb += c;
        if (c > 0) used[m] = 1;
    }
    nt = block_reduce(nt, [](long long a, long long b) { return a + b; }, red);
    nb = block_reduce(nb, [](long long a, long long b) { return a + b; }, red);
    if (threadIdx.x == 0) {
        ntask[img] = (int)nt;
        nblk[img] = (int)nb;
    }
}

// Sparse form: sort the image's keys m*T + t ascending (bitonic, in LDS, written back in place),
// count the distinct rotations and flag them in used[].  An image with more keys than the buffer
// holds counts nothing (the caller redoes the pass with the mask).
__global__ __launch_bounds__(1024) void k_task_count_keys(int* __restrict__ keys, const int* __restrict__ cnt, int cap,
                                                          int T, int* __restrict__ ntask, int* __restrict__ nblk,
                                                          unsigned char* __restrict__ used) {
    __shared__ int sk[kSigCap];
    __shared__ int red[16];
    const int img = blockIdx.x, tid = threadIdx.x;
    const int c = cnt[img];
    if (c > cap) {
        if (tid == 0) ntask[img] = nblk[img] = 0;
        return;
    }
    int* k = keys + (size_t)img * cap;
    int np2 = 1;
    while (np2 < c) np2 <<= 1;
    for (int q = tid; q < np2; q += 1024) sk[q] = q < c ? k[q] : 0x7fffffff;
    __syncthreads();
    for (int ksz = 2; ksz <= np2; ksz <<= 1)
        for (int j = ksz >> 1; j > 0; j >>= 1) {
            for (int q = tid; q < np2; q += 1024) {
                const int ixj = q ^ j;
                if (ixj > q) {
                    const int a1 = sk[q], a2 = sk[ixj];
                    if (((q & ksz) == 0) ? (a1 > a2) : (a1 < a2)) {
                        sk[q] = a2;
                        sk[ixj] = a1;
                    }
                }
            }
            __syncthreads();
        }
    int nt = 0;
    for (int q = tid; q < c; q += 1024) {
        k[q] = sk[q];
        const int m = sk[q] / T;
        if (q == 0 || sk[q - 1] / T != m) {
            ++nt;
            used[m] = 1;
        }
    }
    nt = block_reduce(nt, [](int a, int b) { return a + b; }, red);
    if (tid == 0) {
        ntask[img] = nt;
        nblk[img] = c;
    }
}

// Emission from the sorted keys: block q of the image is key q; a task starts at each new rotation.
__global__ __launch_bounds__(1024) void k_task_emit_keys(const int* __restrict__ keys, const int* __restrict__ cnt,
                                                         int cap, int T, const long long* __restrict__ task_off,
                                                         const long long* __restrict__ blk_off,
                                                         int* __restrict__ task_img, int* __restrict__ task_m,
                                                         long long* __restrict__ task_blk0, int* __restrict__ task_nt,
                                                         int* __restrict__ blk_task, int* __restrict__ blk_t) {
    __shared__ int sc[1024];
    __shared__ int base;
    const int img = blockIdx.x, tid = threadIdx.x;
    const int c = cnt[img];
    const int* k = keys + (size_t)img * cap;
    const long long bt = task_off[img], bb = blk_off[img];
    if (tid == 0) base = 0;
    __syncthreads();
    for (int q0 = 0; q0 < c; q0 += 1024) {
        const int q = q0 + tid;
        const int key = q < c ? k[q] : 0;
        const int m = key / T;
        const int first = (q < c && (q == 0 || k[q - 1] / T != m)) ? 1 : 0;
        sc[tid] = first;
        __syncthreads();
        for (int o = 1; o < 1024; o <<= 1) {  // inclusive Hillis-Steele scan
            const int v = tid >= o ? sc[tid - o] : 0;
            __syncthreads();
            sc[tid] += v;
            __syncthreads();
        }
        if (q < c) {
            const long long tq = bt + base + sc[tid] - 1;
            blk_task[bb + q] = (int)tq;
            blk_t[bb + q] = key - m * T;
            if (first) {
                int e = q + 1;
                while (e < c && k[e] / T == m) ++e;
                task_img[tq] = img;
                task_m[tq] = m;
                task_blk0[tq] = bb + q;
                task_nt[tq] = e - q;
            }
        }
        __syncthreads();
        if (tid == 1023) base += sc[1023];
        __syncthreads();
    }
}

// Exclusive scans of the per-image counts: one 1024-thread block, each thread a contiguous run of
// images, the run totals scanned in LDS.
__global__ __launch_bounds__(1024) void k_scan_images(const int* __restrict__ ntask, const int* __restrict__ nblk,
                                                      int n, long long* __restrict__ task_off,
                                                      long long* __restrict__ blk_off) {
    __shared__ long long sa[1024], sb[1024];
    const int tid = threadIdx.x;
    const int per = (n + 1023) / 1024, i0 = min(n, tid * per), i1 = min(n, i0 + per);
    long long a = 0, b = 0;
    for (int i = i0; i < i1; ++i) {
        a += ntask[i];
        b += nblk[i];
    }
    sa[tid] = a;
    sb[tid] = b;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {  // inclusive Hillis-Steele scan of the run totals
        const long long va = tid >= o ? sa[tid - o] : 0, vb = tid >= o ? sb[tid - o] : 0;
        __syncthreads();
        sa[tid] += va;
        sb[tid] += vb;
        __syncthreads();
    }
    a = sa[tid] - a;
    b = sb[tid] - b;
    for (int i = i0; i < i1; ++i) {
        task_off[i] = a;
        blk_off[i] = b;
        a += ntask[i];
        b += nblk[i];
    }
    if (tid == 1023) {
        task_off[n] = sa[1023];
        blk_off[n] = sb[1023];
    }
}

// Ordered emission of tasks (m ascending) and blocks (t ascending) for one image per block.
__global__ __launch_bounds__(1024) void k_task_emit(const uint32_t* __restrict__ sigmask, int M, int TW, int T,
                                                    const long long* __restrict__ task_off,
                                                    const long long* __restrict__ blk_off, int* __restrict__ task_img,
                                                    int* __restrict__ task_m, long long* __restrict__ task_blk0,
                                                    int* __restrict__ task_nt, int* __restrict__ blk_task,
                                                    int* __restrict__ blk_t) {
    __shared__ int scan_t[1024], scan_b[1024];
    __shared__ long long base_t, base_b;
    const int img = blockIdx.x, tid = threadIdx.x;
    const uint32_t* msk = sigmask + (size_t)img * M * TW;
    if (tid == 0) {
        base_t = task_off[img];
        base_b = blk_off[img];
    }
    __syncthreads();
    for (int m0 = 0; m0 < M; m0 += 1024) {
        const int m = m0 + tid;
        int c = 0;
        if (m < M)
            for (int w = 0; w < TW; ++w) c += __popc(msk[(size_t)m * TW + w]);
        scan_t[tid] = c > 0;
        scan_b[tid] = c;
        __syncthreads();
        for (int o = 1; o < 1024; o <<= 1) {  // inclusive Hillis-Steele scan
            int vt = tid >= o ? scan_t[tid - o] : 0;
            int vb = tid >= o ? scan_b[tid - o] : 0;
            __syncthreads();
            scan_t[tid] += vt;
            scan_b[tid] += vb;
            __syncthreads();
        }
        if (c > 0) {
            const long long q = base_t + scan_t[tid] - 1;
            const long long b = base_b + scan_b[tid] - c;
            task_img[q] = img;
            task_m[q] = m;
            task_blk0[q] = b;
            task_nt[q] = c;
            int j = 0;
            for (int w = 0; w < TW; ++w) {
                uint32_t bits = msk[(size_t)m * TW + w];
                while (bits) {
                    const int bit = __ffs(bits) - 1;
                    bits &= bits - 1;
                    blk_task[b + j] = (int)q;
                    blk_t[b + j] = w * 32 + bit;
                    ++j;
                }
            }
        }
        __syncthreads();
        if (tid == 1023) {
            base_t += scan_t[1023];
            base_b += scan_b[1023];
        }
        __syncthreads();
    }
}

// Compaction of the rotations used by any image: ulist[u] = m, m2u[m] = u (single block, tiles of
// 4096 flags: four per thread, wave scans by shuffles, wave totals through LDS).
__global__ __launch_bounds__(1024) void k_unique_rot(const unsigned char* __restrict__ used, int M,
                                                     int* __restrict__ ulist, int* __restrict__ m2u,
                                                     int* __restrict__ ucount) {
    __shared__ int wt[16];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    int carry = 0;
    for (int t0 = 0; t0 < M; t0 += 4096) {
        const int m = t0 + 4 * tid;
        int f[4], loc = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            f[j] = (m + j < M && used[m + j]) ? 1 : 0;
            loc += f[j];
        }
        int in = loc;
        for (int o = 1; o < 64; o <<= 1) {
            const int v = __shfl_up(in, o, 64);
            if (lane >= o) in += v;
        }
        if (lane == 63) wt[wid] = in;
        __syncthreads();
        int pre = 0, tot = 0;
        for (int w = 0; w < 16; ++w) {
            if (w < wid) pre += wt[w];
            tot += wt[w];
        }
        int u = carry + pre + in - loc;
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (f[j]) {
                ulist[u] = m + j;
                m2u[m + j] = u;
                ++u;
            }
        carry += tot;
        __syncthreads();
    }
    if (tid == 0) *ucount = carry;
}

// ------------------------------------------------------------------ fine squared differences
// One block per task (image, rotation m); for each significant translation t of the
// task and each (over_rot, over_trans): diff2 = 1/2 (sum Minv|CF|^2 - 2 sum Minv Re(CF conj(XS)) + C)
// with FP32 products, per-thread FP32 partials and an FP64 block reduction.
// src/map3d_optimizer.cpp:1261-1372 (fine), 2D src/map2d_optimizer.cpp:964-1051.
template <int NOR, int NOT>
DEV void fine_task(const FineArgs& a, int q) {
    constexpr int NT = 256;
    constexpr int TG = (NOR * NOT >= 32) ? 2 : ((NOR * NOT >= 8) ? 8 : 16);  // translations per group
    constexpr int NJ = TG * NOT;
    __shared__ double red[NT / 64][NOR * NJ + NOR];
    __shared__ int ts[kMaxT];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int img = a.task_img[q], m = a.task_m[q], nt = a.task_nt[q];
    const long long b0 = a.task_blk0[q];
    const int u = a.m2u[m];
    for (int j = tid; j < nt; j += NT) ts[j] = a.blk_t[b0 + j];
    __syncthreads();
    const float2* Pf = a.Pfine + (size_t)u * NOR * a.Pf;
    const float2* X = a.Xf + (size_t)img * a.Pf;
    const float* Mi = a.Minvf + (size_t)img * a.Pf;
    const float* Ct = a.Ctff + (size_t)img * a.Pf;
    const double cimg = a.C2f[img];
    float Aacc[NOR];
#pragma unroll
    for (int o = 0; o < NOR; ++o) Aacc[o] = 0.f;
    for (int g0 = 0; g0 < nt; g0 += TG) {
        const int ng = min(TG, nt - g0);
        float B[NOR][NJ];
#pragma unroll
        for (int o = 0; o < NOR; ++o)
#pragma unroll
            for (int j = 0; j < NJ; ++j) B[o][j] = 0.f;
        const float2* St[TG];  // phase rows of this group's translations (a valid row in unused slots)
#pragma unroll
        for (int tt = 0; tt < TG; ++tt) St[tt] = a.Sf + (size_t)ts[g0 + min(tt, ng - 1)] * NOT * a.Pf;
        for (int p = tid; p < a.Pf; p += NT) {
            // every load of the pixel issued before the first use: the loop is bound by memory latency,
            // not bandwidth (pixels with Minv = 0 add zeros: no branch that the other loads wait for)
            const float minv = Mi[p];
            const float ctf = Ct[p];
            const float2 x = X[p];
            float2 f[NOR], Sv[TG][NOT];
#pragma unroll
            for (int o = 0; o < NOR; ++o) f[o] = Pf[(size_t)o * a.Pf + p];
#pragma unroll
            for (int tt = 0; tt < TG; ++tt)
#pragma unroll
                for (int ot = 0; ot < NOT; ++ot) Sv[tt][ot] = St[tt][(size_t)ot * a.Pf + p];
            __builtin_amdgcn_sched_barrier(0);
            float2 CF[NOR];
#pragma unroll
            for (int o = 0; o < NOR; ++o) {
                CF[o] = make_float2(ctf * f[o].x, ctf * f[o].y);
                if (g0 == 0) Aacc[o] += minv * (CF[o].x * CF[o].x + CF[o].y * CF[o].y);
            }
#pragma unroll
            for (int tt = 0; tt < TG; ++tt) {
                if (tt < ng) {
#pragma unroll
                    for (int ot = 0; ot < NOT; ++ot) {
                        const float2 xs = cmul(x, Sv[tt][ot]);
                        const float yr = minv * xs.x, yi = minv * xs.y;
#pragma unroll
                        for (int o = 0; o < NOR; ++o) B[o][tt * NOT + ot] += CF[o].x * yr + CF[o].y * yi;
                    }
                }
            }
        }
        // reduce: wave shuffles then across waves in FP64
#pragma unroll
        for (int o = 0; o < NOR; ++o)
#pragma unroll
            for (int j = 0; j < NJ; ++j) {
                if (j / NOT >= ng) continue;  // translation slot unused in this group (uniform)
                double v = wave_sum((double)B[o][j]);
                if (lane == 0) red[wid][o * NJ + j] = v;
            }
        if (g0 == 0) {
#pragma unroll
            for (int o = 0; o < NOR; ++o) {
                double v = wave_sum((double)Aacc[o]);
                if (lane == 0) red[wid][NOR * NJ + o] = v;
            }
        }
        __syncthreads();
        for (int idx = tid; idx < NOR * NJ; idx += NT) {
            const int o = idx / NJ, j = idx % NJ;
            const int tt = j / NOT, ot = j % NOT;
            if (tt < ng) {
                double bsum = 0., asum = 0.;
                for (int w = 0; w < NT / 64; ++w) {
                    bsum += red[w][idx];
                    asum += red[w][NOR * NJ + o];
                }
                const double diff2 = 0.5 * (asum - 2.0 * bsum) + cimg;
                a.D2[(size_t)(b0 + g0 + tt) * (NOR * NOT) + o * NOT + ot] = diff2;
            }
        }
        __syncthreads();
    }
}

// The same task with the four waves on different translations (NOR = 8): wave w takes translations
// w*TGW .. w*TGW+TGW-1 of each pass of 4*TGW, every wave runs over all pixels with all 8 orientations, so
// a pass reads the task's projections once from L2 / HBM (the other three waves find them in L1) and a
// task of up to 4*TGW translations makes one pass; each wave sums its own outputs (FP32 products and
// per-lane partials, FP64 wave reduction) and computes the norms |CF|^2 itself.
template <int NOR, int NOT, int TGW>
DEV void fine_task_waves(const FineArgs& a, int q) {
    static_assert(NOR == 8, "four waves of two orientations' outputs");
    constexpr int NJW = TGW * NOT, TGB = 4 * TGW;
    __shared__ int ts[kMaxT];
    __shared__ double asum[4][NOR];  // per wave, lane 0 (the lane that writes the outputs)
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int img = a.task_img[q], m = a.task_m[q], nt = a.task_nt[q];
    const long long b0 = a.task_blk0[q];
    const int u = a.m2u[m];
    for (int j = tid; j < nt; j += 256) ts[j] = a.blk_t[b0 + j];
    __syncthreads();
    const float2* Pf = a.Pfine + (size_t)u * NOR * a.Pf;
    const float2* X = a.Xf + (size_t)img * a.Pf;
    const float* Mi = a.Minvf + (size_t)img * a.Pf;
    const float* Ct = a.Ctff + (size_t)img * a.Pf;
    const double cimg = a.C2f[img];
    for (int g0 = 0; g0 < nt; g0 += TGB) {
        const int tw0 = g0 + wid * TGW;
        const int ngw = min(TGW, nt - tw0);  // translations of this wave in this pass (wave-uniform)
        if (ngw <= 0) break;                  // later passes have none either
        float B[NOR][NJW], Aacc[NOR];
#pragma unroll
        for (int o = 0; o < NOR; ++o) {
            Aacc[o] = 0.f;
#pragma unroll
            for (int j = 0; j < NJW; ++j) B[o][j] = 0.f;
        }
        const float2* St[TGW];
#pragma unroll
        for (int tt = 0; tt < TGW; ++tt) St[tt] = a.Sf + (size_t)ts[tw0 + min(tt, ngw - 1)] * NOT * a.Pf;
        for (int p = lane; p < a.Pf; p += 64) {
            const float minv = Mi[p];
            const float ctf = Ct[p];
            const float2 x = X[p];
            float2 f[NOR], Sv[TGW][NOT];
#pragma unroll
            for (int o = 0; o < NOR; ++o) f[o] = Pf[(size_t)o * a.Pf + p];
#pragma unroll
            for (int tt = 0; tt < TGW; ++tt)
#pragma unroll
                for (int ot = 0; ot < NOT; ++ot) Sv[tt][ot] = St[tt][(size_t)ot * a.Pf + p];
            __builtin_amdgcn_sched_barrier(0);
            float2 CF[NOR];
#pragma unroll
            for (int o = 0; o < NOR; ++o) {
                CF[o] = make_float2(ctf * f[o].x, ctf * f[o].y);
                if (g0 == 0) Aacc[o] += minv * (CF[o].x * CF[o].x + CF[o].y * CF[o].y);
            }
#pragma unroll
            for (int tt = 0; tt < TGW; ++tt) {
                if (tt < ngw) {
#pragma unroll
                    for (int ot = 0; ot < NOT; ++ot) {
                        const float2 xs = cmul(x, Sv[tt][ot]);
                        const float yr = minv * xs.x, yi = minv * xs.y;
#pragma unroll
                        for (int o = 0; o < NOR; ++o) B[o][tt * NOT + ot] += CF[o].x * yr + CF[o].y * yi;
                    }
                }
            }
        }
        if (g0 == 0) {
#pragma unroll
            for (int o = 0; o < NOR; ++o) {
                const double v = wave_sum((double)Aacc[o]);
                if (lane == 0) asum[wid][o] = v;
            }
        }
#pragma unroll
        for (int o = 0; o < NOR; ++o)
#pragma unroll
            for (int j = 0; j < NJW; ++j) {
                if (j / NOT >= ngw) continue;  // translation slot unused in this pass (wave-uniform)
                const double bsum = wave_sum((double)B[o][j]);
                if (lane == 0)
                    a.D2[(size_t)(b0 + tw0 + j / NOT) * (NOR * NOT) + o * NOT + j % NOT] =
                        0.5 * (asum[wid][o] - 2.0 * bsum) + cimg;
            }
    }
}

// One task (image, rotation) per block.  Device-driven pools: the grid is sized from the task totals of
// earlier pools (the hardware dispatcher balances tasks of 1 .. T translations as blocks retire), task
// totals on the device, blocks past the total return at once; the tasks past the grid are left to
// k_fine_diff2_queue.  The body is straight-line: a loop around it keeps more registers live, and at 3
// waves per SIMD the task body then spills.
template <int NOR, int NOT>
__global__ __launch_bounds__(256, 2) void k_fine_diff2(FineArgs a) {
    if (a.ntask_dev && (*a.abort || (long long)blockIdx.x >= *a.ntask_dev)) return;
    if constexpr (NOR == 8) fine_task_waves<NOR, NOT, 2>(a, blockIdx.x);
    else fine_task<NOR, NOT>(a, blockIdx.x);
}

// Device-driven pools, after k_fine_diff2 over the first `first` tasks: a persistent work-queue kernel
// whose blocks take the remaining tasks first, first + 1, ... one at a time from the queue head a.queue
// (zero at launch) until the device task total is reached, so a pool with more tasks than earlier ones
// still runs in one captured graph.  Every block returns at once when there is no remainder.  (2 waves
// per SIMD: the loop's live registers do not fit 3 without spills.)
template <int NOR, int NOT>
__global__ __launch_bounds__(256) void k_fine_diff2_queue(FineArgs a, long long first) {
    if (*a.abort) return;  // the pool is rerun: nothing here is used
    const long long ntask = *a.ntask_dev;
    if (ntask <= first) return;  // the same in every block
    __shared__ long long s_q;
    for (;;) {
        if (threadIdx.x == 0) s_q = first + atomicAdd(a.queue, 1);
        __syncthreads();
        const long long q = s_q;
        if (q >= ntask) break;  // the same value in every thread
        if constexpr (NOR == 8) fine_task_waves<NOR, NOT, 2>(a, (int)q);
        else fine_task<NOR, NOT>(a, (int)q);
        __syncthreads();  // s_q and the translation list are rewritten next
    }
}

// ------------------------------------------------------------------ sigma2 / pdf sums / backprojection
// One block per image.  For every significant fine sample (w >= threshold):
//   sigma2 : wsum_sigma2_noise[g][ires] += w |CF - XS|^2 (masked image)   src/map3d_optimizer.cpp:1426-1602
//   pdfs   : sumw_group, pdf_class, pdf_direction, sigma2_offset           src/map3d_optimizer.cpp:2141-2177
//   bp     : slice(rot) += w * shift(Nomask*Minv*CTF), weight += w Minv CTF^2, then trilinear
//            scatter into the class grid                                      src/map3d_optimizer.cpp:2360-2479
// BINNED: 3D, the slices go to HBM for the binned backprojection (the image's own pixel values stay in
// registers); otherwise the slices are scattered into the grids here.
template <int NOR, int NOT, int PPT, int NT, bool BINNED>
__global__ __launch_bounds__(NT) void k_post(PostArgs a) {
    constexpr int NE = NOR * NOT;
    __shared__ double wts[kMaxT * NE];
    __shared__ double red[NT / 64];
    __shared__ double shell[kMaxShell];
    __shared__ double s_pdfk[kMaxK], s_px[kMaxK], s_py[kMaxK];
    if (a.ntask_dev && (*a.abort || *a.ntask_dev * NOR <= a.slice_cap)) return;  // aborted, or binned
    const int img = blockIdx.x, tid = threadIdx.x;
    const int g = a.group[img];
    const double sumw = a.sumw[img], thr = a.thr[img], dmin = a.dmin[img];
    const long long q0 = a.task_off[img], q1 = a.task_off[img + 1];
    double wsig[PPT];
#pragma unroll
    for (int j = 0; j < PPT; ++j) wsig[j] = 0.;
    for (int s = tid; s < a.oriF; s += NT) shell[s] = 0.;
    for (int k = tid; k < a.K; k += NT) s_pdfk[k] = s_px[k] = s_py[k] = 0.;
    double t_sumw = 0., t_off = 0.;  // thread 0 accumulators
    __syncthreads();
    const float2* X = a.Xf + (size_t)img * a.Pf;
    const float* Ct = a.Ctff + (size_t)img * a.Pf;
    const float2* NM = a.NMf + (size_t)img * a.Pf;
    const float* MC2 = a.MC2bp + (size_t)img * a.Pf;
    const double ox = a.old_offx[img], oy = a.old_offy[img];
    float2 Xr[BINNED ? PPT : 1], NMr[BINNED ? PPT : 1];
    if constexpr (BINNED) {
#pragma unroll
        for (int j = 0; j < PPT; ++j) {
            const int p = tid + j * NT;
            Xr[j] = p < a.Pf ? X[p] : make_float2(0.f, 0.f);
            NMr[j] = p < a.Pf ? NM[p] : make_float2(0.f, 0.f);
        }
    }
    for (long long q = q0; q < q1; ++q) {
        const int m = a.task_m[q], nt = a.task_nt[q];
        const long long b0 = a.task_blk0[q];
        const int k = m / a.R, r = m % a.R, dir = r / a.PSI;
        const int u = a.m2u[m];
        // weights of all entries of the task
        for (int e = tid; e < nt * NE; e += NT) {
            const int tb = e / NE;
            const int t = a.blk_t[b0 + tb];
            const double d2 = a.D2[(size_t)(b0 + tb) * NE + (e % NE)];
            const double dd = d2 - dmin;
            double w = a.pdf_dir[(size_t)k * a.D + dir] *
                       pdf_offset(a.trans_x[t], a.trans_y[t], ox, oy, a.prior_x[k], a.prior_y[k], a.sigma2_offset);
            w = dd > 700. ? 0. : w * exp(-dd);
            wts[e] = w;
        }
        __syncthreads();
        const double* Ai = a.Ai_fine + 6 * ((size_t)r * NOR);
        for (int o = 0; o < NOR; ++o) {
            float2 acc[PPT];
#pragma unroll
            for (int j = 0; j < PPT; ++j) acc[j] = make_float2(0.f, 0.f);
            double wslice = 0.;
            const float2* Pf = a.Pfine + ((size_t)u * NOR + o) * a.Pf;
            float2 CF[PPT];
#pragma unroll
            for (int j = 0; j < PPT; ++j) {
                const int p = tid + j * NT;
                if (p < a.Pf) {
                    const float2 f = Pf[p];
                    const float c = Ct[p];
                    CF[j] = make_float2(c * f.x, c * f.y);
                }
            }
            for (int tb = 0; tb < nt; ++tb) {
                const int t = a.blk_t[b0 + tb];
                for (int ot = 0; ot < NOT; ++ot) {
                    const double w = wts[tb * NE + o * NOT + ot];
                    if (!(w >= thr)) continue;
                    const double wn = w / sumw;
                    wslice += wn;
                    if (tid == 0) {
                        t_sumw += wn;
                        s_pdfk[k] += wn;
                        const double sx = a.shift_x[t * NOT + ot], sy = a.shift_y[t * NOT + ot];
                        const double offx = a.prior_x[k] - ox - sx, offy = a.prior_y[k] - oy - sy;
                        t_off += wn * (offx * offx + offy * offy);
                        if (!a.is3d) {
                            s_px[k] += wn * (ox + sx);
                            s_py[k] += wn * (oy + sy);
                        }
                        atomicAdd(&a.wsum_pdf_dir[(size_t)k * a.D + dir], wn);
                    }
                    const float wf = (float)wn;
                    const float2* S = a.Sf + ((size_t)t * NOT + ot) * a.Pf;
#pragma unroll
                    for (int j = 0; j < PPT; ++j) {
                        const int p = tid + j * NT;
                        if (p < a.Pf) {
                            const float2 s = S[p];
                            const float2 xs = cmul(BINNED ? Xr[j] : X[p], s);
                            const float dr = CF[j].x - xs.x, di = CF[j].y - xs.y;
                            wsig[j] += wn * (double)(dr * dr + di * di);
                            const float2 ns = cmul(BINNED ? NMr[j] : NM[p], s);
                            acc[j].x += wf * ns.x;
                            acc[j].y += wf * ns.y;
                        }
                    }
                }
            }
            if constexpr (BINNED) {  // binned 3D backprojection: the slice goes to HBM, the grids later
                const size_t sl = (size_t)q * NOR + o;
                if (wslice > 0.) {
                    float2* V = a.slice_val + sl * a.Pf;
#pragma unroll
                    for (int j = 0; j < PPT; ++j) {
                        const int p = tid + j * NT;
                        if (p < a.Pf) V[p] = acc[j];
                    }
                }
                if (tid == 0) a.slice_w[sl] = wslice;
            } else if (wslice > 0.) {
                // trilinear scatter of the slice (MyBackProjector::backproject, src/reconstruct.cpp:977-1164)
                float* gd = a.bp_data + a.grid_stride * 2 * k;
                float* gw = a.bp_weight + a.grid_stride * k;
                const double* A6 = Ai + 6 * o;
#pragma unroll
                for (int j = 0; j < PPT; ++j) {
                    const int p = tid + j * NT;
                    if (p >= a.Pf) continue;
                    const int xy = a.f_xy[p];
                    const int x = px_x(xy), y = px_y(xy);
                    if (x * x + y * y > a.r_max * a.r_max) continue;
                    const float wgt = (float)((double)MC2[p] * wslice);
                    if (!(wgt > 0.f)) continue;
                    const Coords c = map_coords(A6, x, y, a.c, a.is3d != 0);
                    float vr = acc[j].x, vi = c.neg ? -acc[j].y : acc[j].y;
                    const float mfx = 1.f - c.fx, mfy = 1.f - c.fy, mfz = 1.f - c.fz;
                    if (a.is3d) {
                        const float dd[8] = {mfz * mfy * mfx, mfz * mfy * c.fx, mfz * c.fy * mfx, mfz * c.fy * c.fx,
                                             c.fz * mfy * mfx, c.fz * mfy * c.fx, c.fz * c.fy * mfx, c.fz * c.fy * c.fx};
#pragma unroll
                        for (int cc = 0; cc < 8; ++cc) {
                            const int dz = cc >> 2, dy = (cc >> 1) & 1, dx = cc & 1;
                            const size_t idx = ((size_t)(c.z0 + dz) * a.pad + (c.y0 + dy)) * a.padF + (c.x0 + dx);
                            atomicAdd(&gd[2 * idx], dd[cc] * vr);
                            atomicAdd(&gd[2 * idx + 1], dd[cc] * vi);
                            atomicAdd(&gw[idx], dd[cc] * wgt);
                        }
                    } else {
                        const float dd[4] = {mfy * mfx, mfy * c.fx, c.fy * mfx, c.fy * c.fx};
#pragma unroll
                        for (int cc = 0; cc < 4; ++cc) {
                            const int dy = cc >> 1, dx = cc & 1;
                            const size_t idx = (size_t)(c.y0 + dy) * a.padF + (c.x0 + dx);
                            atomicAdd(&gd[2 * idx], dd[cc] * vr);
                            atomicAdd(&gd[2 * idx + 1], dd[cc] * vi);
                            atomicAdd(&gw[idx], dd[cc] * wgt);
                        }
                    }
                }
            }
        }
        __syncthreads();
    }
    // sigma2 shells and norm-correction sum
    double tot = 0.;
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
        const int p = tid + j * NT;
        if (p < a.Pf) {
            atomicAdd(&shell[a.f_ires[p]], wsig[j]);
            tot += wsig[j];
        }
    }
    tot = block_reduce(tot, [](double x, double y) { return x + y; }, red);
    __syncthreads();
    for (int s = tid; s < a.oriF; s += NT)
        if (shell[s] != 0.) atomicAdd(&a.wsum_sigma2[(size_t)g * a.oriF + s], shell[s]);
    for (int kk = tid; kk < a.K; kk += NT) {
        if (s_pdfk[kk] != 0.) atomicAdd(&a.wsum_pdf_class[kk], s_pdfk[kk]);
        if (s_px[kk] != 0.) atomicAdd(&a.wsum_prior_x[kk], s_px[kk]);
        if (s_py[kk] != 0.) atomicAdd(&a.wsum_prior_y[kk], s_py[kk]);
    }
    if (tid == 0) {
        a.wnorm[img] = tot;
        atomicAdd(&a.wsum_sumw_group[g], t_sumw);
        atomicAdd(&a.wsum_scalars[0], t_off);
    }
}

// 3D (binned backprojection) form of k_post with one block per (task, over_rot) slice instead of one
// per image: eight times the blocks and short dependency chains.  The slice's 8 blocks go to one XCD
// back to back (workgroups are dispatched to the 8 XCDs round-robin), so the image's pixels and the
// phase rows they share come from that XCD's L2.  Per block: the weights of its entries (the same
// expression as k_post), the slice value and weight in the same sample order, sigma2 shells of its
// pixels added to the group's spectrum, and its share of the per-image / per-class sums (atomics; the
// caller zeroes wnorm first).
template <int NOR, int NOT, int PPT, int NT>
DEV void post_slice_one(const PostArgs& a, long long w, int h) {
    __shared__ double wts[kMaxT * NOT];
    __shared__ double red[NT / 64];
    __shared__ double shell[kMaxShell];
    const int tid = threadIdx.x;
    const int p0 = h * PPT * NT;  // this block's pixel range of the slice: [p0, p0 + PPT * NT)
    const long long q = w / NOR;
    const int o = (int)(w - q * NOR);
    const int img = a.task_img[q];
    const int g = a.group[img];
    const double sumw = a.sumw[img], thr = a.thr[img], dmin = a.dmin[img];
    const int m = a.task_m[q], nt = a.task_nt[q];
    const long long b0 = a.task_blk0[q];
    const int k = m / a.R, r = m % a.R, dir = r / a.PSI;
    const int u = a.m2u[m];
    const double ox = a.old_offx[img], oy = a.old_offy[img];
    for (int s = tid; s < a.oriF; s += NT) shell[s] = 0.;
    int sig = 0;
    for (int e = tid; e < nt * NOT; e += NT) {
        const int tb = e / NOT, ot = e - tb * NOT;
        const int t = a.blk_t[b0 + tb];
        const double d2 = a.D2[(size_t)(b0 + tb) * (NOR * NOT) + o * NOT + ot];
        const double dd = d2 - dmin;
        double wv = a.pdf_dir[(size_t)k * a.D + dir] *
                    pdf_offset(a.trans_x[t], a.trans_y[t], ox, oy, a.prior_x[k], a.prior_y[k], a.sigma2_offset);
        wv = dd > 700. ? 0. : wv * exp(-dd);
        wts[e] = wv;
        sig |= wv >= thr;
    }
    // a slice without a significant sample adds nothing: leave before reading its pixels
    if (!__syncthreads_or(sig)) {
        if (tid == 0 && h == 0) a.slice_w[(size_t)q * NOR + o] = 0.;
        return;
    }
    // every per-pixel array is offset to the block's range [p0, p0 + np)
    const int np = a.Pf - p0;
    const float2* X = a.Xf + (size_t)img * a.Pf + p0;
    const float* Ct = a.Ctff + (size_t)img * a.Pf + p0;
    const float2* NM = a.NMf + (size_t)img * a.Pf + p0;
    const float2* Pr = a.Pfine + ((size_t)u * NOR + o) * a.Pf + p0;
    float2 Xr[PPT], NMr[PPT], CF[PPT], acc[PPT];
    double wsig[PPT];
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
        const int p = tid + j * NT;
        const bool in = p < np;
        Xr[j] = in ? X[p] : make_float2(0.f, 0.f);
        NMr[j] = in ? NM[p] : make_float2(0.f, 0.f);
        const float2 f = in ? Pr[p] : make_float2(0.f, 0.f);
        const float c = in ? Ct[p] : 0.f;
        CF[j] = make_float2(c * f.x, c * f.y);
        acc[j] = make_float2(0.f, 0.f);
        wsig[j] = 0.;
    }
    double wslice = 0., t_off = 0.;
    for (int tb = 0; tb < nt; ++tb) {
        const int t = a.blk_t[b0 + tb];
        for (int ot = 0; ot < NOT; ++ot) {
            const double wv = wts[tb * NOT + ot];
            if (!(wv >= thr)) continue;
            const double wn = wv / sumw;
            wslice += wn;
            if (tid == 0) {
                const double sx = a.shift_x[t * NOT + ot], sy = a.shift_y[t * NOT + ot];
                const double offx = a.prior_x[k] - ox - sx, offy = a.prior_y[k] - oy - sy;
                t_off += wn * (offx * offx + offy * offy);
            }
            const float wf = (float)wn;
            const float2* S = a.Sf + ((size_t)t * NOT + ot) * a.Pf + p0;
#pragma unroll
            for (int j = 0; j < PPT; ++j) {
                const int p = tid + j * NT;
                if (p < np) {
                    const float2 s = S[p];
                    const float2 xs = cmul(Xr[j], s);
                    const float dr = CF[j].x - xs.x, di = CF[j].y - xs.y;
                    wsig[j] += wn * (double)(dr * dr + di * di);
                    const float2 ns = cmul(NMr[j], s);
                    acc[j].x += wf * ns.x;
                    acc[j].y += wf * ns.y;
                }
            }
        }
    }
    const size_t sl = (size_t)q * NOR + o;
    if (tid == 0 && h == 0) a.slice_w[sl] = wslice;
    if (!(wslice > 0.)) return;  // the same value in every thread: the whole block leaves
    float2* V = a.slice_val + sl * a.Pf + p0;
    const int* ires = a.f_ires + p0;
    double tot = 0.;
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
        const int p = tid + j * NT;
        if (p < np) {
            V[p] = acc[j];
            atomicAdd(&shell[ires[p]], wsig[j]);
            tot += wsig[j];
        }
    }
    tot = block_reduce(tot, [](double x, double y) { return x + y; }, red);
    __syncthreads();
    for (int s = tid; s < a.oriF; s += NT)
        if (shell[s] != 0.) atomicAdd(&a.wsum_sigma2[(size_t)g * a.oriF + s], shell[s]);
    if (tid == 0) {
        atomicAdd(&a.wnorm[img], tot);
        if (h == 0) {  // the slice's own sums, once per slice
            atomicAdd(&a.wsum_sumw_group[g], wslice);
            atomicAdd(&a.wsum_pdf_class[k], wslice);
            atomicAdd(&a.wsum_pdf_dir[(size_t)k * a.D + dir], wslice);
            atomicAdd(&a.wsum_scalars[0], t_off);
        }
    }
}

// Work items w = (task * NOR + o) * NH + pixel range, NH = the slice's pixel ranges, over the blocks:
// each XCD takes one contiguous range of items (per items).  TAIL = false: block L takes item j = L >> 3
// of range L & 7 (j < per); device-driven pools read the task total on the device, and the grid covers the
// first jcap = gridDim.x / 8 items of each range.  TAIL = true (device-driven only): a small fixed grid
// takes the items j >= jcap of each range from per-XCD counters xq (usually none: the first launch is
// sized from the pools before).  Two kernels rather than one with both paths: the loop raises the slice
// body's register count past the 128 of 1024-thread blocks.  A pool whose slices exceed the slice buffer
// (slice_cap) goes to the direct scatter instead, and an aborted pool adds nothing.
template <int NOR, int NOT, int PPT, int NT, bool TAIL>
__global__ __launch_bounds__(NT) void k_post_slice(PostArgs a, long long ntask, long long jcap) {
    if (a.ntask_dev) {
        if (*a.abort) return;
        ntask = *a.ntask_dev;
        if (ntask * NOR > a.slice_cap) return;
    }
    const int NH = (a.Pf + PPT * NT - 1) / (PPT * NT);
    const long long G = ntask * NOR * NH, per = (G + 7) / 8;
    const int xcd = blockIdx.x & 7;
    if constexpr (!TAIL) {
        const long long j = blockIdx.x >> 3;
        const long long w = xcd * per + j;
        if (j < per && (jcap == 0 || j < jcap) && w < G) post_slice_one<NOR, NOT, PPT, NT>(a, w / NH, (int)(w % NH));
    } else {
        __shared__ long long s_i;
        for (;;) {
            if (threadIdx.x == 0) s_i = jcap + (long long)atomicAdd(a.xq + xcd, 1ull);
            __syncthreads();
            const long long j = s_i;
            __syncthreads();
            if (j >= per) break;
            const long long w = xcd * per + j;
            if (w < G) post_slice_one<NOR, NOT, PPT, NT>(a, w / NH, (int)(w % NH));
            __syncthreads();  // LDS of the next work item
        }
    }
}

// ------------------------------------------------------------------ storeWeightedSums (src/map3d_optimizer.cpp:1933-2057)
__global__ void k_store_sums(StoreArgs a) {
    const int img = blockIdx.x * blockDim.x + threadIdx.x;
    if (img >= a.n || (a.abort && *a.abort)) return;
    const int g = a.group[img];
    const double* pw = a.power + (size_t)img * a.oriF;
    double wn = a.wnorm[img];
    for (int ires = a.curF; ires < a.oriF; ++ires) {
        atomicAdd(&a.wsum_sigma2[(size_t)g * a.oriF + ires], pw[ires]);
        wn += pw[ires];
    }
    a.wnorm[img] = wn;
    double* meta = a.meta + (size_t)img * 9;
    meta[7] = a.norm[img];
    if (a.do_norm) {
        const double old = a.norm[img] / a.avg_norm;
        meta[7] = old * sqrt(wn * 2.);
        atomicAdd(&a.wsum_scalars[3], meta[7]);
    }
    const double dll = log(a.sumw[img]) - a.dmin[img] - a.logsigma2[g];
    meta[8] = dll;
    atomicAdd(&a.wsum_scalars[1], dll);
    atomicAdd(&a.wsum_scalars[2], meta[0]);
}

// ------------------------------------------------------------------ small helpers
// Max-posterior sample of each image -> metadata row and argmax tuple (one thread per image).  An image
// without one (no weight; the caller reports it) gets zeros.
__global__ void k_meta(MetaArgs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    double* md = a.meta + (size_t)9 * i;
    int* am = a.argmax + (size_t)6 * i;
    for (int c = 0; c < 9; ++c) md[c] = 0.;
    for (int c = 0; c < 6; ++c) am[c] = 0;
    const long long key = a.amax[i];
    if (!(a.sumw[i] > 0.) || key < 0) return;
    const int NE = a.NOR * a.NOT;
    const long long b = a.img_blk[i] + key / NE;
    if (b >= a.img_blk[i + 1]) return;
    const int m = a.task_m[a.blk_task[b]];
    const int t = a.blk_t[b];
    const int o = (int)((key % NE) / a.NOT);
    const int ot = (int)(key % a.NOT);
    const int k = m / a.R, r = m % a.R;
    am[0] = k;
    am[1] = r / a.PSI;
    am[2] = r % a.PSI;
    am[3] = o;
    am[4] = t;
    am[5] = ot;
    md[0] = a.maxw[i] / a.sumw[i];
    md[1] = k + 1;
    const size_t pr = (size_t)r * a.NOR + o;
    if (a.is3d) {
        md[2] = a.pose[3 * pr];
        md[3] = a.pose[3 * pr + 1];
        md[4] = a.pose[3 * pr + 2];
    } else {
        double ps = a.pose[pr];
        if (ps > 180.) ps -= 360.;
        if (ps < -180.) ps += 360.;
        md[4] = ps;
    }
    md[5] = a.offx[i] + a.sx[(size_t)t * a.NOT + ot];
    md[6] = a.offy[i] + a.sy[(size_t)t * a.NOT + ot];
}

// ------------------------------------------------------------------ backprojection of whole images
// MyBackProjector::backproject of one image per block at its own pose (apps/rome_reconstruct.cpp:104-330):
// the transform X (current_size half-plane) shifted by the image's tabulated phase, times the CTF, and the
// weight CTF^2, scattered trilinearly (bilinearly in 2D) into one class grid; the pixel set is the fine
// one (x = 0 column for y >= 0 only, as the reference's first_x rule), radius r_max.
__global__ __launch_bounds__(256) void k_bp_images(BpImageArgs a) {
    const int img = blockIdx.x;
    const float2* X = a.X + (size_t)img * a.F2;
    const float* ctf = a.ctf + (size_t)img * a.F2;
    const float2* S = a.phase + (size_t)img * a.Pf;
    const double* A6 = a.Ai + 6 * (size_t)img;
    float* gd = a.bp_data;
    float* gw = a.bp_weight;
    for (int p = threadIdx.x; p < a.Pf; p += blockDim.x) {
        const int xy = a.f_xy[p];
        const int x = px_x(xy), y = px_y(xy);
        if (x * x + y * y > a.r_max * a.r_max) continue;
        const int nf = a.f_nfine[p];
        const float cv = ctf[nf];
        const float wgt = cv * cv;
        if (!(wgt > 0.f)) continue;
        const float2 v = cmul(X[nf], S[p]);
        const Coords c = map_coords(A6, x, y, a.c, a.is3d != 0);
        const float vr = v.x * cv, vi = (c.neg ? -v.y : v.y) * cv;
        const float mfx = 1.f - c.fx, mfy = 1.f - c.fy, mfz = 1.f - c.fz;
        if (a.is3d) {
            const float dd[8] = {mfz * mfy * mfx, mfz * mfy * c.fx, mfz * c.fy * mfx, mfz * c.fy * c.fx,
                                 c.fz * mfy * mfx, c.fz * mfy * c.fx, c.fz * c.fy * mfx, c.fz * c.fy * c.fx};
#pragma unroll
            for (int cc = 0; cc < 8; ++cc) {
                const int dz = cc >> 2, dy = (cc >> 1) & 1, dx = cc & 1;
                const size_t idx = ((size_t)(c.z0 + dz) * a.pad + (c.y0 + dy)) * a.padF + (c.x0 + dx);
                atomicAdd(&gd[2 * idx], dd[cc] * vr);
                atomicAdd(&gd[2 * idx + 1], dd[cc] * vi);
                atomicAdd(&gw[idx], dd[cc] * wgt);
            }
        } else {
            const float dd[4] = {mfy * mfx, mfy * c.fx, c.fy * mfx, c.fy * c.fx};
#pragma unroll
            for (int cc = 0; cc < 4; ++cc) {
                const int dy = cc >> 1, dx = cc & 1;
                const size_t idx = (size_t)(c.y0 + dy) * a.padF + (c.x0 + dx);
                atomicAdd(&gd[2 * idx], dd[cc] * vr);
                atomicAdd(&gd[2 * idx + 1], dd[cc] * vi);
                atomicAdd(&gw[idx], dd[cc] * wgt);
            }
        }
    }
}

// ------------------------------------------------------------------ launch wrappers
#define LAUNCH_CHECK()                                   \
    do {                                                 \
        hipError_t _e = hipGetLastError();               \
        if (_e != hipSuccess) return _e;                 \
    } while (0)

hipError_t launch_phase_table(hipStream_t s, const double* sx, const double* sy, int nshift, const int* xy, int P, int ori,
                              const double* sintab, const double* costab, float2* out) {
    const size_t tot = (size_t)nshift * P;
    if (tot == 0) return hipSuccess;
    k_phase_table<<<(unsigned)((tot + 255) / 256), 256, 0, s>>>(sx, sy, nshift, xy, P, ori, sintab, costab, out);
    LAUNCH_CHECK();
    return hipSuccess;
}

hipError_t launch_project_coarse(hipStream_t s, const float2* grids, size_t grid_stride, int pad, int padF, int c, int is3d,
                                 const double* Ai, int R, int Mtot, const int* xy, int Pc, int Pc_pad, int my_r_max,
                                 int Mpad, float* FcT) {
    dim3 grid((Mtot + 255) / 256, Pc);
    k_project_coarse<<<grid, 256, 0, s>>>(grids, grid_stride, pad, padF, c, is3d, Ai, R, Mtot, xy, Pc, Pc_pad, my_r_max,
                                          Mpad, FcT);
    LAUNCH_CHECK();
    return hipSuccess;
}

hipError_t launch_build_cube(hipStream_t s, const float2* grids, size_t grid_n, int K, int pad, int padF, float4* cube) {
    if (grid_n == 0 || K == 0) return hipSuccess;
    const size_t cn = cube_records(pad, padF);
    dim3 grid((unsigned)((cn + 255) / 256), K);
    k_build_cube<<<grid, 256, 0, s>>>(grids, grid_n, pad, padF, cube, cn);
    LAUNCH_CHECK();
    return hipSuccess;
}

hipError_t launch_project_fine(hipStream_t s, const float2* grids, const float4* cube, size_t grid_stride, int pad,
                               int padF, int c, int is3d, const double* Ai_fine, int R, int NOR, const int* ulist, int U,
                               const int* U_dev, unsigned long long* xq, int persist_blocks, int U_guess,
                               const int* xy, int Pf, int r_max, float2* Pfine) {
    if (is3d && !cube) return hipErrorInvalidValue;  // 3D projections read the neighbourhood grid
    const float4* cb = is3d ? cube : nullptr;
    const long long nPB = (Pf + 255) / 256;
    auto grid_of = [](long long items, unsigned& gx, unsigned& gy) -> bool {  // a multiple of 8 blocks
        const long long blocks = 8 * ((items + 7) / 8);
        gx = (unsigned)std::min<long long>(blocks, 32768);
        gy = (unsigned)((blocks + gx - 1) / gx);
        return gy <= 65535;
    };
    auto go = [&](dim3 grid, unsigned long long* q, long long jcap) -> hipError_t {
        switch (NOR) {
        case 8: k_project_fine<8><<<grid, 256, 0, s>>>(grids, cb, grid_stride, pad, padF, c, is3d, Ai_fine, R, ulist, U,
                                                       U_dev, q, jcap, xy, Pf, r_max, Pfine); break;
        case 2: k_project_fine<2><<<grid, 256, 0, s>>>(grids, cb, grid_stride, pad, padF, c, is3d, Ai_fine, R, ulist, U,
                                                       U_dev, q, jcap, xy, Pf, r_max, Pfine); break;
        case 1: k_project_fine<1><<<grid, 256, 0, s>>>(grids, cb, grid_stride, pad, padF, c, is3d, Ai_fine, R, ulist, U,
                                                       U_dev, q, jcap, xy, Pf, r_max, Pfine); break;
        default: return hipErrorInvalidValue;
        }
        return hipSuccess;
    };
    unsigned gx, gy;
    hipError_t e;
    if (U_dev) {
        // device-driven: one block per item for the first U_guess rotations' items of each XCD range, then
        // a tail grid on per-XCD counters for the rest
        if (!xq || persist_blocks <= 0 || U_guess <= 0) return hipErrorInvalidValue;
        const long long jcap = ((long long)U_guess * nPB + 7) / 8;
        if (!grid_of(8 * jcap, gx, gy)) return hipErrorInvalidValue;
        if ((e = go(dim3(gx, gy), nullptr, jcap)) != hipSuccess) return e;
        if ((e = go(dim3((unsigned)(8 * std::max(1, persist_blocks / 8))), xq, jcap)) != hipSuccess) return e;
        LAUNCH_CHECK();
        return hipSuccess;
    }
    if (U == 0) return hipSuccess;
    if (!grid_of((long long)U * nPB, gx, gy)) return hipErrorInvalidValue;
    if ((e = go(dim3(gx, gy), nullptr, 0)) != hipSuccess) return e;
    LAUNCH_CHECK();
    return hipSuccess;
}

hipError_t launch_project_plane(hipStream_t s, const float2* grid, int pad, int padF, int c, int is3d, const double* Ai,
                                int nrot, int size, int my_r_max, float2* out) {
    if (nrot == 0) return hipSuccess;
    const int n = size * (size / 2 + 1);
    dim3 grid_(nrot, (n + 255) / 256);
    k_project_plane<<<grid_, 256, 0, s>>>(grid, pad, padF, c, is3d, Ai, size, my_r_max, out);
    LAUNCH_CHECK();
    return hipSuccess;
}

hipError_t launch_prep_images(hipStream_t s, int n, const PrepArgs& a) {
    k_prep_images<<<n, 256, 0, s>>>(a);
    LAUNCH_CHECK();
    return hipSuccess;
}

__global__ void k_pack(PackArgs a) {
    const unsigned long long words = a.off[a.nseg] / 4;
    for (unsigned long long w = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; w < words;
         w += (unsigned long long)gridDim.x * blockDim.x) {
        const unsigned long long b = 4 * w;
        int g = 0;
        while (g + 1 < a.nseg && a.off[g + 1] <= b) ++g;
        const unsigned long long rel = b - a.off[g];
        if (rel < a.len[g])  // the alignment gap after each array is left alone
            reinterpret_cast<unsigned*>(a.dst)[w] = reinterpret_cast<const unsigned*>(a.src[g])[rel / 4];
    }
}

hipError_t launch_pack(hipStream_t s, const PackArgs& a) {
    if (a.nseg <= 0) return hipSuccess;
    if (a.nseg > kPackSegs) return hipErrorInvalidValue;
    const unsigned long long words = a.off[a.nseg] / 4;
    const unsigned blocks = (unsigned)std::min<unsigned long long>((words + 255) / 256, 1024);
    k_pack<<<blocks, 256, 0, s>>>(a);
    return hipGetLastError();
}

hipError_t launch_fct_norm(hipStream_t s, const float* FcT, int Pc, int Pc_pad, int Mpad, int Kq, float* Qc) {
    const size_t tot = (size_t)Kq * Mpad;
    k_fct_norm<<<(unsigned)((tot + 255) / 256), 256, 0, s>>>(FcT, Pc, Pc_pad, Mpad, Kq, Qc);
    LAUNCH_CHECK();
    return hipSuccess;
}

hipError_t launch_coarse_aterm(hipStream_t s, int n, const float* Qc, int Kq, int Mtot, int Mpad, const float* MC2cT,
                               int ldn, float* At) {
    if (n <= 0) return hipSuccess;
    if (Kq % 16 || Mpad % 128 || ldn % 128 || ldn < n) return hipErrorInvalidValue;
    dim3 grid(Mpad / 128, ldn / 128);
    k_coarse_aterm<<<grid, 256, 0, s>>>(Qc, Kq, Mtot, Mpad, MC2cT, ldn, n, At);
    LAUNCH_CHECK();
    return hipSuccess;
}

hipError_t launch_coarse_gemm(hipStream_t s, const GemmArgs& a) {
    const int tiles = (a.Mpad / 128) * (a.Npad / 128);
    if (a.list.count) k_coarse_gemm<true><<<tiles, 256, 0, s>>>(a);
    else if (a.E) k_coarse_gemm<false><<<tiles, 256, 0, s>>>(a);
    else return hipErrorInvalidValue;
    LAUNCH_CHECK();
    return hipSuccess;
}

// Coarse significance given by a caller: per image 1 .. cap keys in [0, MT).  An image that breaks this
// keeps no key (its count becomes 0, so the task build reads nothing out of range) and flags[2], flags[0]
// are set.
__global__ __launch_bounds__(256) void k_check_given_keys(int n, const int* __restrict__ keys, int cap,
                                                          int* __restrict__ cnt, long long MT, int* __restrict__ flags) {
    __shared__ int bad;
    const int img = blockIdx.x;
    if (threadIdx.x == 0) bad = 0;
    __syncthreads();
    const int c = cnt[img];
    if (c < 1 || c > cap) {
        if (threadIdx.x == 0) bad = 1;
    } else {
        for (int q = threadIdx.x; q < c; q += blockDim.x) {
            const int k = keys[(size_t)img * cap + q];
            if (k < 0 || (long long)k >= MT) bad = 1;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0 && bad) {
        cnt[img] = 0;
        atomicOr(&flags[2], 1);
        atomicOr(&flags[0], 1);
    }
}

// After the task build sorted the keys in place: two equal neighbours are a repeated sample (flags[2], [0]).
__global__ __launch_bounds__(256) void k_check_sorted_dups(int n, const int* __restrict__ keys, int cap,
                                                           const int* __restrict__ cnt, int* __restrict__ flags) {
    const int img = blockIdx.x;
    const int c = cnt[img];
    const int* k = keys + (size_t)img * cap;
    bool dup = false;
    for (int q = threadIdx.x + 1; q < c; q += blockDim.x) dup |= k[q] == k[q - 1];
    if (dup) {
        atomicOr(&flags[2], 1);
        atomicOr(&flags[0], 1);
    }
}

hipError_t launch_check_given_keys(hipStream_t s, int n, const int* keys, int cap, int* cnt, long long MT,
                                   int* flags) {
    if (n <= 0) return hipSuccess;
    k_check_given_keys<<<n, 256, 0, s>>>(n, keys, cap, cnt, MT, flags);
    LAUNCH_CHECK();
    return hipSuccess;
}

hipError_t launch_check_sorted_dups(hipStream_t s, int n, const int* keys, int cap, const int* cnt, int* flags) {
    if (n <= 0) return hipSuccess;
    k_check_sorted_dups<<<n, 256, 0, s>>>(n, keys, cap, cnt, flags);
    LAUNCH_CHECK();
    return hipSuccess;
}

hipError_t launch_task_build(hipStream_t s, int n, const SigArgs& sig, int M, int T, int* ntask, int* nblk,
                             long long* task_off, long long* blk_off, unsigned char* used, int* ulist, int* m2u,
                             int* ucount) {
    if (n <= 0) return hipSuccess;
    if (sig.keys) {
        if (sig.cap > kSigCap) return hipErrorInvalidValue;
        k_task_count_keys<<<n, 1024, 0, s>>>(sig.keys, sig.cnt, sig.cap, T, ntask, nblk, used);
    } else {
        k_task_count<<<n, 256, 0, s>>>(sig.mask, M, sig.TW, ntask, nblk, used);
    }
    LAUNCH_CHECK();
    k_scan_images<<<1, 1024, 0, s>>>(ntask, nblk, n, task_off, blk_off);
    LAUNCH_CHECK();
    k_unique_rot<<<1, 1024, 0, s>>>(used, M, ulist, m2u, ucount);
    LAUNCH_CHECK();
    return hipSuccess;
}

hipError_t launch_task_emit(hipStream_t s, int n, const SigArgs& sig, int M, int T, const long long* task_off,
                            const long long* blk_off, int* task_img, int* task_m, long long* task_blk0, int* task_nt,
                            int* blk_task, int* blk_t) {
    if (n <= 0) return hipSuccess;
    if (sig.keys)
        k_task_emit_keys<<<n, 1024, 0, s>>>(sig.keys, sig.cnt, sig.cap, T, task_off, blk_off, task_img, task_m,
                                            task_blk0, task_nt, blk_task, blk_t);
    else
        k_task_emit<<<n, 1024, 0, s>>>(sig.mask, M, sig.TW, T, task_off, blk_off, task_img, task_m, task_blk0, task_nt,
                                        blk_task, blk_t);
    LAUNCH_CHECK();
    return hipSuccess;
}

hipError_t launch_fine_diff2(hipStream_t s, int ntask, int NOR, int NOT, const FineArgs& a, int persist_blocks,
                             int first) {
    unsigned grid;
    if (a.ntask_dev) {  // device-driven: `first` blocks, then the queue kernel (head zeroed by the caller)
        if (!a.abort || !a.queue || persist_blocks <= 0 || first <= 0) return hipErrorInvalidValue;
        grid = (unsigned)first;
    } else {
        if (ntask == 0) return hipSuccess;
        grid = (unsigned)ntask;
    }
    const unsigned pg = (unsigned)std::max(persist_blocks, 1);
    const FineArgs& fa = a;
    if (NOR == 8 && NOT == 4) {
        k_fine_diff2<8, 4><<<grid, 256, 0, s>>>(fa);
        if (a.ntask_dev) k_fine_diff2_queue<8, 4><<<pg, 256, 0, s>>>(fa, first);
    } else if (NOR == 2 && NOT == 4) {
        k_fine_diff2<2, 4><<<grid, 256, 0, s>>>(fa);
        if (a.ntask_dev) k_fine_diff2_queue<2, 4><<<pg, 256, 0, s>>>(fa, first);
    } else if (NOR == 1 && NOT == 1) {
        k_fine_diff2<1, 1><<<grid, 256, 0, s>>>(fa);
        if (a.ntask_dev) k_fine_diff2_queue<1, 1><<<pg, 256, 0, s>>>(fa, first);
    } else {
        return hipErrorInvalidValue;
    }
    LAUNCH_CHECK();
    return hipSuccess;
}

template <int NOR, int NOT>
static hipError_t post_ppt(hipStream_t s, int n, int Pf, const PostArgs& a) {
    if (a.slice_val) {  // 512 threads: room in registers for the image's own pixels
        if (Pf <= 4 * 256) k_post<NOR, NOT, 4, 256, true><<<n, 256, 0, s>>>(a);
        else if (Pf <= 8 * 512) k_post<NOR, NOT, 8, 512, true><<<n, 512, 0, s>>>(a);
        else if (Pf <= 16 * 512) k_post<NOR, NOT, 16, 512, true><<<n, 512, 0, s>>>(a);
        else if (Pf <= 16 * 1024) k_post<NOR, NOT, 16, 1024, true><<<n, 1024, 0, s>>>(a);
        else return hipErrorInvalidValue;
        return hipSuccess;
    }
    if (Pf <= 4 * 256) k_post<NOR, NOT, 4, 256, false><<<n, 256, 0, s>>>(a);
    else if (Pf <= 4 * 1024) k_post<NOR, NOT, 4, 1024, false><<<n, 1024, 0, s>>>(a);
    else if (Pf <= 8 * 1024) k_post<NOR, NOT, 8, 1024, false><<<n, 1024, 0, s>>>(a);
    else if (Pf <= 16 * 1024) k_post<NOR, NOT, 16, 1024, false><<<n, 1024, 0, s>>>(a);
    else return hipErrorInvalidValue;
    return hipSuccess;
}

hipError_t launch_post(hipStream_t s, int n, int NOR, int NOT, const PostArgs& a) {
    hipError_t e;
    if (NOR == 8 && NOT == 4) e = post_ppt<8, 4>(s, n, a.Pf, a);
    else if (NOR == 2 && NOT == 4) e = post_ppt<2, 4>(s, n, a.Pf, a);
    else if (NOR == 1 && NOT == 1) e = post_ppt<1, 1>(s, n, a.Pf, a);
    else return hipErrorInvalidValue;
    if (e != hipSuccess) return e;
    LAUNCH_CHECK();
    return hipSuccess;
}

bool post_slices_supported(int NOR, int NOT, int Pf, int is3d) {
    return NOR == 8 && NOT == 4 && Pf <= 4 * 8 * 1024 && is3d;  // up to 4 pixel ranges of 8192 per slice
}

hipError_t launch_post_slices(hipStream_t s, long long ntask, int NOR, int NOT, const PostArgs& a, int persist_blocks,
                              long long guess) {
    if (!post_slices_supported(NOR, NOT, a.Pf, a.is3d) || !a.slice_val) return hipErrorNotSupported;
    const long long nh = (a.Pf + 8 * 1024 - 1) / (8 * 1024);
    if (a.ntask_dev) {
        // device-driven: the first launch covers guess tasks (one block per item), the tail the rest
        if (!a.abort || !a.xq || persist_blocks <= 0 || guess <= 0) return hipErrorInvalidValue;
        const long long jcap = (guess * NOR * nh + 7) / 8;
        if (8 * jcap > 0x7fffffffLL) return hipErrorInvalidValue;
        k_post_slice<8, 4, 8, 1024, false><<<(unsigned)(8 * jcap), 1024, 0, s>>>(a, ntask, jcap);
        k_post_slice<8, 4, 8, 1024, true><<<(unsigned)(8 * std::max(1, persist_blocks / 8)), 1024, 0, s>>>(a, ntask,
                                                                                                          jcap);
        LAUNCH_CHECK();
        return hipSuccess;
    }
    if (ntask <= 0) return hipSuccess;
    const long long blocks = 8 * ((ntask * NOR * nh + 7) / 8);
    if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    k_post_slice<8, 4, 8, 1024, false><<<(unsigned)blocks, 1024, 0, s>>>(a, ntask, 0);
    LAUNCH_CHECK();
    return hipSuccess;
}

__global__ void k_fill32(uint32_t* __restrict__ p, uint32_t v, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

hipError_t launch_fill32(hipStream_t s, void* p, uint32_t value, size_t words) {
    if (words == 0) return hipSuccess;
    const unsigned blocks = (unsigned)std::min<size_t>((words + 255) / 256, 4096);
    k_fill32<<<blocks, 256, 0, s>>>(static_cast<uint32_t*>(p), value, words);
    LAUNCH_CHECK();
    return hipSuccess;
}

// Device-driven pools: flags[0] = 1 when an image's coarse list overflowed (status 8), it has no
// significant sample (status 1), its significant samples exceed the key buffer, or the pool needs more
// fine projections than the buffer holds; the pool is then rerun host-synchronised.
__global__ void k_pool_check(int n, const int* __restrict__ status, const int* __restrict__ cnt, int cap,
                             const int* __restrict__ ucount, int ucap, int* __restrict__ flags) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool bad = false;
    if (i < n) bad = (status[i] & 9) != 0 || (cnt && cnt[i] > cap);
    if (i == 0) bad = bad || *ucount > ucap;
    if (bad) atomicOr(&flags[0], 1);
}

hipError_t launch_pool_check(hipStream_t s, int n, const int* status, const int* cnt, int cap, const int* ucount,
                             int ucap, int* flags) {
    k_pool_check<<<(std::max(n, 1) + 255) / 256, 256, 0, s>>>(n, status, cnt, cap, ucount, ucap, flags);
    LAUNCH_CHECK();
    return hipSuccess;
}

hipError_t launch_meta(hipStream_t s, const MetaArgs& a) {
    if (a.n <= 0) return hipSuccess;
    k_meta<<<(a.n + 127) / 128, 128, 0, s>>>(a);
    LAUNCH_CHECK();
    return hipSuccess;
}

hipError_t launch_store_sums(hipStream_t s, const StoreArgs& a) {
    k_store_sums<<<(a.n + 127) / 128, 128, 0, s>>>(a);
    LAUNCH_CHECK();
    return hipSuccess;
}

hipError_t launch_bp_images(hipStream_t s, int n, const BpImageArgs& a) {
    if (n <= 0) return hipSuccess;
    k_bp_images<<<n, 256, 0, s>>>(a);
    LAUNCH_CHECK();
    return hipSuccess;
}

}  // namespace rome
