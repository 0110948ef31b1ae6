// Learned pose proposal (no counterpart in the reference): a small MLP, trained on the E-step's own
// posteriors, proposes per image the (class, direction, in-plane bin) cells worth an exact coarse pass.
//
//   x  = 2 * 812 whitened low-resolution coefficients of the masked image (BF16, zero-padded to Fin)
//   h1 = gelu(W1 x + b1), h2 = gelu(W2 h1 + b2), h3 = gelu(W3 h2 + b3)          H = 1024
//   z  = W4 h3 + b4: K * C cell logits (C = HEALPix order-3 directions x in-plane bins) + 2 shift outputs
//
// Every product runs on BF16 MFMA (v_mfma_f32_32x32x16_bf16) with FP32 accumulation; the parameters are
// kept in FP32 (Adam state too) with BF16 copies for the products.  Training: softmax cross entropy
// against the E-step's posterior over cells plus a squared shift error, the batch mean; backward as
// the transposed products.  Inference: the top-k cells and their neighbours become per-image rotation
// lists; the restricted coarse pass computes their diff2 (split three-product, like the recompute of
// marked samples) into the coarse list, and a small audit slice of every pool also runs the exhaustive
// pass to measure the posterior mass the proposal captured.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_util.h"
#include "kernels.h"
#include "proposal.h"

namespace rome {

namespace {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f16v __attribute__((ext_vector_type(16)));

DEV f16v mfma_bf16(bf16x8 a, bf16x8 b, f16v c) { return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0); }

// 8-wide fragments and the 32x32x16 MFMA of the arm's format
template <typename T>
struct LowpMfma;
template <>
struct LowpMfma<__bf16> {
    typedef __bf16 v8 __attribute__((ext_vector_type(8)));
    static DEV f16v run(v8 a, v8 b, f16v c) { return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0); }
};
template <>
struct LowpMfma<_Float16> {
    typedef _Float16 v8 __attribute__((ext_vector_type(8)));
    static DEV f16v run(v8 a, v8 b, f16v c) { return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0); }
};

// ------------------------------------------------------------------ dense products
// C[m][n] = sum_k A[m][k] B[n][k] (+ bias[n]); A, B BF16 with K-contiguous rows (K % 32 == 0), C FP32.
// 256 threads = 4 waves in 2 x 2, tile 128 x 128, each wave 64 x 64 = 2 x 2 accumulator blocks; the
// k chunks of 32 go through LDS (80-byte rows: 16-byte fragment reads without bank conflicts).
__global__ __launch_bounds__(256) void k_mlp_gemm(const bf16* __restrict__ A, int lda, const bf16* __restrict__ B,
                                                  int ldb, float* __restrict__ C, int ldc,
                                                  const float* __restrict__ bias, int M, int N, int K) {
    constexpr int BM = 128, BN = 128, BK = 32, LDW = BK + 8;
    __shared__ __attribute__((aligned(16))) bf16 As[BM][LDW];
    __shared__ __attribute__((aligned(16))) bf16 Bs[BN][LDW];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    f16v acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f16v{};
    const bf16x8 zero = {};
    for (int k0 = 0; k0 < K; k0 += BK) {
        // 128 rows x 32 k = 512 chunks of 8 per operand: two per thread
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const int q = tid + c * 256, r = q >> 2, kc = (q & 3) * 8;
            const bf16x8 va = m0 + r < M ? *reinterpret_cast<const bf16x8*>(A + (size_t)(m0 + r) * lda + k0 + kc) : zero;
            const bf16x8 vb = n0 + r < N ? *reinterpret_cast<const bf16x8*>(B + (size_t)(n0 + r) * ldb + k0 + kc) : zero;
            *reinterpret_cast<bf16x8*>(&As[r][kc]) = va;
            *reinterpret_cast<bf16x8*>(&Bs[r][kc]) = vb;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < BK; kk += 16) {
            bf16x8 fa[2], fb[2];
#pragma unroll
            for (int i = 0; i < 2; ++i)
                fa[i] = *reinterpret_cast<const bf16x8*>(&As[wm * 64 + i * 32 + (lane & 31)][kk + (lane >> 5) * 8]);
#pragma unroll
            for (int j = 0; j < 2; ++j)
                fb[j] = *reinterpret_cast<const bf16x8*>(&Bs[wn * 64 + j * 32 + (lane & 31)][kk + (lane >> 5) * 8]);
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) acc[i][j] = mfma_bf16(fa[i], fb[j], acc[i][j]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int row = m0 + wm * 64 + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                const int col = n0 + wn * 64 + j * 32 + (lane & 31);
                if (row < M && col < N) C[(size_t)row * ldc + col] = acc[i][j][r] + (bias ? bias[col] : 0.f);
            }
}

// out[c][r] = bf16(in[r][c]) for r < R, c < C (out rows have stride ldo); 32 x 32 tiles through LDS.
template <typename Tin>
__global__ void k_transpose_bf16(const Tin* __restrict__ in, int R, int Cn, int ldi, bf16* __restrict__ out, int ldo) {
    __shared__ float tile[32][33];
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 256 threads: 8 rows per pass
    for (int y = ty; y < 32; y += 8) {
        const int r = r0 + y, c = c0 + tx;
        tile[y][tx] = (r < R && c < Cn) ? static_cast<float>(in[(size_t)r * ldi + c]) : 0.f;
    }
    __syncthreads();
    for (int y = ty; y < 32; y += 8) {
        const int c = c0 + y, r = r0 + tx;
        if (c < Cn && r < R) out[(size_t)c * ldo + r] = static_cast<bf16>(tile[tx][y]);
    }
}

DEV float gelu(float a) { return 0.5f * a * (1.f + erff(a * 0.70710678118654752f)); }
DEV float gelu_grad(float a) {
    return 0.5f * (1.f + erff(a * 0.70710678118654752f)) + a * expf(-0.5f * a * a) * 0.39894228040143268f;
}

// h = bf16(gelu(a)), [rows][cols] with row strides lda / ldh
__global__ void k_gelu_fwd(const float* __restrict__ a, int lda, bf16* __restrict__ h, int ldh, int rows, int cols) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)rows * cols) return;
    const int r = (int)(i / cols), c = (int)(i % cols);
    h[(size_t)r * ldh + c] = static_cast<bf16>(gelu(a[(size_t)r * lda + c]));
}

// da = dh * gelu'(a) in place of dh, and its BF16 copy
__global__ void k_gelu_bwd(float* __restrict__ dh, const float* __restrict__ a, bf16* __restrict__ daq, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float v = dh[i] * gelu_grad(a[i]);
    dh[i] = v;
    daq[i] = static_cast<bf16>(v);
}

// db[c] = sum_r d[r][c] (FP32 columns, one thread per column, rows in order)
__global__ void k_colsum(const float* __restrict__ d, int rows, int cols, int ld, float* __restrict__ db) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols) return;
    float s = 0.f;
    for (int r = 0; r < rows; ++r) s += d[(size_t)r * ld + c];
    db[c] = s;
}

// Softmax cross entropy of one image's cell logits against its sparse target (cells, probabilities) and
// the squared shift error: dz = (softmax - target) / B on the cells, lam (z - s) / B on the shift outputs
// (zero in the padding columns), FP32 and BF16 copies; the image's loss / B is added to *loss.
__global__ __launch_bounds__(256) void k_softmax_ce(const float* __restrict__ z, int ldz, int ncell, int B,
                                                    const int* __restrict__ tcell, const float* __restrict__ tprob,
                                                    const int* __restrict__ tcnt, int tcap,
                                                    const float* __restrict__ tshift, float lam, float* __restrict__ dz,
                                                    bf16* __restrict__ dzq, int ldd, float* __restrict__ loss) {
    __shared__ float red[16];
    const int b = blockIdx.x, tid = threadIdx.x;
    const float* zr = z + (size_t)b * ldz;
    float mx = -INFINITY;
    for (int c = tid; c < ncell; c += blockDim.x) mx = fmaxf(mx, zr[c]);
    mx = block_reduce(mx, [](float u, float v) { return fmaxf(u, v); }, red);
    float s = 0.f;
    for (int c = tid; c < ncell; c += blockDim.x) s += expf(zr[c] - mx);
    s = block_reduce(s, [](float u, float v) { return u + v; }, red);
    const float ls = logf(s), invB = 1.f / (float)B;
    float* dr = dz + (size_t)b * ldd;
    bf16* dq = dzq + (size_t)b * ldd;
    for (int c = tid; c < ncell; c += blockDim.x) dr[c] = expf(zr[c] - mx) / s * invB;
    for (int c = ncell + 2 + tid; c < ldd; c += blockDim.x) dr[c] = 0.f;
    __syncthreads();
    const int nt = tcnt[b];
    float l = 0.f;
    for (int q = tid; q < nt; q += blockDim.x) {
        const int c = tcell[(size_t)b * tcap + q];
        const float p = tprob[(size_t)b * tcap + q];
        l -= p * (zr[c] - mx - ls);
        dr[c] -= p * invB;  // target cells are distinct
    }
    if (tid < 2) {
        const float d = zr[ncell + tid] - tshift[2 * b + tid];
        dr[ncell + tid] = lam * d * invB;
        l += 0.5f * lam * d * d;
    }
    l = block_reduce(l, [](float u, float v) { return u + v; }, red);
    __syncthreads();
    for (int c = tid; c < ldd; c += blockDim.x) dq[c] = static_cast<bf16>(dr[c]);
    if (tid == 0) atomicAdd(loss, l * invB);
}

// Adam (bias-corrected moments): w -= lr * (m * bc1) / (sqrt(v * bc2) + eps); BF16 copy of w.
__global__ void k_adam(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v, const float* __restrict__ g,
                       bf16* __restrict__ wq, size_t n, float lr, float b1, float b2, float eps, float bc1, float bc2) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi * bc1) / (sqrtf(vi * bc2) + eps);
    w[i] = wi;
    if (wq) wq[i] = static_cast<bf16>(wi);
}

__global__ void k_to_bf16(const float* __restrict__ in, bf16* __restrict__ out, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = static_cast<bf16>(in[i]);
}

// Rows idx[r] of src [*][ld] -> dst [r][ld] (BF16 rows of 16-byte chunks)
__global__ void k_gather_rows(const bf16* __restrict__ src, const int* __restrict__ idx, int rows, int ld,
                              bf16* __restrict__ dst) {
    const int r = blockIdx.x;
    if (r >= rows) return;
    const int4* s = reinterpret_cast<const int4*>(src + (size_t)idx[r] * ld);
    int4* d = reinterpret_cast<int4*>(dst + (size_t)r * ld);
    for (int q = threadIdx.x; q < ld / 8; q += blockDim.x) d[q] = s[q];
}

// Training-set rows idx[r] of the sparse targets -> batch row r
__global__ void k_gather_targets(const int* __restrict__ idx, int rows, int tcap, const int* __restrict__ tcell,
                                 const float* __restrict__ tprob, const int* __restrict__ tcnt,
                                 const float* __restrict__ tshift, int* __restrict__ ocell, float* __restrict__ oprob,
                                 int* __restrict__ ocnt, float* __restrict__ oshift) {
    const int r = blockIdx.x;
    if (r >= rows) return;
    const size_t src = (size_t)idx[r];
    for (int q = threadIdx.x; q < tcap; q += blockDim.x) {
        ocell[(size_t)r * tcap + q] = tcell[src * tcap + q];
        oprob[(size_t)r * tcap + q] = tprob[src * tcap + q];
    }
    if (threadIdx.x == 0) {
        ocnt[r] = tcnt[src];
        oshift[2 * r] = tshift[2 * src];
        oshift[2 * r + 1] = tshift[2 * src + 1];
    }
}

// ------------------------------------------------------------------ features
// x[img][0 .. 2 ncoef) = the image's masked transform at the coefficient pixels, whitened by the noise
// (X / sigma), then scaled to unit RMS over the coefficients; zero up to ldx.
__global__ __launch_bounds__(256) void k_mlp_features(const float2* __restrict__ fmask, int F2f,
                                                      const int* __restrict__ coef_n, const int* __restrict__ coef_ires,
                                                      int ncoef, const double* __restrict__ sigma2,
                                                      const int* __restrict__ group, int oriF, bf16* __restrict__ x,
                                                      int ldx) {
    __shared__ float red[16];
    const int img = blockIdx.x, tid = threadIdx.x;
    const double* s2 = sigma2 + (size_t)group[img] * oriF;
    const float2* f = fmask + (size_t)img * F2f;
    float ss = 0.f;
    for (int q = tid; q < ncoef; q += blockDim.x) {
        const float2 v = f[coef_n[q]];
        const float w = (float)(1.0 / sqrt(s2[coef_ires[q]]));
        ss += w * w * (v.x * v.x + v.y * v.y);
    }
    ss = block_reduce(ss, [](float u, float v) { return u + v; }, red);
    const float scale = ss > 0.f ? rsqrtf(ss / (float)(2 * ncoef)) : 0.f;
    bf16* xr = x + (size_t)img * ldx;
    for (int q = tid; q < ncoef; q += blockDim.x) {
        const float2 v = f[coef_n[q]];
        const float w = (float)(1.0 / sqrt(s2[coef_ires[q]])) * scale;
        xr[2 * q] = static_cast<bf16>(v.x * w);
        xr[2 * q + 1] = static_cast<bf16>(v.y * w);
    }
    for (int q = 2 * ncoef + tid; q < ldx; q += blockDim.x) xr[q] = static_cast<bf16>(0.f);
}

// ------------------------------------------------------------------ training targets
// The E-step's posterior of the last pool, aggregated over proposal cells: every fine sample's weight
// pdf_direction * pdf_offset * exp(-(diff2 - min)) / sum (the fine pass's own normalisation) goes to the cell
// (class, order-3 direction = NEST child o / 2 of the coarse direction, in-plane bin of the coarse psi);
// the tcap heaviest cells (renormalised) and the weighted mean shift (old offset + fine shift, over
// offset_range) are the image's targets.  One block per image, the cell sums in LDS.
__global__ __launch_bounds__(1024) void k_proposal_targets(ProposalTargetArgs a) {
    extern __shared__ float cell[];  // [ncell]
    __shared__ float red[16];
    const int img = blockIdx.x, tid = threadIdx.x;
    const int ncell = a.K * a.C;
    for (int c = tid; c < ncell; c += blockDim.x) cell[c] = 0.f;
    __syncthreads();
    const long long b0 = a.img_blk[img], b1 = a.img_blk[img + 1];
    const int NE = a.NOR * a.NOT;
    const double dmin = a.dmin[img], sumw = a.sumw[img];
    float sx = 0.f, sy = 0.f;
    for (long long e = b0 * NE + tid; e < b1 * NE; e += blockDim.x) {
        const long long b = e / NE;
        const int o = (int)((e % NE) / a.NOT), ot = (int)(e % a.NOT);
        const int m = a.task_m[a.blk_task[b]], t = a.blk_t[b];
        const int k = m / a.R, r = m % a.R, d2 = r / a.PSI, ps = r % a.PSI;
        const double po = pdf_offset(a.trans_x[t], a.trans_y[t], a.old_offx[img], a.old_offy[img], a.prior_x[k],
                                     a.prior_y[k], a.sigma2_offset);
        const double dd = a.D2[e] - dmin;
        const double w = dd > 700. ? 0. : a.pdf_m[m] * po * exp(-dd) / sumw;
        if (!(w > 0.)) continue;
        const int c = k * a.C + (4 * d2 + (o >> 1)) * a.nbins + ps * a.nbins / a.PSI;
        atomicAdd(&cell[c], (float)w);
        sx += (float)w * (float)(a.old_offx[img] + a.shift_x[t * a.NOT + ot]);
        sy += (float)w * (float)(a.old_offy[img] + a.shift_y[t * a.NOT + ot]);
    }
    sx = block_reduce(sx, [](float u, float v) { return u + v; }, red);
    sy = block_reduce(sy, [](float u, float v) { return u + v; }, red);
    __syncthreads();
    // the tcap heaviest cells: bisection on the mass threshold (counts of cells at or above it)
    float lo = 0.f, hi = 1.f;
    for (int it = 0; it < 40; ++it) {
        const float mid = 0.5f * (lo + hi);
        float cnt = 0.f;
        for (int c = tid; c < ncell; c += blockDim.x) cnt += (cell[c] > 0.f && cell[c] >= mid) ? 1.f : 0.f;
        cnt = block_reduce(cnt, [](float u, float v) { return u + v; }, red);
        if (cnt > (float)a.tcap) lo = mid;
        else hi = mid;
    }
    const float thr = hi;  // at most tcap cells >= thr, and the cut is within 2^-40 of the tcap-th mass
    __shared__ int n_out;
    __shared__ float tot;
    if (tid == 0) {
        n_out = 0;
        tot = 0.f;
    }
    __syncthreads();
    const size_t row = (size_t)a.row0 + img, base = row * a.tcap;
    for (int c = tid; c < ncell; c += blockDim.x) {
        if (cell[c] > 0.f && cell[c] >= thr) {
            const int q = atomicAdd(&n_out, 1);
            if (q < a.tcap) {
                a.tcell[base + q] = c;
                a.tprob[base + q] = cell[c];
                atomicAdd(&tot, cell[c]);
            }
        }
    }
    __syncthreads();
    const int no = min(n_out, a.tcap);
    for (int q = tid; q < no; q += blockDim.x) a.tprob[base + q] /= tot;
    if (tid == 0) {
        a.tcnt[row] = no;
        a.tshift[2 * row] = sx / a.shift_scale;
        a.tshift[2 * row + 1] = sy / a.shift_scale;
    }
}

// ------------------------------------------------------------------ candidates
// Per image: softmax over the K * C cell logits; the topk largest logits (ties included) select cells;
// each selected cell marks its group (class, coarse direction = order-3 direction / 4, in-plane bin), the
// groups of the 8 neighbouring coarse directions in the same bin and the two neighbouring bins; the
// predicted mass is the softmax mass of every cell in a marked group.  Below min_mass every group is
// marked.  The marked groups become the image's candidate rotations m = (k * D + d) * PSI + psi (psi in
// the bin), ascending.
__global__ __launch_bounds__(1024) void k_select_candidates(CandidateArgs a) {
    __shared__ float red[16];
    __shared__ unsigned gm[kMaxGroupWords];
    __shared__ int scan[1024];
    const int img = blockIdx.x, tid = threadIdx.x;
    const int ncell = a.K * a.C, ngroup = a.K * a.D * a.nbins, gw = (ngroup + 31) / 32;
    const float* z = a.z + (size_t)img * a.ldz;
    for (int q = tid; q < gw; q += blockDim.x) gm[q] = 0u;
    float mx = -INFINITY;
    for (int c = tid; c < ncell; c += blockDim.x) mx = fmaxf(mx, z[c]);
    mx = block_reduce(mx, [](float u, float v) { return fmaxf(u, v); }, red);
    float s = 0.f;
    for (int c = tid; c < ncell; c += blockDim.x) s += expf(z[c] - mx);
    s = block_reduce(s, [](float u, float v) { return u + v; }, red);
    // threshold: the largest logit level with >= topk cells at or above it (bisection over 64 nats)
    float lo = mx - 64.f, hi = mx;
    for (int it = 0; it < 32; ++it) {
        const float mid = 0.5f * (lo + hi);
        float cnt = 0.f;
        for (int c = tid; c < ncell; c += blockDim.x) cnt += z[c] >= mid ? 1.f : 0.f;
        cnt = block_reduce(cnt, [](float u, float v) { return u + v; }, red);
        if (cnt >= (float)a.topk) lo = mid;
        else hi = mid;
    }
    const float thr = lo;
    const int per_dir3 = a.nbins;
    for (int c = tid; c < ncell; c += blockDim.x) {
        if (!(z[c] >= thr)) continue;
        const int k = c / a.C, rem = c % a.C, d3 = rem / per_dir3, bin = rem % per_dir3, d = d3 >> 2;
        auto mark = [&](int dd, int bb) {
            const int g = (k * a.D + dd) * a.nbins + bb;
            atomicOr(&gm[g >> 5], 1u << (g & 31));
        };
        mark(d, bin);
        mark(d, (bin + 1) % a.nbins);
        mark(d, (bin + a.nbins - 1) % a.nbins);
        for (int q = 0; q < 8; ++q) {
            const int nb = a.nbr[d * 8 + q];
            if (nb >= 0) mark(nb, bin);
        }
    }
    __syncthreads();
    float pm = 0.f;
    for (int c = tid; c < ncell; c += blockDim.x) {
        const int k = c / a.C, rem = c % a.C, d = (rem / per_dir3) >> 2, bin = rem % per_dir3;
        const int g = (k * a.D + d) * a.nbins + bin;
        if (gm[g >> 5] >> (g & 31) & 1u) pm += expf(z[c] - mx);
    }
    pm = block_reduce(pm, [](float u, float v) { return u + v; }, red) / s;
    const bool all = !(pm >= a.min_mass);
    if (all)
        for (int q = tid; q < gw; q += blockDim.x) gm[q] = q < ngroup / 32 ? ~0u : (1u << (ngroup & 31)) - 1u;
    __syncthreads();
    // ordered emission: groups in (k, d, bin) order, PSI / nbins rotations each
    const int pb = a.PSI / a.nbins;
    const int per = (ngroup + 1023) / 1024;  // groups per thread, contiguous
    const int g0 = tid * per, g1 = min(g0 + per, ngroup);
    int cntg = 0;
    for (int g = g0; g < g1; ++g) cntg += gm[g >> 5] >> (g & 31) & 1u;
    scan[tid] = cntg;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const int v = tid >= o ? scan[tid - o] : 0;
        __syncthreads();
        scan[tid] += v;
        __syncthreads();
    }
    int pos = (scan[tid] - cntg) * pb;
    int* out = a.cand_m + (size_t)img * a.cap;
    for (int g = g0; g < g1; ++g) {
        if (!(gm[g >> 5] >> (g & 31) & 1u)) continue;
        const int k = g / (a.D * a.nbins), rem = g % (a.D * a.nbins), d = rem / a.nbins, bin = rem % a.nbins;
        for (int j = 0; j < pb; ++j)
            if (pos < a.cap) out[pos++] = (k * a.D + d) * a.PSI + bin * pb + j;
    }
    if (tid == 1023) a.cand_cnt[img] = min(scan[1023] * pb, a.cap);
    unsigned* gout = a.gmask + (size_t)img * a.gwords;
    for (int q = tid; q < a.gwords; q += blockDim.x) gout[q] = q < gw ? gm[q] : 0u;
    if (tid == 0) {
        a.pmass[img] = pm;
        a.all[img] = all ? 1 : 0;
    }
}

// ------------------------------------------------------------------ restricted coarse pass
// One block per image: its candidate rotations 32 at a time x all translations (one wave per tile of 32),
// split three-product MFMA as in the recompute of marked samples; E = At - C / (sF s_img).  Every entry
// with v = E - ln pdf_direction below the image's running minimum of v + reach goes to the coarse list
// (entry t * Mtot + m, value E; a superset of the weight window, as the cross-term epilogue keeps), and
// the minimum of E over rotations with a prior goes to emin_key.
template <typename T>
__global__ __launch_bounds__(256) void k_cand_coarse(CandCoarseArgs a) {
    typedef typename LowpMfma<T>::v8 v8;
    __shared__ int rowm[32];
    __shared__ float rowsg[32];
    __shared__ int s_vmin, s_emin, s_cnt;
    __shared__ float red[16];
    const int img = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int ncand = a.cand_cnt[img];
    const int* cm = a.cand_m + (size_t)img * a.cap_m;
    const int KH = a.K16 / 2, Tn = a.T, ntt = (Tn + 31) / 32;
    const T* Wh = reinterpret_cast<const T*>(a.Whi) + (size_t)img * Tn * a.K16;
    const T* Wl = reinterpret_cast<const T*>(a.Wlo) + (size_t)img * Tn * a.K16;
    const T* Ahi = reinterpret_cast<const T*>(a.Ahi);
    const T* Alo = reinterpret_cast<const T*>(a.Alo);
    const float inv = a.inv_scale[img];
    const float* At = a.At + (size_t)img * a.Mpad;
    const float reach = a.reach[img];
    unsigned* ent = a.list.ent + (size_t)img * a.list.cap;
    float* val = a.list.val + (size_t)img * a.list.cap;
    if (tid == 0) {
        s_vmin = 0x7fffffff;
        s_emin = 0x7fffffff;
        s_cnt = 0;
    }
    const v8 zero = {};
    for (int q0 = 0; q0 < ncand; q0 += 32) {
        const int nr = min(32, ncand - q0);
        if (tid < 32) {
            const int m = tid < nr ? cm[q0 + tid] : -1;
            const bool mirrored = m >= 0 && a.psih > 0 && m % (2 * a.psih) >= a.psih;
            rowm[tid] = m;
            rowsg[tid] = mirrored ? -1.f : 1.f;
        }
        __syncthreads();
        const bool bounded = s_vmin != 0x7fffffff;
        const float vbound = float_from_order_key(s_vmin);  // bound before this tile: the kept set is a superset
        const int r = lane & 31;
        const int m = rowm[r];
        int h = m;
        if (m >= 0 && a.psih > 0) {
            const int P2 = 2 * a.psih, qd = m / P2, ps = m - qd * P2;
            h = qd * a.psih + (ps >= a.psih ? ps - a.psih : ps);
        }
        const T* ah_row = Ahi + (size_t)(m >= 0 ? h : 0) * a.K16 + (lane >> 5) * 8;
        const T* al_row = Alo + (size_t)(m >= 0 ? h : 0) * a.K16 + (lane >> 5) * 8;
        float tile_vmin = INFINITY, tile_emin = INFINITY;
        for (int tt = wave; tt < ntt; tt += 4) {
            const int t = tt * 32 + r;
            const bool tok = t < Tn;
            const T* bh_row = Wh + (size_t)(tok ? t : 0) * a.K16 + (lane >> 5) * 8;
            const T* bl_row = Wl + (size_t)(tok ? t : 0) * a.K16 + (lane >> 5) * 8;
            f16v P = {}, Q = {};
            for (int k0 = 0; k0 < KH; k0 += 16) {
                const v8 ah = m >= 0 ? *reinterpret_cast<const v8*>(ah_row + k0) : zero;
                const v8 al = m >= 0 ? *reinterpret_cast<const v8*>(al_row + k0) : zero;
                const v8 bh = tok ? *reinterpret_cast<const v8*>(bh_row + k0) : zero;
                const v8 bl = tok ? *reinterpret_cast<const v8*>(bl_row + k0) : zero;
                P = LowpMfma<T>::run(ah, bh, P);
                P = LowpMfma<T>::run(ah, bl, P);
                P = LowpMfma<T>::run(al, bh, P);
            }
            for (int k0 = KH; k0 < a.K16; k0 += 16) {
                const v8 ah = m >= 0 ? *reinterpret_cast<const v8*>(ah_row + k0) : zero;
                const v8 al = m >= 0 ? *reinterpret_cast<const v8*>(al_row + k0) : zero;
                const v8 bh = tok ? *reinterpret_cast<const v8*>(bh_row + k0) : zero;
                const v8 bl = tok ? *reinterpret_cast<const v8*>(bl_row + k0) : zero;
                Q = LowpMfma<T>::run(ah, bh, Q);
                Q = LowpMfma<T>::run(ah, bl, Q);
                Q = LowpMfma<T>::run(al, bh, Q);
            }
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const int row = 8 * (j >> 2) + 4 * (lane >> 5) + (j & 3);
                const int tc = tt * 32 + (lane & 31);
                if (row >= nr || tc >= Tn) continue;
                const int mm = rowm[row];
                const float lpd = a.list.lpd[mm];
                if (!(lpd > -INFINITY)) continue;  // no prior: no entry (the reference skips the rotation)
                const float c = P[j] + rowsg[row] * Q[j];
                const float e = __builtin_fmaf(-c, inv, At[mm]);
                const float v = e - lpd;
                tile_emin = fminf(tile_emin, e);
                tile_vmin = fminf(tile_vmin, v);
                if (!bounded || v < vbound + reach + 1e-6f * (fabsf(vbound) + reach) + 1e-3f) {
                    const int q = atomicAdd(&s_cnt, 1);
                    if (q < a.list.cap) {
                        ent[q] = (unsigned)tc * (unsigned)a.Mtot + (unsigned)mm;
                        val[q] = e;
                    }
                }
            }
        }
        tile_vmin = block_reduce(tile_vmin, [](float u, float v) { return fminf(u, v); }, red);
        tile_emin = block_reduce(tile_emin, [](float u, float v) { return fminf(u, v); }, red);
        if (tid == 0) {
            if (tile_vmin < INFINITY) s_vmin = min(s_vmin, float_order_key(tile_vmin + 1e-6f * fabsf(tile_vmin) + 1e-4f));
            if (tile_emin < INFINITY) s_emin = min(s_emin, float_order_key(tile_emin));
        }
        __syncthreads();
    }
    if (tid == 0) {
        a.list.count[img] = s_cnt;
        a.emin_key[img] = s_emin;
    }
}

// ------------------------------------------------------------------ audit
// Exact posterior mass the proposal captured for an audited image, from its exhaustive pose term E
// [T][Mtot] (the full coarse pass): sum of pdf_direction * pdf_offset * exp(-(d - min)) over the rotations
// in the image's marked groups, over the sum over every rotation.
__global__ __launch_bounds__(1024) void k_audit_mass(AuditArgs a) {
    __shared__ double red[16];
    __shared__ double lpo[kMaxKT];
    const int ai = blockIdx.x, img = a.img0 + ai, tid = threadIdx.x;
    const float* E = a.E + (size_t)ai * a.T * a.Mtot;
    const unsigned* gm = a.gmask + (size_t)img * a.gwords;
    const WeightParams& wp = a.wp;
    for (int q = tid; q < wp.K * wp.T; q += blockDim.x) {
        const int k = q / wp.T, t = q % wp.T;
        const double po = pdf_offset(wp.trans_x[t], wp.trans_y[t], wp.old_offx[img], wp.old_offy[img], wp.prior_x[k],
                                     wp.prior_y[k], wp.sigma2_offset);
        lpo[q] = po > 0 ? log(po) : -INFINITY;
    }
    __syncthreads();
    const size_t cnt = (size_t)a.T * a.Mtot;
    double umin = INFINITY;
    for (size_t e = tid; e < cnt; e += blockDim.x) {
        const int t = (int)(e / a.Mtot), m = (int)(e % a.Mtot);
        if (!(wp.pdf_m[m] > 0.)) continue;
        umin = fmin(umin, (double)E[e] - wp.lpdf_m[m] - lpo[wp.k_m[m] * wp.T + t]);
    }
    umin = block_reduce(umin, [](double u, double v) { return fmin(u, v); }, red);
    double all = 0., in = 0.;
    const int pb = wp.PSI / a.nbins;
    for (size_t e = tid; e < cnt; e += blockDim.x) {
        const int t = (int)(e / a.Mtot), m = (int)(e % a.Mtot);
        if (!(wp.pdf_m[m] > 0.)) continue;
        const double du = (double)E[e] - wp.lpdf_m[m] - lpo[wp.k_m[m] * wp.T + t] - umin;
        if (du > 700.) continue;
        const double w = exp(-du);
        all += w;
        const int k = m / wp.R, r = m % wp.R, d = r / wp.PSI, bin = (r % wp.PSI) / pb;
        const int g = (k * wp.D + d) * a.nbins + bin;
        if (gm[g >> 5] >> (g & 31) & 1u) in += w;
    }
    all = block_reduce(all, [](double u, double v) { return u + v; }, red);
    in = block_reduce(in, [](double u, double v) { return u + v; }, red);
    if (tid == 0) a.mass[ai] = all > 0. ? in / all : 0.;
}

}  // namespace

// ------------------------------------------------------------------ launchers
hipError_t mlp_gemm(hipStream_t s, const void* A, int lda, const void* B, int ldb, float* C, int ldc, const float* bias,
                    int M, int N, int K) {
    if (M <= 0 || N <= 0) return hipSuccess;
    if (K % 32 || lda % 8 || ldb % 8) return hipErrorInvalidValue;
    dim3 g((N + 127) / 128, (M + 127) / 128);
    if (g.y > 65535) return hipErrorInvalidValue;
    k_mlp_gemm<<<g, 256, 0, s>>>((const bf16*)A, lda, (const bf16*)B, ldb, C, ldc, bias, M, N, K);
    return hipGetLastError();
}

hipError_t mlp_transpose(hipStream_t s, const void* in, bool in_f32, int R, int Cn, int ldi, void* out, int ldo) {
    if (R <= 0 || Cn <= 0) return hipSuccess;
    dim3 g((Cn + 31) / 32, (R + 31) / 32);
    if (in_f32) k_transpose_bf16<float><<<g, 256, 0, s>>>((const float*)in, R, Cn, ldi, (bf16*)out, ldo);
    else k_transpose_bf16<bf16><<<g, 256, 0, s>>>((const bf16*)in, R, Cn, ldi, (bf16*)out, ldo);
    return hipGetLastError();
}

hipError_t mlp_gelu_fwd(hipStream_t s, const float* a, int lda, void* h, int ldh, int rows, int cols) {
    const size_t n = (size_t)rows * cols;
    if (!n) return hipSuccess;
    k_gelu_fwd<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(a, lda, (bf16*)h, ldh, rows, cols);
    return hipGetLastError();
}

hipError_t mlp_gelu_bwd(hipStream_t s, float* dh, const float* a, void* daq, size_t n) {
    if (!n) return hipSuccess;
    k_gelu_bwd<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(dh, a, (bf16*)daq, n);
    return hipGetLastError();
}

hipError_t mlp_colsum(hipStream_t s, const float* d, int rows, int cols, int ld, float* db) {
    k_colsum<<<(cols + 255) / 256, 256, 0, s>>>(d, rows, cols, ld, db);
    return hipGetLastError();
}

hipError_t mlp_softmax_ce(hipStream_t s, const float* z, int ldz, int ncell, int B, const int* tcell, const float* tprob,
                          const int* tcnt, int tcap, const float* tshift, float lam, float* dz, void* dzq, int ldd,
                          float* loss) {
    k_softmax_ce<<<B, 256, 0, s>>>(z, ldz, ncell, B, tcell, tprob, tcnt, tcap, tshift, lam, dz, (bf16*)dzq, ldd, loss);
    return hipGetLastError();
}

hipError_t mlp_adam(hipStream_t s, float* w, float* m, float* v, const float* g, void* wq, size_t n, float lr, float b1,
                    float b2, float eps, float bc1, float bc2) {
    k_adam<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(w, m, v, g, (bf16*)wq, n, lr, b1, b2, eps, bc1, bc2);
    return hipGetLastError();
}

hipError_t mlp_to_bf16(hipStream_t s, const float* in, void* out, size_t n) {
    if (!n) return hipSuccess;
    k_to_bf16<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(in, (bf16*)out, n);
    return hipGetLastError();
}

hipError_t mlp_gather_rows(hipStream_t s, const void* src, const int* idx, int rows, int ld, void* dst) {
    if (rows <= 0) return hipSuccess;
    if (ld % 8) return hipErrorInvalidValue;
    k_gather_rows<<<rows, 128, 0, s>>>((const bf16*)src, idx, rows, ld, (bf16*)dst);
    return hipGetLastError();
}

hipError_t mlp_gather_targets(hipStream_t s, const int* idx, int rows, int tcap, const int* tcell, const float* tprob,
                              const int* tcnt, const float* tshift, int* ocell, float* oprob, int* ocnt, float* oshift) {
    if (rows <= 0) return hipSuccess;
    k_gather_targets<<<rows, 64, 0, s>>>(idx, rows, tcap, tcell, tprob, tcnt, tshift, ocell, oprob, ocnt, oshift);
    return hipGetLastError();
}

hipError_t mlp_features(hipStream_t s, int n, const float2* fmask, int F2f, const int* coef_n, const int* coef_ires,
                        int ncoef, const double* sigma2, const int* group, int oriF, void* x, int ldx) {
    if (n <= 0) return hipSuccess;
    if (2 * ncoef > ldx) return hipErrorInvalidValue;
    k_mlp_features<<<n, 256, 0, s>>>(fmask, F2f, coef_n, coef_ires, ncoef, sigma2, group, oriF, (bf16*)x, ldx);
    return hipGetLastError();
}

hipError_t proposal_targets(hipStream_t s, int n, const ProposalTargetArgs& a) {
    if (n <= 0) return hipSuccess;
    const size_t lds = (size_t)a.K * a.C * sizeof(float);
    if (lds > 64 * 1024 || a.C % a.nbins || a.PSI % a.nbins) return hipErrorInvalidValue;
    k_proposal_targets<<<n, 1024, lds, s>>>(a);
    return hipGetLastError();
}

hipError_t select_candidates(hipStream_t s, int n, const CandidateArgs& a) {
    if (n <= 0) return hipSuccess;
    const int ngroup = a.K * a.D * a.nbins;
    if ((ngroup + 31) / 32 > kMaxGroupWords || a.gwords < (ngroup + 31) / 32 || a.PSI % a.nbins) return hipErrorInvalidValue;
    k_select_candidates<<<n, 1024, 0, s>>>(a);
    return hipGetLastError();
}

hipError_t cand_coarse(hipStream_t s, int precision, int n, const CandCoarseArgs& a) {
    if (n <= 0) return hipSuccess;
    if (a.K16 % 32) return hipErrorInvalidValue;
    if (precision == 1) k_cand_coarse<__bf16><<<n, 256, 0, s>>>(a);
    else if (precision == 2) k_cand_coarse<_Float16><<<n, 256, 0, s>>>(a);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t audit_mass(hipStream_t s, int na, const AuditArgs& a) {
    if (na <= 0) return hipSuccess;
    if (a.wp.K * a.wp.T > kMaxKT) return hipErrorInvalidValue;
    k_audit_mass<<<na, 1024, 0, s>>>(a);
    return hipGetLastError();
}

}  // namespace rome
