// Mixed-precision arm of the coarse cross term (K2): BF16 or FP16 MFMA with FP32 accumulation.
//
//   C[m][j] = sum_k A16[m][k] * B16[j][k],   E[img][m][t] = At[img][m] - C[m][img*T+t] / (sF * s_img)
//
// A16 = sF * coarse projections (power-of-two scale over the whole table), B16 = s_img * Minv*CTF*X*S
// (power-of-two scale per image so its largest value is in [0.5, 1)): FP16 keeps 11 significant bits
// over the image's dynamic range, BF16 8 bits over FP32's range.  The norm term At stays FP32 and the
// per-image constant FP64, so only the cross term carries the reduced precision; the fine pass that
// decides poses, weights and the maps is unchanged (FP32 products, FP64 sums).
//
// v_mfma_f32_32x32x16_{bf16,f16}: lane l supplies A[l%32][8*(l/32) .. +7] and B[8*(l/32) .. +7][l%32];
// D as for the FP32 instruction (row 8*(r/4) + 4*(l/32) + r%4, col l%32).  K-contiguous operands so
// each lane loads its 8 values with one 16-byte LDS read.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_util.h"
#include "gemm_common.h"
#include "kernels.h"

namespace rome {

namespace {

template <typename T>
struct Lowp;
template <>
struct Lowp<__bf16> {
    typedef __bf16 v8 __attribute__((ext_vector_type(8)));
    static DEV f16v mfma(v8 a, v8 b, f16v c) { return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0); }
};
template <>
struct Lowp<_Float16> {
    typedef _Float16 v8 __attribute__((ext_vector_type(8)));
    static DEV f16v mfma(v8 a, v8 b, f16v c) { return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0); }
};

// Full pose term E: 256x256 output tile per block of 8 waves (4 along rotations x 2 along columns); each wave owns
// 64x128 = 2x4 accumulator blocks (128 accumulator registers, two waves per SIMD).  BK = 64: per
// 16-k step a wave reads 6 fragments from LDS for 8 MFMAs; the global loads of chunk c+1 (held in
// eight named 16-byte registers: an array here would live in scratch) are in flight while chunk c is
// multiplied by both waves of the SIMD.  Two LDS stages of 72 KB, one block per CU; after the main
// loop the same 144 KB stage the E rows.  (4x2 waves of 64x128 beat 2x4 waves of 128x64 by 15 % in
// the main loop at K = 512.)
template <typename T>
__global__ __launch_bounds__(512, 1) void k_gemm_lowp(LowpGemmArgs a) {
    constexpr int BM = 256, BN = 256, BK = 64, LDW = BK + 8;  // 144-byte LDS rows: conflict-free b128 reads
    typedef typename Lowp<T>::v8 v8;
    constexpr int TI = 2, TJ = 4;  // accumulator blocks per wave
    // operand stages [A|B][stage][row][k]; the epilogue reuses the whole array
    __shared__ __attribute__((aligned(16))) T ABs[2][2][BM][LDW];
    __shared__ int slot_min[epi_slots<BN>()];
    auto& As = ABs[0];
    auto& Bs = ABs[1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    int bm, bn;
    gemm_tile(blockIdx.x, a.Mpad / BM, a.Npad / BN, a.sband, bm, bn);
    const int m0 = bm * BM, n0 = bn * BN;
    const T* A = reinterpret_cast<const T*>(a.A);
    const T* B = reinterpret_cast<const T*>(a.B);
    f16v acc[TI][TJ];
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;
    const int nk = a.K16 / BK;
    // chunk = 256 rows x 64 k of each operand = 2048 16-byte pieces: 4 per thread per operand
    // (rows tid / 8 + 64 h, piece tid % 8)
    const int prow = tid >> 3, pcol = (tid & 7) * 8;
    const T* Ag = A + (size_t)(m0 + prow) * a.K16 + pcol;
    const T* Bg = B + (size_t)(n0 + prow) * a.K16 + pcol;
    const size_t q64 = (size_t)64 * a.K16;
#define LOWP_LOAD(kc)                                                        \
    do {                                                                             \
        ra0 = *reinterpret_cast<const uint4*>(Ag + (kc) * BK);                     \
        ra1 = *reinterpret_cast<const uint4*>(Ag + q64 + (kc) * BK);               \
        ra2 = *reinterpret_cast<const uint4*>(Ag + 2 * q64 + (kc) * BK);           \
        ra3 = *reinterpret_cast<const uint4*>(Ag + 3 * q64 + (kc) * BK);           \
        rb0 = *reinterpret_cast<const uint4*>(Bg + (kc) * BK);                     \
        rb1 = *reinterpret_cast<const uint4*>(Bg + q64 + (kc) * BK);               \
        rb2 = *reinterpret_cast<const uint4*>(Bg + 2 * q64 + (kc) * BK);           \
        rb3 = *reinterpret_cast<const uint4*>(Bg + 3 * q64 + (kc) * BK);           \
    } while (0)
#define LOWP_STORE(st)                                                       \
    do {                                                                             \
        *reinterpret_cast<uint4*>(&As[st][prow][pcol]) = ra0;                      \
        *reinterpret_cast<uint4*>(&As[st][prow + 64][pcol]) = ra1;                 \
        *reinterpret_cast<uint4*>(&As[st][prow + 128][pcol]) = ra2;                \
        *reinterpret_cast<uint4*>(&As[st][prow + 192][pcol]) = ra3;                \
        *reinterpret_cast<uint4*>(&Bs[st][prow][pcol]) = rb0;                      \
        *reinterpret_cast<uint4*>(&Bs[st][prow + 64][pcol]) = rb1;                 \
        *reinterpret_cast<uint4*>(&Bs[st][prow + 128][pcol]) = rb2;                \
        *reinterpret_cast<uint4*>(&Bs[st][prow + 192][pcol]) = rb3;                \
    } while (0)
#define LOWP_COMPUTE(st)                                                                                    \
    do {                                                                                                    \
        _Pragma("unroll") for (int kk = 0; kk < BK; kk += 16) {                                             \
            v8 av[TI], bv[TJ];                                                                              \
            _Pragma("unroll") for (int i = 0; i < TI; ++i) av[i] =                                          \
                *reinterpret_cast<const v8*>(&As[st][wm * 64 + i * 32 + (lane & 31)][kk + 8 * (lane >> 5)]); \
            _Pragma("unroll") for (int j = 0; j < TJ; ++j) bv[j] =                                          \
                *reinterpret_cast<const v8*>(&Bs[st][wn * 128 + j * 32 + (lane & 31)][kk + 8 * (lane >> 5)]); \
            _Pragma("unroll") for (int i = 0; i < TI; ++i)                                                  \
                _Pragma("unroll") for (int j = 0; j < TJ; ++j) acc[i][j] = Lowp<T>::mfma(av[i], bv[j], acc[i][j]); \
        }                                                                                                   \
    } while (0)
    uint4 ra0, ra1, ra2, ra3, rb0, rb1, rb2, rb3;
    LOWP_LOAD(0);
    LOWP_STORE(0);
    __syncthreads();
    for (int kc = 0; kc < nk; ++kc) {
        const int st = kc & 1;
        if (kc + 1 < nk) LOWP_LOAD(kc + 1);
        LOWP_COMPUTE(st);
        if (kc + 1 < nk) LOWP_STORE(st ^ 1);
        __syncthreads();
    }
#undef LOWP_LOAD
#undef LOWP_STORE
#undef LOWP_COMPUTE
    // the operand stages are free now: 8 waves x 128 columns x 36 floats of staging
    float* stage = reinterpret_cast<float*>(&ABs[0][0][0][0]) + wave * (TJ * 32 * 36);
    static_assert(sizeof(ABs) >= 8 * TJ * 32 * 36 * sizeof(float), "epilogue staging does not fit");
    gemm_epilogue_staged<true, TI, TJ, BN>(acc, m0 + wm * 64, n0 + wn * 128, n0, lane, tid, a.At, a.inv_scale, a.E,
                                           a.valid, a.emin_key, a.Mtot, a.Mpad, a.Ncols, a.T, slot_min, stage);
}

// List output: 256x128 tiles per block of 4 waves (one per SIMD, each 64x128 = 2x4 accumulator
// blocks), one 55 KB LDS stage (the next chunk waits in registers), two blocks per CU: one block's
// epilogue runs beside the other's MFMA loop.
template <typename T>
__global__ __launch_bounds__(256, 2) void k_gemm_lowp_list(LowpGemmArgs a) {
    constexpr int BM = 256, BN = 128, BK = 64, LDW = BK + 8;
    typedef typename Lowp<T>::v8 v8;
    constexpr int TI = 2, TJ = 4;
    __shared__ __attribute__((aligned(16))) T As[BM][LDW];
    __shared__ __attribute__((aligned(16))) T Bs[BN][LDW];
    __shared__ int slot_min[epi_slots<BN>()];
    __shared__ int slot_vb[epi_slots<BN>()];
    __shared__ float slot_lim[epi_slots<BN>()];
    __shared__ int slot_cnt[epi_slots<BN>()];
    const int tid = threadIdx.x, lane = tid & 63, wm = tid >> 6;
    int bm, bn;
    gemm_tile(blockIdx.x, a.Mpad / BM, a.Npad / BN, a.sband, bm, bn);
    const int m0 = bm * BM, n0 = bn * BN;
    const T* A = reinterpret_cast<const T*>(a.A);
    const T* B = reinterpret_cast<const T*>(a.B);
    f16v acc[TI][TJ];
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;
    const int nk = a.K16 / BK;
    // chunk: A 256 x 64 (8 pieces of 16 bytes per thread), B 128 x 64 (4 per thread); rows tid / 8 + 32 h
    const int prow = tid >> 3, pcol = (tid & 7) * 8;
    const T* Ag = A + (size_t)(m0 + prow) * a.K16 + pcol;
    const T* Bg = B + (size_t)(n0 + prow) * a.K16 + pcol;
    const size_t q32 = (size_t)32 * a.K16;
    uint4 ra0, ra1, ra2, ra3, ra4, ra5, ra6, ra7, rb0, rb1, rb2, rb3;
#define L2_LOAD(kc)                                                          \
    do {                                                                     \
        const T* ag = Ag + (kc) * BK;                                        \
        const T* bg = Bg + (kc) * BK;                                        \
        ra0 = *reinterpret_cast<const uint4*>(ag);                           \
        ra1 = *reinterpret_cast<const uint4*>(ag + q32);                     \
        ra2 = *reinterpret_cast<const uint4*>(ag + 2 * q32);                 \
        ra3 = *reinterpret_cast<const uint4*>(ag + 3 * q32);                 \
        ra4 = *reinterpret_cast<const uint4*>(ag + 4 * q32);                 \
        ra5 = *reinterpret_cast<const uint4*>(ag + 5 * q32);                 \
        ra6 = *reinterpret_cast<const uint4*>(ag + 6 * q32);                 \
        ra7 = *reinterpret_cast<const uint4*>(ag + 7 * q32);                 \
        rb0 = *reinterpret_cast<const uint4*>(bg);                           \
        rb1 = *reinterpret_cast<const uint4*>(bg + q32);                     \
        rb2 = *reinterpret_cast<const uint4*>(bg + 2 * q32);                 \
        rb3 = *reinterpret_cast<const uint4*>(bg + 3 * q32);                 \
    } while (0)
#define L2_STORE()                                                           \
    do {                                                                     \
        *reinterpret_cast<uint4*>(&As[prow][pcol]) = ra0;                    \
        *reinterpret_cast<uint4*>(&As[prow + 32][pcol]) = ra1;               \
        *reinterpret_cast<uint4*>(&As[prow + 64][pcol]) = ra2;               \
        *reinterpret_cast<uint4*>(&As[prow + 96][pcol]) = ra3;               \
        *reinterpret_cast<uint4*>(&As[prow + 128][pcol]) = ra4;              \
        *reinterpret_cast<uint4*>(&As[prow + 160][pcol]) = ra5;              \
        *reinterpret_cast<uint4*>(&As[prow + 192][pcol]) = ra6;              \
        *reinterpret_cast<uint4*>(&As[prow + 224][pcol]) = ra7;              \
        *reinterpret_cast<uint4*>(&Bs[prow][pcol]) = rb0;                    \
        *reinterpret_cast<uint4*>(&Bs[prow + 32][pcol]) = rb1;               \
        *reinterpret_cast<uint4*>(&Bs[prow + 64][pcol]) = rb2;               \
        *reinterpret_cast<uint4*>(&Bs[prow + 96][pcol]) = rb3;               \
    } while (0)
    L2_LOAD(0);
    L2_STORE();
    __syncthreads();
    for (int kc = 0; kc < nk; ++kc) {
        if (kc + 1 < nk) L2_LOAD(kc + 1);
#pragma unroll
        for (int kk = 0; kk < BK; kk += 16) {
            v8 av[TI], bv[TJ];
#pragma unroll
            for (int i = 0; i < TI; ++i)
                av[i] = *reinterpret_cast<const v8*>(&As[wm * 64 + i * 32 + (lane & 31)][kk + 8 * (lane >> 5)]);
#pragma unroll
            for (int j = 0; j < TJ; ++j)
                bv[j] = *reinterpret_cast<const v8*>(&Bs[j * 32 + (lane & 31)][kk + 8 * (lane >> 5)]);
#pragma unroll
            for (int i = 0; i < TI; ++i)
#pragma unroll
                for (int j = 0; j < TJ; ++j) acc[i][j] = Lowp<T>::mfma(av[i], bv[j], acc[i][j]);
        }
        __syncthreads();
        if (kc + 1 < nk) {
            L2_STORE();
            __syncthreads();
        }
    }
#undef L2_LOAD
#undef L2_STORE
    gemm_epilogue_list<true, TI, TJ, BN>(acc, m0 + wm * 64, n0, n0, lane, tid, a.At, a.inv_scale, a.emin_key, a.Mtot,
                                         a.Mpad, a.Ncols, a.T, a.list, slot_min, slot_vb, slot_lim, slot_cnt);
}

// sF: power of two that brings max |FcT| into [0.5, 1) (exact rescaling)
DEV float table_scale(const int* absmax_bits) {
    const float mx = __int_as_float(*absmax_bits);
    return mx > 0.f ? ldexpf(1.f, -ilogbf(mx) - 1) : 1.f;
}

// A16[m][k] = sF * FcT[k][m] for k < Kfp = 2*Pc_pad, zero beyond and for rows m >= Mtot.
template <typename T>
__global__ void k_fct_to_lowp(const float* __restrict__ FcT, int Kfp, int Mpad, int Mtot,
                              const int* __restrict__ absmax_bits, T* __restrict__ A16, int K16) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)Mpad * K16) return;
    const float sF = table_scale(absmax_bits);
    const int m = (int)(idx / K16), k = (int)(idx % K16);
    const float v = (m < Mtot && k < Kfp) ? FcT[(size_t)k * Mpad + m] * sF : 0.f;
    A16[idx] = static_cast<T>(v);
}

// max |x| over an array as float bits (non-negative floats order like their bit patterns); *out starts at 0
__global__ void k_absmax(const float* __restrict__ x, size_t n, int* __restrict__ out) {
    __shared__ float red[16];
    float m = 0.f;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        m = fmaxf(m, fabsf(x[i]));
    m = block_reduce(m, [](float u, float v) { return fmaxf(u, v); }, red);
    if (threadIdx.x == 0) atomicMax(out, __float_as_int(m));
}

// W16[img*T + t][k]: s_img * Minv*CTF*X*S (Re at k = p, Im at k = Pc_pad + p, zero in the padding);
// one block per image.  s_img is the power of two that brings max |Minv*CTF*X| into [0.5, 1).
template <typename T>
__global__ void k_prep_lowp(PrepLowpArgs a) {
    __shared__ float red[16];
    const int img = blockIdx.x;
    const int g = a.group[img];
    const double* s2 = a.sigma2 + (size_t)g * a.ori_Fsize;
    const float2* fm = a.fmask + (size_t)img * a.F2f;
    const float* fc = a.fctf + (size_t)img * a.F2f;
    float mx = 0.f;
    for (int p = threadIdx.x; p < a.Pc; p += blockDim.x) {
        const int nf = a.c_nfine[p];
        const float2 X = fm[nf];
        const float mc = (float)(1.0 / (a.sigma2_fudge * s2[a.c_ires[p]])) * (a.ctf_in_diff ? fc[nf] : 1.f);
        mx = fmaxf(mx, fabsf(mc) * sqrtf(X.x * X.x + X.y * X.y));
    }
    mx = block_reduce(mx, [](float u, float v) { return fmaxf(u, v); }, red);
    const float s = mx > 0.f ? ldexpf(1.f, -ilogbf(mx) - 1) : 1.f;
    if (threadIdx.x == 0) a.inv_scale[img] = 1.f / (s * table_scale(a.fct_absmax));
    T* W = reinterpret_cast<T*>(a.W16) + (size_t)img * a.T * a.K16;
    for (int p = threadIdx.x; p < a.Pc; p += blockDim.x) {
        const int nf = a.c_nfine[p];
        const float2 X = fm[nf];
        const float mc = (float)(1.0 / (a.sigma2_fudge * s2[a.c_ires[p]])) * (a.ctf_in_diff ? fc[nf] : 1.f) * s;
        for (int t = 0; t < a.T; ++t) {
            const float2 XS = cmul(X, a.Sc[(size_t)t * a.Pc + p]);
            W[(size_t)t * a.K16 + p] = static_cast<T>(mc * XS.x);
            W[(size_t)t * a.K16 + a.Pc_pad + p] = static_cast<T>(mc * XS.y);
        }
    }
    // padding columns: [Pc, Pc_pad), [Pc_pad + Pc, 2 Pc_pad), [2 Pc_pad, K16)
    const int npad = a.K16 - 2 * a.Pc;
    for (int q = threadIdx.x; q < a.T * npad; q += blockDim.x) {
        const int t = q / npad, j = q - t * npad;
        const int d = a.Pc_pad - a.Pc;
        const int k = j < d ? a.Pc + j : (j < 2 * d ? a.Pc_pad + a.Pc + (j - d) : 2 * a.Pc_pad + (j - 2 * d));
        W[(size_t)t * a.K16 + k] = static_cast<T>(0.f);
    }
}

}  // namespace

hipError_t launch_gemm_lowp(hipStream_t s, int precision, const LowpGemmArgs& a) {
    if (a.Mpad % 256 || a.Npad % 256 || a.K16 % 64 || a.K16 <= 0) return hipErrorInvalidValue;
    const int tiles = (a.Mpad / 256) * (a.Npad / 256);
    const bool list = a.list.count != nullptr;
    if (!list && !a.E) return hipErrorInvalidValue;
    if (precision == 1) {
        if (list) k_gemm_lowp_list<__bf16><<<2 * tiles, 256, 0, s>>>(a);
        else k_gemm_lowp<__bf16><<<tiles, 512, 0, s>>>(a);
    } else if (precision == 2) {
        if (list) k_gemm_lowp_list<_Float16><<<2 * tiles, 256, 0, s>>>(a);
        else k_gemm_lowp<_Float16><<<tiles, 512, 0, s>>>(a);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_fct_to_lowp(hipStream_t s, int precision, const float* FcT, int Kfp, int Mpad, int Mtot,
                              const int* fct_absmax, void* A16, int K16) {
    const size_t tot = (size_t)Mpad * K16;
    const unsigned blocks = (unsigned)((tot + 255) / 256);
    if (precision == 1)
        k_fct_to_lowp<__bf16><<<blocks, 256, 0, s>>>(FcT, Kfp, Mpad, Mtot, fct_absmax, (__bf16*)A16, K16);
    else if (precision == 2)
        k_fct_to_lowp<_Float16><<<blocks, 256, 0, s>>>(FcT, Kfp, Mpad, Mtot, fct_absmax, (_Float16*)A16, K16);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_absmax(hipStream_t s, const float* x, size_t n, int* out) {
    k_absmax<<<512, 256, 0, s>>>(x, n, out);
    return hipGetLastError();
}

hipError_t launch_prep_lowp(hipStream_t s, int precision, int n, const PrepLowpArgs& a) {
    if (2 * a.Pc_pad > a.K16 || a.Pc > a.Pc_pad) return hipErrorInvalidValue;
    if (precision == 1) k_prep_lowp<__bf16><<<n, 256, 0, s>>>(a);
    else if (precision == 2) k_prep_lowp<_Float16><<<n, 256, 0, s>>>(a);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

}  // namespace rome
