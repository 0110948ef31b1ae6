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
    if (n0 >= a.Ncols) return;  // a tile of padding columns only (block-uniform, before any barrier)
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
    if (!gemm_tile_grouped(blockIdx.x, a.Mpad / BM, a.Npad / BN, a.sband, bm, bn)) return;
    const int m0 = bm * BM, n0 = bn * BN;
    if (n0 >= a.Ncols) return;  // a tile of padding columns only (block-uniform, before any barrier)
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
    gemm_epilogue_list<true, TI, TJ, BN>(acc, ContigRows{m0 + wm * 64, lane >> 5}, n0, n0, lane, tid, a.At, a.inv_scale, a.emin_key, a.Mtot,
                                         a.Mpad, a.Ncols, a.T, a.list, slot_min, slot_vb, slot_lim, slot_cnt);
}

// Rows of the paired cross term: block 0 = rotations m of the half-table rows, block 1 = m + psih.
// The tile's At rows and ln pdf_direction sit in LDS (at: [image slot][kTileRows], lp: [kTileRows],
// tile-local row = 128 i + half-table row - m0), so the epilogue needs no per-row global addresses.
struct PairRows {
    static constexpr bool kPadding = false;  // Mh is a multiple of the tile
    static constexpr bool kPaired = true;    // block 0 holds P, block 1 holds Q
    static constexpr int kTileRows = 256;
    int m[4], psih, l0;                       // rotation of the lane's quad q of block 0; its local row
    const float *at, *lp;
    DEV PairRows(int hbase, int lrow, int ph, const float* at_, const float* lp_) : psih(ph), l0(lrow), at(at_),
                                                                                    lp(lp_) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int hh = hbase + lrow + 8 * q;
            const unsigned uh = (unsigned)hh, up = (unsigned)ph;  // non-negative: the shorter unsigned division
            m[q] = (int)((uh / up) * 2 * up + uh % up);
        }
    }
    DEV int operator()(int i, int q) const { return m[q] + i * psih; }
    DEV int local(int i, int q) const { return 128 * i + l0 + 8 * q; }
};

// E output of the paired cross term (block 0 of acc holds P, block 1 Q): E[img][t][m] for both
// rotations of each pair (16-byte stores of four consecutive rotations), and the per-image minimum of
// E over rotations with a prior.
template <int NI, int NJ, int TN, typename Rows>
DEV void gemm_epilogue_e(const f16v (&acc)[NI][NJ], const Rows& rows, int c0, int n0, int lane, int tid,
                         const float* __restrict__ At, const float* __restrict__ inv_scale,
                         const float* __restrict__ lpd_row, float* __restrict__ E, int* __restrict__ emin_key,
                         int Mtot, int Mpad, int Ncols, int T, int* slot_min) {
    const int img0 = n0 / T;
    const int last = min(n0 + TN, Ncols) - 1;
    const int nslot = last >= n0 ? last / T - img0 + 1 : 0;
    for (int q = tid; q < nslot; q += blockDim.x) slot_min[q] = INT_MAX;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
        const int col = c0 + j * 32 + (lane & 31);
        if (col >= Ncols) continue;
        const int img = col / T, t = col - img * T;
        const float s = inv_scale[img];
        float emin = INFINITY;
#pragma unroll
        for (int i = 0; i < NI; ++i)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int base = rows(i, q);
                const float4 lp = *reinterpret_cast<const float4*>(lpd_row + base);
                const float4 at = *reinterpret_cast<const float4*>(At + (size_t)img * Mpad + base);
                const float lpv[4] = {lp.x, lp.y, lp.z, lp.w}, atv[4] = {at.x, at.y, at.z, at.w};
                float ev[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const float P = acc[0][j][4 * q + u], Q = acc[1][j][4 * q + u];
                    ev[u] = e_value<true>(atv[u], i == 0 ? P + Q : P - Q, s);
                    emin = fminf(emin, lpv[u] > -INFINITY ? ev[u] : INFINITY);
                }
                *reinterpret_cast<float4*>(E + ((size_t)img * T + t) * Mtot + base) =
                    make_float4(ev[0], ev[1], ev[2], ev[3]);
            }
        if (emin < INFINITY) atomicMin(&slot_min[img - img0], float_order_key(emin));
    }
    __syncthreads();
    for (int q = tid; q < nslot; q += blockDim.x)
        if (slot_min[q] != INT_MAX) atomicMin(&emin_key[img0 + q], slot_min[q]);
}

// Paired cross term: 128 half-table rows x 128 columns per block of 4 waves (one per SIMD), each wave
// 32 half-table rows x 128 columns with P and Q accumulators (2 x 4 blocks of
// v_mfma_f32_32x32x16_{f16,bf16}); 32-k chunks of both halves (Re and Im) of both operands in one
// 40 KB LDS stage, the next chunk waiting in registers; two blocks per CU.  The epilogue turns P, Q
// into the two rotations' cross terms P + Q, P - Q and writes the list (LIST) or E.
constexpr int kPqBN = 128;         // columns per block: 64 or 128
// one 16-byte global load per lane straight into LDS: lane L writes lds + 16 L (lds wave-uniform)
DEV void lds_load16(const void* g, void* lds) {
    __builtin_amdgcn_global_load_lds(g, (__attribute__((address_space(3))) void*)lds, 16, 0, 0);
}
DEV void wait_vmem_all() { __builtin_amdgcn_s_waitcnt(0x0F70); }  // vmcnt(0), expcnt / lgkmcnt untouched
constexpr int kPqBlocksPerCU = 2;  // blocks per CU the register budget is sized for
template <typename T, bool LIST>
__global__ __launch_bounds__(256, kPqBlocksPerCU) void k_gemm_lowp_pq(LowpGemmArgs a) {
    constexpr int BM = 128, BN = kPqBN, BK = 32, LDW = BK + 8;  // 80-byte LDS rows: conflict-free b128 reads
    typedef typename Lowp<T>::v8 v8;
    constexpr int TJ = BN / 32;
    static_assert(BN == 64 || BN == 128, "column tile");
    __shared__ __attribute__((aligned(16))) T As[2][BM][LDW];  // [Re | Im][half-table row][k]
    __shared__ __attribute__((aligned(16))) T Bs[2][BN][LDW];  // [Re | Im][column][k]
    __shared__ int slot_min[epi_slots<BN>()];
    __shared__ int slot_vb[LIST ? epi_slots<BN>() : 1];
    __shared__ float slot_lim[LIST ? epi_slots<BN>() : 1];
    __shared__ int slot_cnt[LIST ? epi_slots<BN>() : 1];
    const int tid = threadIdx.x, lane = tid & 63, wm = tid >> 6;
    int bm, bn;
    if (!gemm_tile_grouped(blockIdx.x, a.Mh / BM, a.Npad / BN, a.sband, bm, bn)) return;
    const int m0 = bm * BM, n0 = bn * BN;
    if (n0 >= a.Ncols) return;  // a tile of padding columns only (block-uniform, before any barrier)
    // the list epilogue's rows of At (one per image slot of the tile) and of ln pdf_direction: with at most
    // three image slots, wave w loads row w straight into LDS before the main loop (no registers held,
    // no global round trip after it); otherwise they are staged after the loop through the operand stage
    __shared__ __attribute__((aligned(16))) float AtD[4][256];
    const int img0 = n0 / a.T, nsl = (min(n0 + BN, a.Ncols) - 1) / a.T - img0 + 1;
    const bool at_early = LIST && nsl < 4;
    if (at_early && wm <= nsl) {
        const int lr = lane * 4, hh = m0 + (lr & 127);
        const unsigned uh = (unsigned)hh, up = (unsigned)a.psih;
        const int m = (int)((uh / up) * 2 * up + uh % up) + (lr >> 7) * a.psih;
        lds_load16(wm < nsl ? static_cast<const void*>(a.At + (size_t)(img0 + wm) * a.Mpad + m)
                            : static_cast<const void*>(a.list.lpd + m),
                   &AtD[wm][0]);
    }
    const T* A = reinterpret_cast<const T*>(a.A);
    const T* B = reinterpret_cast<const T*>(a.B);
    f16v acc[2][TJ];  // [0]: P = sum Re Re, [1]: Q = sum Im Im
#pragma unroll
    for (int j = 0; j < TJ; ++j)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[0][j][r] = acc[1][j][r] = 0.f;
    const int KH = a.K16 / 2, nk = KH / BK;
    // chunk: per operand 128 rows x (32 Re + 32 Im) k = 1024 pieces of 16 bytes, 4 per thread:
    // piece tid + 256 v -> half (v >> 1), row tid / 4 + 64 (v & 1), k offset 8 (tid % 4)
    const int prow = tid >> 2, pcol = (tid & 3) * 8;
    const T* Ag = A + (size_t)(m0 + prow) * a.K16 + pcol;
    const T* Bg = B + (size_t)(n0 + prow) * a.K16 + pcol;
    const size_t r64 = (size_t)64 * a.K16;
    uint4 ra0, ra1, ra2, ra3, rb0, rb1, rb2, rb3;
#define PQ_LOAD(kc)                                                          \
    do {                                                                     \
        const T* ag = Ag + (kc) * BK;                                        \
        const T* bg = Bg + (kc) * BK;                                        \
        ra0 = *reinterpret_cast<const uint4*>(ag);                           \
        ra1 = *reinterpret_cast<const uint4*>(ag + r64);                     \
        ra2 = *reinterpret_cast<const uint4*>(ag + KH);                      \
        ra3 = *reinterpret_cast<const uint4*>(ag + r64 + KH);                \
        rb0 = *reinterpret_cast<const uint4*>(bg);                           \
        if constexpr (BN == 128) rb1 = *reinterpret_cast<const uint4*>(bg + r64); \
        rb2 = *reinterpret_cast<const uint4*>(bg + KH);                      \
        if constexpr (BN == 128) rb3 = *reinterpret_cast<const uint4*>(bg + r64 + KH); \
    } while (0)
#define PQ_STORE()                                                           \
    do {                                                                     \
        *reinterpret_cast<uint4*>(&As[0][prow][pcol]) = ra0;                 \
        *reinterpret_cast<uint4*>(&As[0][prow + 64][pcol]) = ra1;            \
        *reinterpret_cast<uint4*>(&As[1][prow][pcol]) = ra2;                 \
        *reinterpret_cast<uint4*>(&As[1][prow + 64][pcol]) = ra3;            \
        *reinterpret_cast<uint4*>(&Bs[0][prow][pcol]) = rb0;                 \
        if constexpr (BN == 128) *reinterpret_cast<uint4*>(&Bs[0][prow + 64][pcol]) = rb1; \
        *reinterpret_cast<uint4*>(&Bs[1][prow][pcol]) = rb2;                 \
        if constexpr (BN == 128) *reinterpret_cast<uint4*>(&Bs[1][prow + 64][pcol]) = rb3; \
    } while (0)
    PQ_LOAD(0);
    PQ_STORE();
    __syncthreads();
    for (int kc = 0; kc < nk; ++kc) {
        if (kc + 1 < nk) PQ_LOAD(kc + 1);
#pragma unroll
        for (int kk = 0; kk < BK; kk += 16) {
            const int ko = kk + 8 * (lane >> 5);
            const v8 are = *reinterpret_cast<const v8*>(&As[0][wm * 32 + (lane & 31)][ko]);
            const v8 aim = *reinterpret_cast<const v8*>(&As[1][wm * 32 + (lane & 31)][ko]);
            v8 bre[TJ], bim[TJ];
#pragma unroll
            for (int j = 0; j < TJ; ++j) {
                bre[j] = *reinterpret_cast<const v8*>(&Bs[0][j * 32 + (lane & 31)][ko]);
                bim[j] = *reinterpret_cast<const v8*>(&Bs[1][j * 32 + (lane & 31)][ko]);
            }
            // the MFMA burst at raised wave priority: the other block of the CU, often in its VALU-heavy
            // epilogue, then does not delay the issue of these (3.44 -> 3.37 ms; the whole k loop at raised
            // priority measured 3.50 ms)
            __builtin_amdgcn_s_setprio(1);
#pragma unroll
            for (int j = 0; j < TJ; ++j) {
                acc[0][j] = Lowp<T>::mfma(are, bre[j], acc[0][j]);
                acc[1][j] = Lowp<T>::mfma(aim, bim[j], acc[1][j]);
            }
            __builtin_amdgcn_s_setprio(0);
        }
        __syncthreads();
        if (kc + 1 < nk) {
            PQ_STORE();
            __syncthreads();
        }
    }
#undef PQ_LOAD
#undef PQ_STORE
    // the tile's 256 rotations (128 pairs): At of its images and ln pdf_direction in LDS (the operand
    // stage is free: the loop ended with a barrier)
    float* Ats = at_early ? &AtD[0][0] : reinterpret_cast<float*>(&As[0][0][0]);
    if (at_early) {
        wait_vmem_all();  // this wave's row has landed; the barrier makes every wave's visible
        __syncthreads();
        for (int v = tid; v < nsl * 64; v += 256) {  // +inf on rows without a prior
            const int sl = v >> 6, lr = (v & 63) * 4;
            float4 x = *reinterpret_cast<const float4*>(Ats + sl * 256 + lr);
            const float4 l = *reinterpret_cast<const float4*>(Ats + nsl * 256 + lr);
            x.x += l.x > -INFINITY ? 0.f : INFINITY;
            x.y += l.y > -INFINITY ? 0.f : INFINITY;
            x.z += l.z > -INFINITY ? 0.f : INFINITY;
            x.w += l.w > -INFINITY ? 0.f : INFINITY;
            *reinterpret_cast<float4*>(Ats + sl * 256 + lr) = x;
        }
    } else
    for (int v = tid; v < (nsl + 1) * 64; v += 256) {
        const int sl = v >> 6, lr = (v & 63) * 4, hh = m0 + (lr & 127);
        const int m = (hh / a.psih) * 2 * a.psih + hh % a.psih + (lr >> 7) * a.psih;
        const float4 l = *reinterpret_cast<const float4*>(a.list.lpd + m);
        float4 x = l;
        if (sl < nsl) {  // At, +inf on rows without a prior: their E is +inf, outside every minimum and list
            x = *reinterpret_cast<const float4*>(a.At + (size_t)(img0 + sl) * a.Mpad + m);
            x.x += l.x > -INFINITY ? 0.f : INFINITY;
            x.y += l.y > -INFINITY ? 0.f : INFINITY;
            x.z += l.z > -INFINITY ? 0.f : INFINITY;
            x.w += l.w > -INFINITY ? 0.f : INFINITY;
        }
        *reinterpret_cast<float4*>(Ats + sl * 256 + lr) = x;
    }
    __syncthreads();
    // the cross terms of rotation m and of m + psih are P + Q and P - Q
    const PairRows rows(m0, wm * 32 + 4 * (lane >> 5), a.psih, Ats, Ats + nsl * 256);
    if constexpr (LIST)
        gemm_epilogue_list<true, 2, TJ, BN>(acc, rows, n0, n0, lane, tid, a.At, a.inv_scale, a.emin_key, a.Mtot,
                                            a.Mpad, a.Ncols, a.T, a.list, slot_min, slot_vb, slot_lim, slot_cnt);
    else
        gemm_epilogue_e<2, TJ, BN>(acc, rows, n0, n0, lane, tid, a.At, a.inv_scale, a.list.lpd, a.E, a.emin_key,
                                   a.Mtot, a.Mpad, a.Ncols, a.T, slot_min);
}

// sF: power of two that brings max |FcT| into [0.5, 1) (exact rescaling)
DEV float table_scale(const int* absmax_bits) {
    const float mx = __int_as_float(*absmax_bits);
    return mx > 0.f ? ldexpf(1.f, -ilogbf(mx) - 1) : 1.f;
}

// A16[row][k] = sF * FcT[.][m] for rotation m = rowmap(row): Re of pixel p at k = p, Im at k = K16/2 + p
// (p < Pc_pad), zero elsewhere and for rows past the table.  psih = 0: row = m (rows m >= Mtot zero);
// psih > 0: the half table of the paired cross term, row h = rotation (h / psih) * 2 psih + h % psih.
template <typename T>
__global__ void k_fct_to_lowp(const float* __restrict__ FcT, int Pc_pad, int Mpad, int Mtot, int rows, int psih,
                              const int* __restrict__ absmax_bits, T* __restrict__ A16, T* __restrict__ A16lo, int K16) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)rows * K16) return;
    const float sF = table_scale(absmax_bits);
    const int row = (int)(idx / K16), k = (int)(idx % K16), KH = K16 / 2;
    const int m = psih > 0 ? (row / psih) * 2 * psih + row % psih : row;
    const int kf = k < Pc_pad ? k : (k >= KH && k < KH + Pc_pad ? Pc_pad + (k - KH) : -1);
    const float v = (m < Mtot && kf >= 0) ? FcT[(size_t)kf * Mpad + m] * sF : 0.f;
    const T hi = static_cast<T>(v);
    A16[idx] = hi;
    if (A16lo) A16lo[idx] = static_cast<T>(v - static_cast<float>(hi));  // v = hi + lo to 2x the format's bits
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

// W16[img*T + t][k]: s_img * Minv*CTF*X*S (Re at k = p, Im at k = K16/2 + p, zero in the padding);
// one block per image.  s_img is the power of two that brings max |Minv*CTF*X| into [0.5, 1).
template <typename T>
__global__ void k_prep_lowp(PrepLowpArgs a) {
    __shared__ float red[16];
    T* Wlo = a.W16lo ? reinterpret_cast<T*>(a.W16lo) + (size_t)blockIdx.x * a.T * a.K16 : nullptr;
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
            const float vr = mc * XS.x, vi = mc * XS.y;
            const T hr = static_cast<T>(vr), hi = static_cast<T>(vi);
            W[(size_t)t * a.K16 + p] = hr;
            W[(size_t)t * a.K16 + a.K16 / 2 + p] = hi;
            if (Wlo) {
                Wlo[(size_t)t * a.K16 + p] = static_cast<T>(vr - static_cast<float>(hr));
                Wlo[(size_t)t * a.K16 + a.K16 / 2 + p] = static_cast<T>(vi - static_cast<float>(hi));
            }
        }
    }
    // padding columns: [Pc, K16/2), [K16/2 + Pc, K16)
    const int npad = a.K16 - 2 * a.Pc;
    for (int q = threadIdx.x; q < a.T * npad; q += blockDim.x) {
        const int t = q / npad, j = q - t * npad;
        const int d = a.K16 / 2 - a.Pc;
        const int k = j < d ? a.Pc + j : a.K16 / 2 + a.Pc + (j - d);
        W[(size_t)t * a.K16 + k] = static_cast<T>(0.f);
        if (Wlo) Wlo[(size_t)t * a.K16 + k] = static_cast<T>(0.f);
    }
}

// Split three-product recompute of the marked coarse samples of the mixed arms (one block per image).
// Both operands are held as hi + lo in the arm's format (hi = the operand the coarse GEMM used, lo = the
// rounding of the rest), and C = A_hi B_hi + A_hi B_lo + A_lo B_hi (the A_lo B_lo term is below the
// format's squared precision): 2^-16 relative per product in BF16 instead of 2^-8, on MFMA.  The
// image's tasks (rotations m, sorted) go 32 at a time: rows = those rotations (half-table rows with the
// sign of the mirrored Im part in the paired layout), columns = every translation in tiles of 32 (one
// wave per tile); the epilogue keeps the (m, t) of the image's marked samples: D2[b] = At - C / (sF s_img)
// + the image constant, for block b of the task list.
template <typename T>
__global__ __launch_bounds__(256) void k_recompute_split(RecomputeArgs a) {
    typedef typename Lowp<T>::v8 v8;
    __shared__ short bidx[32][kMaxT];  // tile row x translation -> block index - tile base (-1: unmarked)
    __shared__ int rowm[32];
    __shared__ float rowsg[32];
    const int img = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const long long q0 = a.task_off[img], q1 = a.task_off[img + 1];
    const int KH = a.K16 / 2, Tn = a.T, ntt = (Tn + 31) / 32;
    const T* Wh = reinterpret_cast<const T*>(a.Whi) + (size_t)img * Tn * a.K16;
    const T* Wl = reinterpret_cast<const T*>(a.Wlo) + (size_t)img * Tn * a.K16;
    const T* Ahi = reinterpret_cast<const T*>(a.Ahi);
    const T* Alo = reinterpret_cast<const T*>(a.Alo);
    const float inv = a.inv_scale[img];
    const float* At = a.At + (size_t)img * a.Mpad;
    const double cimg = a.C2c[img];
    const v8 zero = {};
    for (long long qt = q0; qt < q1; qt += 32) {
        const int nr = (int)min(32LL, q1 - qt);
        const long long base = a.task_blk0[qt];
        for (int i = tid; i < 32 * kMaxT; i += 256) (&bidx[0][0])[i] = -1;
        if (tid < 32) {
            const int m = tid < nr ? a.task_m[qt + tid] : -1;
            // paired layout: the rotation with psi >= 180 deg has the conjugate of psi - 180's projection
            const bool mirrored = m >= 0 && a.psih > 0 && m % (2 * a.psih) >= a.psih;
            rowm[tid] = m;
            rowsg[tid] = mirrored ? -1.f : 1.f;
        }
        __syncthreads();
        for (int r = wave; r < nr; r += 4) {
            const long long b0 = a.task_blk0[qt + r];
            const int nt = a.task_nt[qt + r];
            for (int j = lane; j < nt; j += 64) bidx[r][a.blk_t[b0 + j]] = (short)(b0 + j - base);
        }
        __syncthreads();
        const int r = lane & 31;
        const int m = rowm[r];
        int h = m;  // the rotation's row of the (half) table
        if (m >= 0 && a.psih > 0) {
            const int P2 = 2 * a.psih, qd = m / P2, ps = m - qd * P2;
            h = qd * a.psih + (ps >= a.psih ? ps - a.psih : ps);
        }
        const T* ah_row = Ahi + (size_t)(m >= 0 ? h : 0) * a.K16 + (lane >> 5) * 8;
        const T* al_row = Alo + (size_t)(m >= 0 ? h : 0) * a.K16 + (lane >> 5) * 8;
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
                P = Lowp<T>::mfma(ah, bh, P);
                P = Lowp<T>::mfma(ah, bl, P);
                P = Lowp<T>::mfma(al, bh, P);
            }
            for (int k0 = KH; k0 < a.K16; k0 += 16) {
                const v8 ah = m >= 0 ? *reinterpret_cast<const v8*>(ah_row + k0) : zero;
                const v8 al = m >= 0 ? *reinterpret_cast<const v8*>(al_row + k0) : zero;
                const v8 bh = tok ? *reinterpret_cast<const v8*>(bh_row + k0) : zero;
                const v8 bl = tok ? *reinterpret_cast<const v8*>(bl_row + k0) : zero;
                Q = Lowp<T>::mfma(ah, bh, Q);
                Q = Lowp<T>::mfma(ah, bl, Q);
                Q = Lowp<T>::mfma(al, bh, Q);
            }
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const int row = 8 * (j >> 2) + 4 * (lane >> 5) + (j & 3);
                const int tc = tt * 32 + (lane & 31);
                if (row >= nr || tc >= Tn) continue;
                const int rb = bidx[row][tc];
                if (rb < 0) continue;
                const float c = P[j] + rowsg[row] * Q[j];
                const float e = __builtin_fmaf(-c, inv, At[rowm[row]]);
                a.D2[base + rb] = (double)e + cimg;
            }
        }
        __syncthreads();
    }
}

}  // namespace

hipError_t launch_gemm_lowp(hipStream_t s, int precision, const LowpGemmArgs& a) {
    if (a.Mpad % 256 || a.Npad % 256 || a.K16 % 64 || a.K16 <= 0) return hipErrorInvalidValue;
    if (a.psih > 0) {  // paired cross term
        if (a.psih % 4 || a.Mh % 128 || 2 * a.Mh != a.Mtot || a.Mtot % 4 || a.T < 8 || !a.list.lpd)
            return hipErrorInvalidValue;
        const bool lst = a.list.count != nullptr;
        if (!lst && (!a.E || !a.list.lpd)) return hipErrorInvalidValue;
        const unsigned tiles = gemm_grouped_blocks(a.Mh / 128, a.Npad / kPqBN, a.sband);
        if (precision == 1) {
            if (lst) k_gemm_lowp_pq<__bf16, true><<<tiles, 256, 0, s>>>(a);
            else k_gemm_lowp_pq<__bf16, false><<<tiles, 256, 0, s>>>(a);
        } else if (precision == 2) {
            if (lst) k_gemm_lowp_pq<_Float16, true><<<tiles, 256, 0, s>>>(a);
            else k_gemm_lowp_pq<_Float16, false><<<tiles, 256, 0, s>>>(a);
        } else {
            return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    const int tiles = (a.Mpad / 256) * (a.Npad / 256);
    const unsigned list_blocks = gemm_grouped_blocks(a.Mpad / 256, a.Npad / 128, a.sband);
    const bool list = a.list.count != nullptr;
    if (!list && !a.E) return hipErrorInvalidValue;
    if (precision == 1) {
        if (list) k_gemm_lowp_list<__bf16><<<list_blocks, 256, 0, s>>>(a);
        else k_gemm_lowp<__bf16><<<tiles, 512, 0, s>>>(a);
    } else if (precision == 2) {
        if (list) k_gemm_lowp_list<_Float16><<<list_blocks, 256, 0, s>>>(a);
        else k_gemm_lowp<_Float16><<<tiles, 512, 0, s>>>(a);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_fct_to_lowp(hipStream_t s, int precision, const float* FcT, int Pc_pad, int Mpad, int Mtot,
                              int rows, int psih, const int* fct_absmax, void* A16, void* A16lo, int K16) {
    if (K16 % 2 || 2 * Pc_pad > K16) return hipErrorInvalidValue;
    const size_t tot = (size_t)rows * K16;
    const unsigned blocks = (unsigned)((tot + 255) / 256);
    if (precision == 1)
        k_fct_to_lowp<__bf16><<<blocks, 256, 0, s>>>(FcT, Pc_pad, Mpad, Mtot, rows, psih, fct_absmax, (__bf16*)A16,
                                                      (__bf16*)A16lo, K16);
    else if (precision == 2)
        k_fct_to_lowp<_Float16><<<blocks, 256, 0, s>>>(FcT, Pc_pad, Mpad, Mtot, rows, psih, fct_absmax, (_Float16*)A16,
                                                        (_Float16*)A16lo, K16);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_recompute_split(hipStream_t s, int precision, int n, const RecomputeArgs& a) {
    if (n <= 0) return hipSuccess;
    if (a.T > kMaxT || a.K16 % 32) return hipErrorInvalidValue;  // bidx rows hold kMaxT translations; KH % 16 == 0
    if (precision == 1) k_recompute_split<__bf16><<<n, 256, 0, s>>>(a);
    else if (precision == 2) k_recompute_split<_Float16><<<n, 256, 0, s>>>(a);
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
