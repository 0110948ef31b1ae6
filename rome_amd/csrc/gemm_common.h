// Pieces shared by the coarse cross-term GEMMs: FP32 MFMA (kernels.hip) and BF16/FP16 MFMA
// (gemm_lowp.hip).  Both use 128x128 output tiles, 4 waves in a 2x2 grid, each wave 2x2 blocks of
// 32x32 MFMA accumulators (lane l, register r -> row 8*(r/4) + 4*(l/32) + r%4, col l%32).
#pragma once
#include <climits>

#include "device_util.h"
#include "kernels.h"

namespace rome {

typedef float f16v __attribute__((ext_vector_type(16)));

// blockIdx -> (bm, bn).  Workgroups are dispatched round-robin over the 8 XCDs (XCD = bid % 8).
// When the m-tiles split evenly, XCD x owns the band [x*band, (x+1)*band) of m-tiles and walks it in
// sub-bands of `sband` tiles, n-tile after n-tile: the sub-band's rows of the rotation operand stay
// in that XCD's L2 while the image operand streams past once per sub-band.  A bijection on
// [0, nbm*nbn) in both branches.
DEV void gemm_tile(int bid, int nbm, int nbn, int sband, int& bm, int& bn) {
    const int band = nbm >> 3;
    if ((nbm & 7) == 0 && sband > 0 && band % sband == 0) {
        const int x = bid & 7, i = bid >> 3;
        const int per = sband * nbn;
        const int sb = i / per, rem = i - sb * per;
        bm = x * band + sb * sband + rem % sband;
        bn = rem / sband;
    } else {
        bm = bid % nbm;
        bn = bid / nbm;
    }
}

// Images one tile of TN columns can touch (T >= 1): TN + 2.
template <int TN>
constexpr int epi_slots() { return TN + 2; }

// E[img][m][t] = At[img][m] - s_img * C[m][img*T + t] (s_img = inv_scale[img] when SCALED, else 1)
// for a tile whose first column is n0 (TN columns); this wave holds NI x NJ accumulator blocks
// starting at row r0 and column c0.  Also the per-image minimum of E over rotations with a positive
// prior into emin_key (ordered-int keys): LDS atomics per image slot, then one global atomic per
// image the tile touches.  At rows and the valid flags are padded to Mpad (16-byte aligned rows).
// Called by all threads of the block (contains barriers).
template <bool SCALED, int NI, int NJ, int TN>
DEV void gemm_epilogue(const f16v (&acc)[NI][NJ], int r0, int c0, int n0, int lane, int tid,
                       const float* __restrict__ At, const float* __restrict__ inv_scale, float* __restrict__ E,
                       const unsigned char* __restrict__ valid, int* __restrict__ emin_key, int Mtot, int Mpad,
                       int Ncols, int T, int* slot_min) {
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
        const float s = SCALED ? inv_scale[img] : 1.f;
        const float* Ar = At + (size_t)img * Mpad;
        float* Er = E + (size_t)img * Mtot * T + t;
        float emin = INFINITY;
#pragma unroll
        for (int i = 0; i < NI; ++i)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                // registers 4q..4q+3 hold four consecutive rows: one 16-byte At load, one 4-byte valid load
                const int base = r0 + i * 32 + 8 * q + 4 * (lane >> 5);
                const float4 at = *reinterpret_cast<const float4*>(Ar + base);
                const uint32_t v4 = *reinterpret_cast<const uint32_t*>(valid + base);
                const float atv[4] = {at.x, at.y, at.z, at.w};
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int row = base + u;
                    if (row < Mtot) {
                        const float c = SCALED ? acc[i][j][4 * q + u] * s : acc[i][j][4 * q + u];
                        const float e = atv[u] - c;
                        Er[(size_t)row * T] = e;
                        if ((v4 >> (8 * u)) & 0xff) emin = fminf(emin, e);
                    }
                }
            }
        if (emin < INFINITY) atomicMin(&slot_min[img - img0], float_order_key(emin));
    }
    __syncthreads();
    for (int q = tid; q < nslot; q += blockDim.x)
        if (slot_min[q] != INT_MAX) atomicMin(&emin_key[img0 + q], slot_min[q]);
}

}  // namespace rome
