// Pieces shared by the coarse cross-term GEMMs: FP32 MFMA (kernels.hip) and BF16/FP16 MFMA
// (gemm_lowp.hip).  Both use 128x128 output tiles, 4 waves in a 2x2 grid, each wave 2x2 blocks of
// 32x32 MFMA accumulators (lane l, register r -> row 8*(r/4) + 4*(l/32) + r%4, col l%32).
#pragma once
#include <climits>

#include "device_util.h"
#include "kernels.h"

namespace rome {

typedef float f16v __attribute__((ext_vector_type(16)));
typedef float f4v __attribute__((ext_vector_type(4)));

// blockIdx -> (bm, bn).  Workgroups are dispatched round-robin over the 8 XCDs (XCD = bid % 8).
// When the m-tiles split evenly, XCD x owns the band [x*band, (x+1)*band) of m-tiles and walks it in
// sub-bands of `sband` tiles, n-tile after n-tile: the sub-band's rows of the rotation operand stay
// in that XCD's L2 while the image operand streams past once per sub-band.  When the band is one
// sub-band (one pass over the n-tiles), XCD x starts its pass x/8 of the way along the n-tiles, so
// that the 8 bands reach an image's columns at 8 different times: the later ones find the image's
// list bound set by the earlier ones and keep fewer entries (in the same order, all 8 bands of an
// n-tile run at once and each sees only its own minimum).  A bijection on [0, nbm*nbn) in both
// branches.
DEV void gemm_tile(int bid, int nbm, int nbn, int sband, int& bm, int& bn) {
    const int band = nbm >> 3;
    if ((nbm & 7) == 0 && sband > 0 && band % sband == 0) {
        const int x = bid & 7, i = bid >> 3;
        const int per = sband * nbn;
        const int sb = i / per, rem = i - sb * per;
        bm = x * band + sb * sband + rem % sband;
        bn = rem / sband;
        if (sband == band) bn = (bn + x * ((nbn + 7) >> 3)) % nbn;
    } else {
        bm = bid % nbm;
        bn = bid / nbm;
    }
}

// gemm_tile, except that without row bands (sband == 0) the tiles are grouped by XCD: blocks reach the
// 8 XCDs round-robin (XCD = bid % 8) and each XCD takes whole column tiles, the nbm row tiles of one
// column tile in a row, so that each column tile of the image operand is read into one L2 only.  The
// grid holds gemm_grouped_blocks(nbm, nbn, sband) blocks; false: a padding block, which has no tile.
DEV bool gemm_tile_grouped(int bid, int nbm, int nbn, int sband, int& bm, int& bn) {
    if (sband > 0) {
        gemm_tile(bid, nbm, nbn, sband, bm, bn);
        return true;
    }
    const int x = bid & 7, i = bid >> 3;
    bm = i % nbm;
    bn = (i / nbm) * 8 + x;
    return bn < nbn;
}
inline unsigned gemm_grouped_blocks(int nbm, int nbn, int sband) {
    return (unsigned)nbm * (unsigned)(sband > 0 ? nbn : (nbn + 7) / 8 * 8);
}

// Images one tile of TN columns can touch (T >= 1): TN + 2.
template <int TN>
constexpr int epi_slots() { return TN + 2; }

// One E value from an accumulator: At - C (FP32 arm) or At - C * inv_scale as one fused
// multiply-add (scaled arms), the same rounding in every epilogue.
template <bool SCALED>
DEV float e_value(float at, float c, float s) {
    return SCALED ? __builtin_fmaf(-c, s, at) : at - c;
}

// E[img][t][m] = At[img][m] - s_img * C[m][img*T + t] (s_img = inv_scale[img] when SCALED, else 1)
// for a tile whose first column is n0 (TN columns); this wave holds NI x NJ accumulator blocks
// starting at row r0 and column c0.  Also the per-image minimum of E over rotations with a positive
// prior into emin_key (ordered-int keys): LDS atomics per image slot, then one global atomic per
// image the tile touches.  At rows and the valid flags are padded to Mpad (16-byte aligned rows).
// Rotation-contiguous E rows make each 32-row accumulator block one whole 128-byte line per column,
// written by one wave with 16-byte stores (an unaligned line would be read back from HBM to merge).
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
        float* Er = E + ((size_t)img * T + t) * Mtot;
        const bool vec = (Mtot & 3) == 0;
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
                float ev[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    ev[u] = e_value<SCALED>(atv[u], acc[i][j][4 * q + u], s);
                    if (base + u < Mtot && ((v4 >> (8 * u)) & 0xff)) emin = fminf(emin, ev[u]);
                }
                if (vec && base + 3 < Mtot) {
                    *reinterpret_cast<float4*>(Er + base) = make_float4(ev[0], ev[1], ev[2], ev[3]);
                } else {
#pragma unroll
                    for (int u = 0; u < 4; ++u)
                        if (base + u < Mtot) Er[base + u] = ev[u];
                }
            }
        if (emin < INFINITY) atomicMin(&slot_min[img - img0], float_order_key(emin));
    }
    __syncthreads();
    for (int q = tid; q < nslot; q += blockDim.x)
        if (slot_min[q] != INT_MAX) atomicMin(&emin_key[img0 + q], slot_min[q]);
}

// Same results as gemm_epilogue, with the E rows staged through LDS: for each 32-row accumulator block
// the wave writes its values column-major into `stage` (NJ*32 columns x 36 floats, this wave's own
// region), then eight lanes store one column's 32 rows, one whole 128-byte line per store instruction,
// as non-temporal (streaming) stores: complete lines that neither merge in L2 nor evict the operand
// panels there.  All threads of the block call it (barriers).
template <bool SCALED, int NI, int NJ, int TN>
DEV void gemm_epilogue_staged(const f16v (&acc)[NI][NJ], int r0, int c0, int n0, int lane, int tid,
                              const float* __restrict__ At, const float* __restrict__ inv_scale,
                              float* __restrict__ E, const unsigned char* __restrict__ valid,
                              int* __restrict__ emin_key, int Mtot, int Mpad, int Ncols, int T, int* slot_min,
                              float* stage) {
    constexpr int LD = 36;  // column stride in floats (16-byte aligned)
    const int img0 = n0 / T;
    const int last = min(n0 + TN, Ncols) - 1;
    const int nslot = last >= n0 ? last / T - img0 + 1 : 0;
    for (int q = tid; q < nslot; q += blockDim.x) slot_min[q] = INT_MAX;
    const bool vec = (Mtot & 3) == 0;
    const int h = lane >> 5;
    int img_j[NJ];
    float s_j[NJ], emin_j[NJ];
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
        const int col = min(c0 + j * 32 + (lane & 31), Ncols - 1);  // clamped: outputs of padding columns are dropped
        img_j[j] = col / T;
        s_j[j] = SCALED ? inv_scale[img_j[j]] : 1.f;
        emin_j[j] = INFINITY;
    }
#pragma unroll
    for (int i = 0; i < NI; ++i) {
        // registers -> LDS (column-major), running minimum of this lane's column
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
            const float* Ar = At + (size_t)img_j[j] * Mpad;
            const bool colok = c0 + j * 32 + (lane & 31) < Ncols;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int base = r0 + i * 32 + 8 * q + 4 * h;
                const float4 at = *reinterpret_cast<const float4*>(Ar + base);
                const uint32_t v4 = *reinterpret_cast<const uint32_t*>(valid + base);
                const float atv[4] = {at.x, at.y, at.z, at.w};
                float ev[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    ev[u] = e_value<SCALED>(atv[u], acc[i][j][4 * q + u], s_j[j]);
                    if (colok && base + u < Mtot && ((v4 >> (8 * u)) & 0xff)) emin_j[j] = fminf(emin_j[j], ev[u]);
                }
                *reinterpret_cast<float4*>(&stage[(j * 32 + (lane & 31)) * LD + 8 * q + 4 * h]) =
                    make_float4(ev[0], ev[1], ev[2], ev[3]);
            }
        }
        __syncthreads();
        // LDS -> E: lane group of 8 = one column's 32 rows = one 128-byte line
        const int row0 = r0 + i * 32 + 4 * (lane & 7);
#pragma unroll
        for (int p = 0; p < NJ * 4; ++p) {
            const int c = p * 8 + (lane >> 3);
            const int col = c0 + c;
            if (col < Ncols) {
                const int img = col / T, t = col - img * T;
                const float4 v = *reinterpret_cast<const float4*>(&stage[c * LD + 4 * (lane & 7)]);
                float* Er = E + ((size_t)img * T + t) * Mtot + row0;
                if (vec && row0 + 3 < Mtot) {
                    const f4v w = {v.x, v.y, v.z, v.w};
                    __builtin_nontemporal_store(w, reinterpret_cast<f4v*>(Er));
                } else {
                    const float vv[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                    for (int u = 0; u < 4; ++u)
                        if (row0 + u < Mtot) Er[u] = vv[u];
                }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < NJ; ++j)
        if (c0 + j * 32 + (lane & 31) < Ncols && emin_j[j] < INFINITY)
            atomicMin(&slot_min[img_j[j] - img0], float_order_key(emin_j[j]));
    __syncthreads();
    for (int q = tid; q < nslot; q += blockDim.x)
        if (slot_min[q] != INT_MAX) atomicMin(&emin_key[img0 + q], slot_min[q]);
}

// List epilogue (ListArgs): the per-image minimum of E as gemm_epilogue, and instead of E the
// entries whose v = E - ln pdf_direction can lie within reach[img] of the image's minimum v.  The
// E values never leave the registers:
//   1. E in place of the accumulators; per lane and column the minimum of E over rows with a
//      prior (finite ln pdf_direction) and the minimum vf of v in FP32;
//   2. per image of the tile (LDS slots, then one global atomic each) the bound B >= min v, where a
//      lane offers vf_min + 1e-6 |vf_min| + 1e-4 >= its entry's exact v = (double)E - ln pdf (FP32
//      rounding of E - l: < 6e-8 |vf|; of l = (float)ln pdf: < 6e-8 * 746);
//   3. the test vf < lim = B + reach + 1e-6 (|B| + reach) + 1e-3 keeps every entry with exact
//      v < B + reach (the margin exceeds both roundings and that of lim itself); counts per
//      accumulator block -> LDS offsets per image -> one global atomic per image to reserve the
//      range -> each lane stores its own entries.
// B only falls, so every entry the weights need passes, whatever the order in which tiles run.  All
// threads of the block call it (barriers).  slot_* arrays hold epi_slots<TN>() entries.
// Rotation rows of a wave's accumulator blocks: rows(i, q) = the rotation of the four rows
// 8q + 4(lane/32) .. +3 of block i (consecutive rotations), or -1 for padding rows.
struct ContigRows {  // block i = rotations r0 + 32 i ..
    static constexpr bool kPadding = false;  // never -1
    static constexpr bool kPaired = false;   // the accumulators hold the cross terms themselves
    int r0, h;
    DEV int operator()(int i, int q) const { return r0 + i * 32 + 8 * q + 4 * h; }
};

template <bool SCALED, int NI, int NJ, int TN, typename Rows>
DEV void gemm_epilogue_list(f16v (&acc)[NI][NJ], const Rows& rows, int c0, int n0, int lane, int tid,
                            const float* __restrict__ At, const float* __restrict__ inv_scale,
                            int* __restrict__ emin_key, int Mtot, int Mpad, int Ncols, int T,
                            const ListArgs& L, int* slot_min, int* slot_vb, float* slot_lim, int* slot_cnt) {
    const int img0 = n0 / T;
    const int last = min(n0 + TN, Ncols) - 1;
    const int nslot = last >= n0 ? last / T - img0 + 1 : 0;
    for (int q = tid; q < nslot; q += blockDim.x) {
        slot_min[q] = INT_MAX;
        slot_vb[q] = INT_MAX;
        slot_cnt[q] = 0;
    }
    __syncthreads();
    const int h = lane >> 5;
    int sl_j[NJ];
    bool ok_j[NJ];
    float s_j[NJ], emin_j[NJ], vmin_ij[NI][NJ];  // vmin per accumulator block: the count skips whole blocks
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
        const int col = c0 + j * 32 + (lane & 31);
        ok_j[j] = col < Ncols;
        const int img = (int)((unsigned)min(col, Ncols - 1) / (unsigned)T);  // clamped: padding columns are dropped
        sl_j[j] = img - img0;
        s_j[j] = SCALED ? inv_scale[img] : 1.f;
        emin_j[j] = INFINITY;
#pragma unroll
        for (int i = 0; i < NI; ++i) vmin_ij[i][j] = INFINITY;
    }
    // 1. E in place, minima
    if constexpr (Rows::kPaired) {
        // paired cross term: block 0 holds P, block 1 holds Q; rotation rows(0, q) has cross term P + Q,
        // rows(1, q) has P - Q (both formed here, in place)
        static_assert(NI == 2, "paired blocks");
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            // the tile's rows of At and ln pdf_direction were staged in LDS (rows.at / rows.lp)
            const int c0l = rows.local(0, q), c1l = rows.local(1, q);
            const float4 l0 = *reinterpret_cast<const float4*>(rows.lp + c0l);
            const float4 l1 = *reinterpret_cast<const float4*>(rows.lp + c1l);
            const float lp0[4] = {l0.x, l0.y, l0.z, l0.w}, lp1[4] = {l1.x, l1.y, l1.z, l1.w};
#pragma unroll
            for (int j = 0; j < NJ; ++j) {
                const float* Ar = rows.at + sl_j[j] * rows.kTileRows;
                const float4 a0 = *reinterpret_cast<const float4*>(Ar + c0l);
                const float4 a1 = *reinterpret_cast<const float4*>(Ar + c1l);
                const float at0[4] = {a0.x, a0.y, a0.z, a0.w}, at1[4] = {a1.x, a1.y, a1.z, a1.w};
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const float P = acc[0][j][4 * q + u], Q = acc[1][j][4 * q + u];
                    const float e0 = e_value<SCALED>(at0[u], P + Q, s_j[j]);
                    const float e1 = e_value<SCALED>(at1[u], P - Q, s_j[j]);
                    acc[0][j][4 * q + u] = e0;
                    acc[1][j][4 * q + u] = e1;
                    // the staged At is +inf on rows without a prior, so is their E: out of the minimum
                    emin_j[j] = fminf(fminf(emin_j[j], e0), e1);  // v_min3
                    vmin_ij[0][j] = fminf(vmin_ij[0][j], e0 - lp0[u]);
                    vmin_ij[1][j] = fminf(vmin_ij[1][j], e1 - lp1[u]);
                }
            }
        }
    } else {
#pragma unroll
    for (int i = 0; i < NI; ++i)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int base = rows(i, q);  // padding rows (-1): E = +inf, no prior, never kept
            float4 lp = make_float4(-INFINITY, -INFINITY, -INFINITY, -INFINITY);
            if (!Rows::kPadding || base >= 0) lp = *reinterpret_cast<const float4*>(L.lpd + base);
            const float lpd[4] = {lp.x, lp.y, lp.z, lp.w};
            float pen[4];  // +inf for rows without a prior (no part in the minimum of E), else 0
#pragma unroll
            for (int u = 0; u < 4; ++u) pen[u] = lpd[u] > -INFINITY ? 0.f : INFINITY;
#pragma unroll
            for (int j = 0; j < NJ; ++j) {
                float4 at = make_float4(INFINITY, INFINITY, INFINITY, INFINITY);
                if (!Rows::kPadding || base >= 0)
                    at = *reinterpret_cast<const float4*>(At + (size_t)(img0 + sl_j[j]) * Mpad + base);
                const float atv[4] = {at.x, at.y, at.z, at.w};
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const float e = (!Rows::kPadding || base >= 0) ? e_value<SCALED>(atv[u], acc[i][j][4 * q + u], s_j[j])
                                                                    : INFINITY;
                    acc[i][j][4 * q + u] = e;
                    emin_j[j] = fminf(emin_j[j], e + pen[u]);  // rows with a prior
                    vmin_ij[i][j] = fminf(vmin_ij[i][j], e - lpd[u]);  // +inf without a prior
                }
            }
        }
    }
    // 2. bounds per image
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
        // branch-free: padding columns and infinite minima contribute INT_MAX (no change)
        float vmin = vmin_ij[0][j];
#pragma unroll
        for (int i = 1; i < NI; ++i) vmin = fminf(vmin, vmin_ij[i][j]);
        const int kmin = ok_j[j] && emin_j[j] < INFINITY ? float_order_key(emin_j[j]) : INT_MAX;
        const int kvb = ok_j[j] && vmin < INFINITY ? float_order_key(vmin + 1e-6f * fabsf(vmin) + 1e-4f) : INT_MAX;
        atomicMin(&slot_min[sl_j[j]], kmin);
        atomicMin(&slot_vb[sl_j[j]], kvb);
    }
    __syncthreads();
    for (int q = tid; q < nslot; q += blockDim.x) {
        const int img = img0 + q;
        if (slot_min[q] != INT_MAX) atomicMin(&emin_key[img], slot_min[q]);
        const int mine = slot_vb[q];
        const int glob = __hip_atomic_load(&L.vbound[img], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (mine < glob) atomicMin(&L.vbound[img], mine);
        const int b = min(mine, glob);  // a stale global value is larger: still a valid bound
        const float bv = float_from_order_key(b), r = L.reach[img];
        slot_lim[q] = b == INT_MAX ? -INFINITY : bv + r + 1e-6f * (fabsf(bv) + r) + 1e-3f;
    }
    __syncthreads();
    // 3. count, reserve, store
    float lim_j[NJ];
#pragma unroll
    for (int j = 0; j < NJ; ++j) lim_j[j] = ok_j[j] ? slot_lim[sl_j[j]] : -INFINITY;
    int off[NI][NJ];
#pragma unroll
    for (int i = 0; i < NI; ++i) {
        float lpd[16];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            float4 lp = make_float4(-INFINITY, -INFINITY, -INFINITY, -INFINITY);
            if constexpr (Rows::kPaired) {
                lp = *reinterpret_cast<const float4*>(rows.lp + rows.local(i, q));
            } else {
                const int base = rows(i, q);
                if (!Rows::kPadding || base >= 0) lp = *reinterpret_cast<const float4*>(L.lpd + base);
            }
            lpd[4 * q] = lp.x, lpd[4 * q + 1] = lp.y, lpd[4 * q + 2] = lp.z, lpd[4 * q + 3] = lp.w;
        }
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
            int cnt = 0;
            // vmin_ij is the minimum of the same differences over the lane's rows of this block: none passes
            // unless it does
            if (vmin_ij[i][j] < lim_j[j]) {
#pragma unroll
                for (int r = 0; r < 16; ++r) cnt += acc[i][j][r] - lpd[r] < lim_j[j];
            }
            off[i][j] = cnt ? atomicAdd(&slot_cnt[sl_j[j]], cnt) : -1;
        }
    }
    __syncthreads();
    for (int q = tid; q < nslot; q += blockDim.x) {
        const int c = slot_cnt[q];
        slot_cnt[q] = c ? atomicAdd(&L.count[img0 + q], c) : 0;  // now the base of this block's range
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NI; ++i) {
        bool any = false;
#pragma unroll
        for (int j = 0; j < NJ; ++j) any |= off[i][j] >= 0;
        if (!any) continue;
        float lpd[16];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            float4 lp = make_float4(-INFINITY, -INFINITY, -INFINITY, -INFINITY);
            if constexpr (Rows::kPaired) {
                lp = *reinterpret_cast<const float4*>(rows.lp + rows.local(i, q));
            } else {
                const int base = rows(i, q);
                if (!Rows::kPadding || base >= 0) lp = *reinterpret_cast<const float4*>(L.lpd + base);
            }
            lpd[4 * q] = lp.x, lpd[4 * q + 1] = lp.y, lpd[4 * q + 2] = lp.z, lpd[4 * q + 3] = lp.w;
        }
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
            if (off[i][j] < 0) continue;
            const int img = img0 + sl_j[j];
            const unsigned tm = (unsigned)(c0 + j * 32 + (lane & 31) - img * T) * (unsigned)Mtot;
            int pos = slot_cnt[sl_j[j]] + off[i][j];
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                if (!(acc[i][j][r] - lpd[r] < lim_j[j])) continue;
                if (pos < L.cap) {
                    const size_t o = (size_t)img * L.cap + pos;
                    L.ent[o] = tm + (unsigned)(rows(i, r >> 2) + (r & 3));
                    L.val[o] = acc[i][j][r];
                }
                ++pos;
            }
        }
    }
}

}  // namespace rome
