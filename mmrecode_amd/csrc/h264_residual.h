// H.264 residual reconstruction shared by the inter (h264.hip) and intra (h264_intra.hip) kernels:
// dequantisation with scaling lists, 4x4 / 8x8 inverse transforms, chroma DC Hadamard and the
// Intra16x16 luma DC Hadamard, run by one wave per macroblock.
//
// Reference (crates/codecs/h264/src/decoder.rs): transform_residual_4x4 / _8x8 /
// inverse_transform_8 / level_scale_* :14185-14442, transform_chroma_dc / chroma_qp :14102-14140,
// transform_intra16_luma_dc :14142-14183, add_luma_residual_blocks / _8x8_blocks /
// add_chroma_residual / add_intra16_luma_residual :12580-12789.
#pragma once

#include "device_common.h"
#include "recon_internal.h"

namespace mmr {

__constant__ static const uint8_t kZz4[16] = {0, 1, 4, 8, 5, 2, 3, 6, 9, 12, 13, 10, 7, 11, 14, 15};  // row*4+col
__constant__ static const uint8_t kZz8[64] = {
    0, 1, 8, 16, 9, 2, 3, 10, 17, 24, 32, 25, 18, 11, 4, 5, 12, 19, 26, 33, 40, 48, 41, 34, 27, 20,
    13, 6, 7, 14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51, 58, 59,
    52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};
__constant__ static const int32_t kNorm4[6][3] = {{10, 16, 13}, {11, 18, 14}, {13, 20, 16}, {14, 23, 18}, {16, 25, 20}, {18, 29, 23}};
__constant__ static const int32_t kNorm8[6][6] = {{20, 18, 32, 19, 25, 24}, {22, 19, 35, 21, 28, 26}, {26, 23, 42, 24, 33, 31},
                                                 {28, 25, 45, 26, 35, 33}, {32, 28, 51, 30, 40, 38}, {36, 32, 58, 34, 46, 43}};
__constant__ static const int32_t kChromaQp[22] = {29, 30, 31, 32, 32, 33, 34, 34, 35, 35, 36, 36, 37, 37, 37, 38, 38, 38, 39, 39, 39, 39};

MMR_DEV int chroma_qp(int luma_qp, int offset) {
    const int index = clampi(luma_qp + offset, 0, 51);
    return index < 30 ? index : kChromaQp[index - 30];
}

MMR_DEV int32_t level_scale_4x4(int qp, int row, int col, const MMR_GLOBAL uint8_t* scaling) {
    const int cat = ((row & 1) == 0 && (col & 1) == 0) ? 0 : (((row & 1) && (col & 1)) ? 1 : 2);
    return kNorm4[qp % 6][cat] * (int32_t)scaling[row * 4 + col];
}

MMR_DEV int32_t level_scale_8x8(int qp, int row, int col, const MMR_GLOBAL uint8_t* scaling) {
    const int rm = row & 3, cm = col & 3;
    const bool re = (rm & 1) == 0, ce = (cm & 1) == 0;
    int cat;
    if (!re && !ce) cat = 1;
    else if (re != ce) cat = (rm == 2 || cm == 2) ? 5 : 3;
    else cat = (rm == 0 && cm == 0) ? 0 : ((rm == 2 && cm == 2) ? 2 : 4);
    return kNorm8[qp % 6][cat] * (int32_t)scaling[row * 8 + col];
}

MMR_DEV int clip255(int v) { return clampi(v, 0, 255); }

// transform_residual_4x4 (decoder.rs:14185-14233) on one lane; out in raster order. T is int64_t
// (the reference's arithmetic) or int32_t when the macroblock's levels are small enough that no
// intermediate can leave the int32 range (then both give the same integers).
template <typename T>
MMR_DEV void residual_4x4(const int32_t* lv, int qp, bool dc_scaled, const MMR_GLOBAL uint8_t* scaling, int32_t out[16]) {
    T s[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int rc = kZz4[i];
        const T v = lv[i];
        if (i == 0 && dc_scaled) {
            s[rc] = v;
        } else {
            const T scale = level_scale_4x4(qp, rc >> 2, rc & 3, scaling);
            s[rc] = qp >= 24 ? (v * scale) << (qp / 6 - 4) : (v * scale + ((T)1 << (3 - qp / 6))) >> (4 - qp / 6);
        }
    }
    T h[16];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int b = r * 4;
        const T e0 = s[b] + s[b + 2], e1 = s[b] - s[b + 2];
        const T e2 = (s[b + 1] >> 1) - s[b + 3], e3 = s[b + 1] + (s[b + 3] >> 1);
        h[b] = e0 + e3;
        h[b + 1] = e1 + e2;
        h[b + 2] = e1 - e2;
        h[b + 3] = e0 - e3;
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const T g0 = h[c] + h[8 + c], g1 = h[c] - h[8 + c];
        const T g2 = (h[4 + c] >> 1) - h[12 + c], g3 = h[4 + c] + (h[12 + c] >> 1);
        out[0 * 4 + c] = (int32_t)((g0 + g3 + 32) >> 6);
        out[1 * 4 + c] = (int32_t)((g1 + g2 + 32) >> 6);
        out[2 * 4 + c] = (int32_t)((g1 - g2 + 32) >> 6);
        out[3 * 4 + c] = (int32_t)((g0 - g3 + 32) >> 6);
    }
}

// inverse_transform_8 (decoder.rs:14280-14308).
template <typename T>
MMR_DEV void itrans8(const T in[8], T out[8]) {
    const T a0 = in[0] + in[4], a2 = in[0] - in[4];
    const T a4 = (in[2] >> 1) - in[6], a6 = in[2] + (in[6] >> 1);
    const T b0 = a0 + a6, b2 = a2 + a4, b4 = a2 - a4, b6 = a0 - a6;
    const T a1 = -in[3] + in[5] - in[7] - (in[7] >> 1);
    const T a3 = in[1] + in[7] - in[3] - (in[3] >> 1);
    const T a5 = -in[1] + in[7] + in[5] + (in[5] >> 1);
    const T a7 = in[3] + in[5] + in[1] + (in[1] >> 1);
    const T b1 = a1 + (a7 >> 2), b7 = a7 - (a1 >> 2), b3 = a3 + (a5 >> 2), b5 = (a3 >> 2) - a5;
    out[0] = b0 + b7;
    out[1] = b2 + b5;
    out[2] = b4 + b3;
    out[3] = b6 + b1;
    out[4] = b6 - b1;
    out[5] = b4 - b3;
    out[6] = b2 - b5;
    out[7] = b0 - b7;
}

struct MbShared {
    uint8_t pred_y[256];
    uint8_t pred_c[2][64];
    int32_t res_y[256];
    int32_t res_c[2][64];
    int64_t s8[4][64];
    int32_t dc[2][4];
    int32_t dc16[16];   // Intra16x16 scaled luma DC, [block_y * 4 + block_x]
};

// Scaling lists of a macroblock's residual (indices into the picture's [6][16] 4x4 and [2][64]
// 8x8 lists): luma 4x4, luma 8x8, first chroma list; inter 3 / 1 / 4, intra 0 / 0 / 1. i16:
// Intra16x16, whose luma DC comes from its own Hadamard transform (levels [16 k]).
struct ResidualLists {
    int luma4, luma8, chroma;
    bool i16;
};
constexpr ResidualLists kInterLists = {3, 1, 4, false};

// Transform bypass of one macroblock (qpprime_y_zero_transform_bypass_flag): luma for the 4x4
// blocks of an Intra4x4 or inter macroblock with QP 0, chroma when the chroma QP is 0 (the
// reference decides chroma on the chroma QP, decoder.rs:12739). luma_modes packs the 16 DPCM
// modes by block index (0 vertical, 1 horizontal, else none: the Intra4x4 modes, 2 for inter);
// chroma_mode likewise (intra chroma mode 2 -> 0, 1 -> 1, else 2).
struct Bypass {
    bool luma, chroma;
    uint32_t luma_modes[4];
    int chroma_mode;
};

MMR_DEV int bypass_luma_mode(const Bypass& by, int k) {
    const uint32_t lo = (k & 4) ? by.luma_modes[1] : by.luma_modes[0], hi = (k & 4) ? by.luma_modes[3] : by.luma_modes[2];
    return (int)((((k & 8) ? hi : lo) >> (8 * (k & 3))) & 0xffu);
}

// transform_bypass_residual_4x4 (decoder.rs:14339-14380): levels placed by the zig-zag scan, then
// vertical (mode 0) or horizontal (mode 1) running sums inside the block.
MMR_DEV void bypass_residual_4x4(const int32_t* lv, int mode, int32_t out[16]) {
#pragma unroll
    for (int i = 0; i < 16; ++i) out[kZz4[i]] = lv[i];
    if (mode == 0) {
#pragma unroll
        for (int r = 1; r < 4; ++r)
#pragma unroll
            for (int c = 0; c < 4; ++c) out[r * 4 + c] += out[(r - 1) * 4 + c];
    } else if (mode == 1) {
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int c = 1; c < 4; ++c) out[r * 4 + c] += out[r * 4 + c - 1];
    }
}

// Residual of one macroblock (phases 2-4 of the kernel) into S.res_y / S.res_c, in integer type T,
// from its level record lv (global memory, or a copy in LDS).
template <typename T, typename LV>
MMR_DEV void residual_phases(MbShared& S, LV lv, int qp, int cqp, bool t8, int l, const MMR_GLOBAL uint8_t* scaling,
                             ResidualLists rl, uint32_t* status) {
    // ---- phase 1 (Intra16x16): transform_intra16_luma_dc, lane l = output (row l >> 2, column l & 3)
    if (rl.i16) {
        if (l < 16) {
            T c[16];
#pragma unroll
            for (int k = 0; k < 16; ++k) c[kZz4[k]] = lv[16 * k];
            const int r = l >> 2, col = l & 3;
            // f = H c H with H rows (1 1 1 1) (1 1 -1 -1) (1 -1 -1 1) (1 -1 1 -1)
            const int hs[4][4] = {{1, 1, 1, 1}, {1, 1, -1, -1}, {1, -1, -1, 1}, {1, -1, 1, -1}};
            T f = 0;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                T h = 0;
#pragma unroll
                for (int i = 0; i < 4; ++i) h += hs[r][i] * c[i * 4 + j];
                f += h * hs[j][col];
            }
            const T scale = level_scale_4x4(qp, 0, 0, scaling + rl.luma4 * 16);
            const T v = qp >= 36 ? (f * scale) << (qp / 6 - 6) : (f * scale + ((T)1 << (5 - qp / 6))) >> (6 - qp / 6);
            if (v < (T)INT32_MIN || v > (T)INT32_MAX) atomicOr(status, kDevStatusInvalidData);
            S.dc16[l] = (int32_t)v;
        }
        wave_lds_sync();
    }
    // ---- phase 2
    {
        if (!t8) {
            if (l < 16) {
                int32_t lev[16], res[16];
                const int kbx = ((l >> 2) & 1) * 2 + (l & 1), kby = (l >> 3) * 2 + ((l >> 1) & 1);
#pragma unroll
                for (int i = 0; i < 16; ++i) lev[i] = lv[l * 16 + i];
                if (rl.i16) lev[0] = S.dc16[kby * 4 + kbx];
                residual_4x4<T>(lev, qp, rl.i16, scaling + rl.luma4 * 16, res);
#pragma unroll
                for (int i = 0; i < 16; ++i) S.res_y[(kby * 4 + (i >> 2)) * 16 + kbx * 4 + (i & 3)] = res[i];
            }
        } else {
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int idx = l * 4 + t, blk8 = idx >> 6, i = idx & 63, rc = kZz8[i];
                const T v = (T)lv[idx] * (T)level_scale_8x8(qp, rc >> 3, rc & 7, scaling + 96 + rl.luma8 * 64);
                const int shift = qp / 6;
                T sv = shift >= 6 ? v << (shift - 6) : (v + ((T)1 << (5 - shift))) >> (6 - shift);
                if (rc == 0) sv += 32;
                S.s8[blk8][rc] = sv;
            }
        }
        const int dcl = t8 ? l : l - 16;
        if (dcl == 0 || dcl == 1) {
            // transform_chroma_dc (decoder.rs:14114-14140) with the chroma list of the component.
            const int comp = dcl;
            const T c0 = lv[256 + comp * 4], c1 = lv[257 + comp * 4], c2 = lv[258 + comp * 4], c3 = lv[259 + comp * 4];
            const T f[4] = {c0 + c1 + c2 + c3, c0 - c1 + c2 - c3, c0 + c1 - c2 - c3, c0 - c1 - c2 + c3};
            const T scale = level_scale_4x4(cqp, 0, 0, scaling + (rl.chroma + comp) * 16);
#pragma unroll
            for (int i = 0; i < 4; ++i) S.dc[comp][i] = (int32_t)(((f[i] * scale) << (cqp / 6)) >> 5);
        }
    }
    wave_lds_sync();

    // ---- phase 3
    {
        if (t8 && l < 32) {
            const int blk8 = l >> 3, r = l & 7;
            T in[8], out[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) in[i] = S.s8[blk8][r * 8 + i];
            itrans8(in, out);
#pragma unroll
            for (int i = 0; i < 8; ++i) S.s8[blk8][r * 8 + i] = out[i];
        }
        if (l >= 32 && l < 40) {
            const int comp = (l - 32) >> 2, b = (l - 32) & 3;
            int32_t co[16], res[16];
            co[0] = S.dc[comp][b];
#pragma unroll
            for (int i = 1; i < 16; ++i) co[i] = lv[264 + (comp * 4 + b) * 15 + (i - 1)];
            residual_4x4<T>(co, cqp, true, scaling + (rl.chroma + comp) * 16, res);
#pragma unroll
            for (int i = 0; i < 16; ++i) S.res_c[comp][((b >> 1) * 4 + (i >> 2)) * 8 + (b & 1) * 4 + (i & 3)] = res[i];
        }
    }
    wave_lds_sync();

    // ---- phase 4
    if (t8 && l < 32) {
        const int blk8 = l >> 3, c = l & 7;
        T in[8], out[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) in[i] = S.s8[blk8][i * 8 + c];
        itrans8(in, out);
#pragma unroll
        for (int i = 0; i < 8; ++i)
            S.res_y[((blk8 >> 1) * 8 + i) * 16 + (blk8 & 1) * 8 + c] = (int32_t)(out[i] >> 6);
    }
    wave_lds_sync();
}

// Transform bypass after residual_phases (rare: a wave-uniform branch of its own, so the common
// path keeps its registers): the bypassed residuals replace the transformed ones in S -- luma 4x4
// blocks on lanes 0-15 (lane = block index), chroma blocks on lanes 32-39 (DC levels as they are
// in coefficient 0), then continue_transform_bypass_chroma (decoder.rs:14382-14420): the lower /
// right blocks of a component add the upper block's last row / the left block's last column.
template <typename LV>
MMR_DEV void bypass_phases(MbShared& S, LV lv, int l, const Bypass& by) {
    if (by.luma && l < 16) {
        int32_t lev[16], res[16];
        const int kbx = ((l >> 2) & 1) * 2 + (l & 1), kby = (l >> 3) * 2 + ((l >> 1) & 1);
#pragma unroll
        for (int i = 0; i < 16; ++i) lev[i] = lv[l * 16 + i];
        bypass_residual_4x4(lev, bypass_luma_mode(by, l), res);
#pragma unroll
        for (int i = 0; i < 16; ++i) S.res_y[(kby * 4 + (i >> 2)) * 16 + kbx * 4 + (i & 3)] = res[i];
    }
    if (by.chroma && l >= 32 && l < 40) {
        const int comp = (l - 32) >> 2, b = (l - 32) & 3;
        int32_t co[16], res[16];
        co[0] = lv[256 + comp * 4 + b];
#pragma unroll
        for (int i = 1; i < 16; ++i) co[i] = lv[264 + (comp * 4 + b) * 15 + (i - 1)];
        bypass_residual_4x4(co, by.chroma_mode, res);
#pragma unroll
        for (int i = 0; i < 16; ++i) S.res_c[comp][((b >> 1) * 4 + (i >> 2)) * 8 + (b & 1) * 4 + (i & 3)] = res[i];
    }
    wave_lds_sync();
    if (by.chroma && by.chroma_mode != 2 && l < 16) {
        const int comp = l >> 3, j = l & 7;
        int32_t* rc = S.res_c[comp];
        if (by.chroma_mode == 0) {
            const int32_t carry = rc[3 * 8 + j];
            for (int r = 4; r < 8; ++r) rc[r * 8 + j] += carry;
        } else {
            const int32_t carry = rc[j * 8 + 3];
            for (int c = 4; c < 8; ++c) rc[j * 8 + c] += carry;
        }
    }
    wave_lds_sync();
}

}  // namespace mmr
