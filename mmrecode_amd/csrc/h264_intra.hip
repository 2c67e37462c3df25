// H.264 intra macroblocks of frame pictures: Intra4x4 / Intra8x8 / Intra16x16 / chroma prediction,
// I_PCM, and their residuals, after the inter macroblocks (h264.hip) and before deblocking.
//
// Reference (crates/codecs/h264/src/decoder.rs): predict_intra4_block and vertical-right /
// horizontal-down :13311-13532, predict_intra8_block / filter_intra8_neighbors :13030-13309,
// predict_block / dc_prediction / plane_prediction :14583-14698, predict_chroma_block /
// predict_chroma_dc :14499-14581, the macroblock order of prediction and residual in
// reconstruct_intra4_luma_block / reconstruct_intra8_luma_block / predict_intra16_macroblock /
// add_intra16_luma_residual / predict_chroma_macroblock :12224-12628 and place_pcm_macroblock
// :10982-11004; availability :10801-10814.
//
// Ordering. An intra macroblock reads the reconstructed (not yet deblocked) samples of its left,
// top-left, top and top-right neighbours. Inter macroblocks are final when this kernel starts, so
// only intra ones order among themselves: (x, y) may run once (x-1, y) and (x+1, y-1) are done, a
// wavefront as in deblocking. Pictures are cut into bands of kIntraRows macroblock rows; a
// workgroup takes band tickets from a device counter (in picture and band order, so the band it
// waits for is held by a running workgroup or done) and wave w walks row w of the band left to
// right. Before an intra macroblock x a row publishes progress x (every macroblock before x is
// final) and waits until the row above has published min(x + 2, width): inside the band through an
// LDS counter with workgroup-scope release / acquire; across bands the band's last row stores its
// samples coherently across the XCDs (agent_store) and publishes to global memory every `publish`
// macroblocks and at its end once they have completed (publish_agent), and the next band's first
// row polls with relaxed agent-scope loads and reads the row above with coherent loads
// (agent_load): no L2 write-back or invalidate. A wave keeps its own row's previous intra
// macroblock in LDS for the left context.
//
// Inside a macroblock the residuals of all blocks are computed first (in parallel, shared with
// the inter kernel); then Intra4x4 walks its 16 blocks and Intra8x8 its four in block order, each
// predicted from an LDS tile holding the macroblock and its context, plus its residual.
#include "device_common.h"
#include "h264_residual.h"
#include "recon_internal.h"

namespace mmr {

constexpr uint32_t kIntraSpinLimit = 1u << 25;
constexpr int kIntraRows = 8;   // rows per band = waves per workgroup
constexpr int kIntraRing = 16;  // bottom rows of the last macroblocks a wave hands to the row below

struct IntraShared {
    MbShared S;
    // Luma tile: row 0 = the row above (column 0 the top-left sample, 1..16 the top row, 17..24
    // top-right); rows 1..16: column 0 the left column, 1..16 the macroblock.
    uint8_t ty[17][28];
    // Chroma tiles: row 0 (column 0 top-left, 1..8 top); rows 1..8: column 0 left, 1..8 macroblock.
    uint8_t tc[2][9][12];
    // Neighbours of the current 4x4 / 8x8 block: index 0 = top-left, 1.. = top (tn) or left (ln);
    // Intra8x8 also the filtered lines (ftn, fln).
    uint8_t tn[20];
    uint8_t ln[12];
    uint8_t ftn[20];
    uint8_t fln[12];
    // Right column of the previous macroblock of this row when it was intra.
    uint8_t left_y[16];
    uint8_t left_c[2][8];
    // Bottom rows of this row's last kIntraRing intra macroblocks (slot x % kIntraRing), the top
    // context of the row below in the same band.
    uint8_t ring_y[kIntraRing][16];
    uint8_t ring_c[kIntraRing][2][8];
    // Level record (or I_PCM samples) of the current macroblock, copied from global memory.
    __align__(16) int16_t lvs[MMR_H264_LEVELS_PER_MB];
};

// modes[k] of an intra record for a lane-varying k, by selects on the four packed words (an
// indexed byte load would put the record in scratch memory).
MMR_DEV uint32_t intra_mode_of(const mmr_recon_h264_intra& in, int k) {
    uint32_t w[4];
    __builtin_memcpy(w, in.modes, sizeof w);
    const uint32_t lo = (k & 4) ? w[1] : w[0], hi = (k & 4) ? w[3] : w[2];
    return (((k & 8) ? hi : lo) >> (8 * (k & 3))) & 0xffu;
}

MMR_DEV uint32_t ip_avg2(uint32_t a, uint32_t b) { return (a + b + 1) >> 1; }
MMR_DEV uint32_t ip_f121(uint32_t a, uint32_t b, uint32_t c) { return (a + 2 * b + c + 2) >> 2; }
MMR_DEV uint32_t ip_avg13(uint32_t a, uint32_t b) { return (a + 3 * b + 2) >> 2; }

// Intra4x4 / Intra8x8 sample (x, y) of an n x n block (n = 4 or 8) from the neighbour lines:
// T(i) = top i, L(i) = left i, T(-1) = L(-1) = top-left -- the formulas of predict_intra4_block /
// predict_intra8_block with top_or_corner(i) = T(i), left_or_corner(i) = L(i).
template <class TF, class LF>
MMR_DEV uint32_t pred_nxn(int mode, int n, int x, int y, TF T, LF L) {
    switch (mode) {
    case 0: return T(x);
    case 1: return L(y);
    case 3:
        return (x == n - 1 && y == n - 1) ? ip_avg13(T(2 * n - 2), T(2 * n - 1)) : ip_f121(T(x + y), T(x + y + 1), T(x + y + 2));
    case 4:
        if (x > y) return ip_f121(T(x - y - 2), T(x - y - 1), T(x - y));
        if (x < y) return ip_f121(L(y - x - 2), L(y - x - 1), L(y - x));
        return ip_f121(T(0), T(-1), L(0));
    case 5: {
        const int z = 2 * x - y;
        if (z >= 0 && (z & 1) == 0) return ip_avg2(T(x - y / 2 - 1), T(x - y / 2));
        if (z >= 0) return ip_f121(T(x - y / 2 - 2), T(x - y / 2 - 1), T(x - y / 2));
        if (z == -1) return ip_f121(L(0), T(-1), T(0));
        return ip_f121(L(y - 2 * x - 1), L(y - 2 * x - 2), L(y - 2 * x - 3));
    }
    case 6: {
        const int z = 2 * y - x;
        if (z >= 0 && (z & 1) == 0) return ip_avg2(L(y - x / 2 - 1), L(y - x / 2));
        if (z >= 0) return ip_f121(L(y - x / 2 - 2), L(y - x / 2 - 1), L(y - x / 2));
        if (z == -1) return ip_f121(T(0), T(-1), L(0));
        return ip_f121(T(x - 2 * y - 1), T(x - 2 * y - 2), T(x - 2 * y - 3));
    }
    case 7: {
        const int s = x + y / 2;
        return (y & 1) == 0 ? ip_avg2(T(s), T(s + 1)) : ip_f121(T(s), T(s + 1), T(s + 2));
    }
    case 8: {
        const int z = x + 2 * y, last = 2 * n - 3;   // 5 for 4x4, 13 for 8x8
        if (z < last && (z & 1) == 0) return ip_avg2(L(y + x / 2), L(y + x / 2 + 1));
        if (z < last) return ip_f121(L(y + x / 2), L(y + x / 2 + 1), L(y + x / 2 + 2));
        if (z == last) return ip_avg13(L(n - 2), L(n - 1));
        return L(n - 1);
    }
    default: return 128;   // mode 2 (DC) is computed by the caller
    }
}

// DC of an n x n block from the lines above (sum of n) and left (sum of n): dc_prediction.
MMR_DEV uint32_t dc_nxn(int n, bool has_top, bool has_left, uint32_t st, uint32_t sl) {
    if (has_top && has_left) return (st + sl + (uint32_t)n) / (uint32_t)(2 * n);
    if (has_top) return (st + (uint32_t)n / 2) / (uint32_t)n;
    if (has_left) return (sl + (uint32_t)n / 2) / (uint32_t)n;
    return 128;
}

// Which neighbours a 4x4 / 8x8 mode reads: bit 0 top, bit 1 left, bit 2 top-left.
MMR_DEV uint32_t nxn_needs(int mode) {
    constexpr uint32_t k = 1u << (0 * 3) | 2u << (1 * 3) | 0u << (2 * 3) | 1u << (3 * 3) | 7u << (4 * 3) |
                           7u << (5 * 3) | 7u << (6 * 3) | 1u << (7 * 3) | 2u << (8 * 3);
    return (k >> (3 * mode)) & 7u;
}

// Plane prediction of an n x n block (16 luma / 8 chroma) at (x, y) from tile row 0 / column 0,
// with the macroblock at tile rows / columns 1..n (plane_prediction).
template <int STRIDE>
MMR_DEV uint32_t plane_sample(const uint8_t (*tile)[STRIDE], int n, int x, int y) {
    const int half = n / 2;
    int h = 0, v = 0;
    for (int i = 1; i <= half; ++i) {
        h += i * ((int)tile[0][half + i] - (int)tile[0][half - i]);     // top[half-1+i] - top[half-1-i] (i = half: corner)
        v += i * ((int)tile[half + i][0] - (int)tile[half - i][0]);
    }
    const int scale = n == 16 ? 5 : 17, shift = n == 16 ? 6 : 5;
    const int a = 16 * ((int)tile[0][n] + (int)tile[n][0]);
    const int b = (scale * h + (1 << (shift - 1))) >> shift;
    const int c = (scale * v + (1 << (shift - 1))) >> shift;
    return (uint32_t)clip255((a + b * (x - (half - 1)) + c * (y - (half - 1)) + 16) >> 5);
}

// Waits for the row above, bounded (a protocol error reports TIMEOUT instead of hanging). Inside the
// band: the LDS counter of the row above, then an acquire of LDS only -- everything a row hands to
// the row below inside the band is in LDS (its ring of bottom rows); the global samples it reads
// there were written by an earlier kernel (inter macroblocks). Across bands: relaxed
// agent-scope polls of the previous band's published progress; the samples it covers are then read
// with coherent loads (the compiler barrier keeps them after the poll).
MMR_DEV void intra_wait_lds(uint32_t* counter, uint32_t need, uint32_t* status) {
    uint32_t spins = 0;
    while (poll_workgroup(counter) < need) {
        if (++spins > kIntraSpinLimit) {
            atomicOr(status, kDevStatusTimeout);
            break;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}
MMR_DEV void intra_wait_band(uint32_t* counter, uint32_t need, uint32_t& seen, uint32_t* status) {
    if (seen >= need) return;
    uint32_t spins = 0;
    for (;;) {
        seen = poll_agent(counter);
        if (seen >= need) break;
        if (++spins > kIntraSpinLimit) {
            atomicOr(status, kDevStatusTimeout);
            seen = need;
            break;
        }
        __builtin_amdgcn_s_sleep(2);
    }
    asm volatile("" ::: "memory");
}

// Residuals of every block of an intra macroblock (luma 4x4 / 8x8 / Intra16x16, chroma) into
// sh.S, zero without a level record; they do not depend on the neighbours, so they are computed
// before the wait for the row above. Also checks the type against the transform flag.
MMR_DEV void intra_residual(IntraShared& sh, const H264Job& job, const mmr_recon_h264_macroblock& m, const mmr_recon_h264_intra& in,
                            int l, uint32_t* status) {
    MbShared& S = sh.S;
    const int type = in.type;
    const bool pcm = type == MMR_H264_INTRA_PCM, i16 = type == MMR_H264_INTRA_16X16, i8 = type == MMR_H264_INTRA_8X8;
    if (type > MMR_H264_INTRA_PCM || i8 != ((m.flags & MMR_H264_MB_TRANSFORM_8X8) != 0)) {
        if (l == 0) atomicOr(status, kDevStatusInvalidData);
    }
    const bool has_record = pcm || i16 || m.cbp != 0;
    const int16_t* lv = sh.lvs;   // staged by the caller when has_record
    // ---- residuals of every block (luma 4x4 / 8x8 / Intra16x16, chroma), before prediction
    if (has_record && !pcm) {
        uint32_t mx = 0;
#pragma unroll
        for (int i = 0; i < MMR_H264_LEVELS_PER_MB / 64; ++i) {
            const int v = lv[l + 64 * i];
            mx = max(mx, (uint32_t)(v < 0 ? -v : v));
        }
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) mx = max(mx, (uint32_t)__shfl_xor((int)mx, d, 64));
        const int qp = m.qp, cqp = chroma_qp(qp, job.chroma_qp_offset[0]);   // first offset for both
        const ResidualLists rl = {0, 0, 1, i16};
        if (mx < 250) residual_phases<int32_t>(S, lv, qp, cqp, i8, l, as_global(job.scaling), rl, status);
        else residual_phases<int64_t>(S, lv, qp, cqp, i8, l, as_global(job.scaling), rl, status);
        if (job.transform_bypass && (qp == 0 || cqp == 0)) {
            Bypass by;
            by.luma = qp == 0 && type == MMR_H264_INTRA_4X4;
            by.chroma = cqp == 0;
            __builtin_memcpy(by.luma_modes, in.modes, sizeof by.luma_modes);
            by.chroma_mode = in.chroma_mode == 1 ? 1 : (in.chroma_mode == 2 ? 0 : 2);
            bypass_phases(S, lv, l, by);
        }
    } else {
        for (int i = l; i < 256; i += 64) S.res_y[i] = 0;
        for (int i = l; i < 128; i += 64) S.res_c[i >> 6][i & 63] = 0;
    }
    wave_lds_sync();
}

// Prediction of one intra macroblock (x, y) of `job`, mode record `in`, plus the residual already
// in sh.S; stores the macroblock and keeps its right column for the next one. `above` is the
// LDS of the row above when it is in the same band (else null: the samples come from global
// memory); bits 0 / 1 / 2 of `above_intra` say that macroblock x - 1 / x / x + 1 of the row above
// is intra, i.e. its bottom row is in above's ring (inter macroblocks: global memory).
MMR_DEV void intra_macroblock(IntraShared& sh, const H264Job& job, int mbx, int mby, const mmr_recon_h264_macroblock& m,
                              const mmr_recon_h264_intra& in, bool left_mb, bool top_mb, bool left_in_lds,
                              const IntraShared* above, uint32_t above_intra, bool agent_out, int l, uint32_t* status) {
    MbShared& S = sh.S;
    const int W = (int)job.width_mbs * 16, CW = W / 2;
    const MMR_GLOBAL uint8_t* py = as_global(job.planes[0]);
    const MMR_GLOBAL uint8_t* pc[2] = {as_global(job.planes[1]), as_global(job.planes[2])};
    const int type = in.type;
    const bool pcm = type == MMR_H264_INTRA_PCM, i16 = type == MMR_H264_INTRA_16X16, i8 = type == MMR_H264_INTRA_8X8;
    const int16_t* lv = sh.lvs;   // I_PCM samples, staged by the caller

    // ---- context tile
    const int x0 = mbx * 16, y0 = mby * 16, cx0 = mbx * 8, cy0 = mby * 8;
    const uint32_t ring_mid = above && (above_intra & 2u) ? (uint32_t)(mbx % kIntraRing) : ~0u;
    // A band's first row reads the row above from the picture, written by the previous band (another
    // workgroup) with coherent stores: coherent loads (the row above in the band is in LDS, or an
    // inter macroblock an earlier kernel wrote).
    auto top_sample = [&](const MMR_GLOBAL uint8_t* p) -> uint8_t { return above ? *p : agent_load(p); };
    if (top_mb) {
        if (l < 16) {
            sh.ty[0][1 + l] = ring_mid != ~0u ? above->ring_y[ring_mid][l] : top_sample(py + (size_t)(y0 - 1) * W + x0 + l);
        } else if (l < 24 && x0 + l < W) {
            sh.ty[0][1 + l] = above && (above_intra & 4u) ? above->ring_y[(mbx + 1) % kIntraRing][l - 16]
                                                          : top_sample(py + (size_t)(y0 - 1) * W + x0 + l);
        } else if (l >= 48) {
            const int comp = (l - 48) >> 3, i = (l - 48) & 7;
            sh.tc[comp][0][1 + i] = ring_mid != ~0u ? above->ring_c[ring_mid][comp][i] : top_sample(pc[comp] + (size_t)(cy0 - 1) * CW + cx0 + i);
        }
    }
    if (left_mb) {
        if (l >= 32 && l < 48) {
            const int r = l - 32;
            sh.ty[1 + r][0] = left_in_lds ? sh.left_y[r] : py[(size_t)(y0 + r) * W + x0 - 1];
        } else if (l >= 16 && l < 32) {
            const int comp = (l - 16) >> 3, r = (l - 16) & 7;
            sh.tc[comp][1 + r][0] = left_in_lds ? sh.left_c[comp][r] : pc[comp][(size_t)(cy0 + r) * CW + cx0 - 1];
        }
    }
    if (top_mb && left_mb && l >= 24 && l < 27) {
        const bool ring = above && (above_intra & 1u);
        const int slot = (mbx - 1) % kIntraRing;
        if (l == 24) sh.ty[0][0] = ring ? above->ring_y[slot][15] : top_sample(py + (size_t)(y0 - 1) * W + x0 - 1);
        else sh.tc[l - 25][0][0] = ring ? above->ring_c[slot][l - 25][7] : top_sample(pc[l - 25] + (size_t)(cy0 - 1) * CW + cx0 - 1);
    }

    wave_lds_sync();   // context tile

    // ---- luma
    if (pcm) {
        const int r = l >> 2, c0 = (l & 3) * 4;
#pragma unroll
        for (int c = 0; c < 4; ++c) sh.ty[1 + r][1 + c0 + c] = (uint8_t)lv[r * 16 + c0 + c];
    } else if (i16) {
        const int mode = in.luma_mode;
        const bool ok = mode <= 3 && (mode != 0 || top_mb) && (mode != 1 || left_mb) && (mode != 3 || (top_mb && left_mb));
        if (!ok && l == 0) atomicOr(status, kDevStatusInvalidData);
        uint32_t dc = 128;
        if (ok && mode == 2) {
            uint32_t st = 0, sl = 0;
            for (int i = 0; i < 16; ++i) {
                st += top_mb ? sh.ty[0][1 + i] : 0u;
                sl += left_mb ? sh.ty[1 + i][0] : 0u;
            }
            dc = dc_nxn(16, top_mb, left_mb, st, sl);
        }
        const int r = l >> 2, c0 = (l & 3) * 4;
        uint8_t out[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            uint32_t p = 128;
            if (ok) {
                if (mode == 0) p = sh.ty[0][1 + c0 + c];
                else if (mode == 1) p = sh.ty[1 + r][0];
                else if (mode == 2) p = dc;
                else p = plane_sample<28>(sh.ty, 16, c0 + c, r);
            }
            out[c] = (uint8_t)clip255((int)p + S.res_y[r * 16 + c0 + c]);
        }
        wave_lds_sync();   // plane prediction read the context; the macroblock interior is written now
#pragma unroll
        for (int c = 0; c < 4; ++c) sh.ty[1 + r][1 + c0 + c] = out[c];
    } else if (!i8) {
        // Intra4x4 in 10 steps: the blocks with bx + 2 by = t (up to two, lanes 0-15 and 16-31).
        // Every sample a block reads (left, top, top-left, and the top-right ones the rules below
        // keep, which all precede it in block order) lies in a block of a smaller t, and the two
        // blocks of a step do not read each other, so this equals the reference's block order.
        for (int t = 0; t < 10; ++t) {
            const int by = (t < 3 ? 0 : (t - 2) >> 1) + (l >> 4), bx = t - 2 * by;
            const bool active = l < 32 && by <= min(3, t >> 1) && bx >= 0 && bx <= 3;
            if (active) {
                const int k = (by >> 1) * 8 + (bx >> 1) * 4 + (by & 1) * 2 + (bx & 1);   // luma_block_index
                const int ox = bx * 4, oy = by * 4;
                const bool ht = by > 0 || top_mb, hl = bx > 0 || left_mb;
                const int mode = (int)intra_mode_of(in, k);
                const uint32_t needs = nxn_needs(mode);
                const bool ok = mode <= 8 && (!(needs & 1) || ht) && (!(needs & 2) || hl) && (!(needs & 4) || (ht && hl));
                if (!ok && (l & 15) == 0) atomicOr(status, kDevStatusInvalidData);
                // top-right (top 4..7): not for blocks 3 and 11, not into the next macroblock below
                // the first block row, not past the picture's right edge; else copies of top 3
                const bool tr_ok = k != 3 && k != 11 && !(by > 0 && ox + 4 >= 16) && x0 + ox + 4 < W;
                auto T = [&](int i) -> uint32_t { return sh.ty[oy][1 + ox + (i < 4 || tr_ok ? i : 3)]; };
                auto L = [&](int i) -> uint32_t { return sh.ty[oy + 1 + i][ox]; };
                const int x = l & 3, y = (l >> 2) & 3;
                uint32_t p = 128;
                if (ok && mode == 2) {
                    uint32_t st = 0, sl = 0;
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        st += ht ? T(i) : 0u;
                        sl += hl ? L(i) : 0u;
                    }
                    p = dc_nxn(4, ht, hl, st, sl);
                } else if (ok) {
                    p = pred_nxn(mode, 4, x, y, T, L);
                }
                sh.ty[1 + oy + y][1 + ox + x] = (uint8_t)clip255((int)p + S.res_y[(oy + y) * 16 + ox + x]);
            }
            wave_lds_sync();
        }
    } else {
        // Intra8x8: 4 blocks in order, neighbours filtered first (filter_intra8_neighbors)
        for (int g = 0; g < 4; ++g) {
            const int ox = (g & 1) * 8, oy = (g >> 1) * 8;
            const bool ht = g >= 2 || top_mb, hl = (g & 1) != 0 || left_mb, hc = ht && hl;
            const int mode = (int)intra_mode_of(in, g);
            if (l < 17) {
                uint32_t v = 0;
                if (l == 0) {
                    v = hc ? sh.ty[oy][ox] : 0u;
                } else if (ht) {
                    const int xi = l - 1;   // top-right: not for the last block, not past the right edge
                    const bool tr_ok = xi < 8 || (g != 3 && x0 + ox + xi < W);
                    v = sh.ty[oy][1 + ox + (tr_ok ? xi : 7)];
                }
                sh.tn[l] = (uint8_t)v;
            } else if (l < 26) {
                const int i = l - 17;
                sh.ln[i] = (uint8_t)(i == 0 ? (hc ? sh.ty[oy][ox] : 0u) : (hl ? sh.ty[oy + i][ox] : 0u));
            }
            wave_lds_sync();
            if (l < 17) {
                uint32_t v = 0;
                if (l == 0) {
                    v = hc ? ip_f121(sh.ln[1], sh.tn[0], sh.tn[1]) : 0u;
                } else if (ht) {
                    const int i = l - 1;
                    if (i == 0) v = hc ? ip_f121(sh.tn[0], sh.tn[1], sh.tn[2]) : ip_avg13(sh.tn[2], sh.tn[1]);
                    else if (i < 15) v = ip_f121(sh.tn[i], sh.tn[i + 1], sh.tn[i + 2]);
                    else v = ip_avg13(sh.tn[15], sh.tn[16]);
                }
                sh.ftn[l] = (uint8_t)v;
            } else if (l < 26) {
                const int i = l - 17;
                uint32_t v = 0;
                if (i == 0) {
                    v = hc ? ip_f121(sh.ln[1], sh.tn[0], sh.tn[1]) : 0u;
                } else if (hl) {
                    const int j = i - 1;
                    if (j == 0) v = hc ? ip_f121(sh.ln[0], sh.ln[1], sh.ln[2]) : ip_avg13(sh.ln[2], sh.ln[1]);
                    else if (j < 7) v = ip_f121(sh.ln[j], sh.ln[j + 1], sh.ln[j + 2]);
                    else v = ip_avg13(sh.ln[7], sh.ln[8]);
                }
                sh.fln[i] = (uint8_t)v;
            }
            wave_lds_sync();
            const uint32_t needs = nxn_needs(mode);
            const bool ok = mode <= 8 && (!(needs & 1) || ht) && (!(needs & 2) || hl) && (!(needs & 4) || hc);
            if (!ok && l == 0) atomicOr(status, kDevStatusInvalidData);
            {
                const int x = l & 7, y = l >> 3;
                uint32_t p = 128;
                if (ok && mode == 2) {
                    uint32_t st = 0, sl = 0;
#pragma unroll
                    for (int i = 0; i < 8; ++i) {
                        st += sh.ftn[1 + i];
                        sl += sh.fln[1 + i];
                    }
                    p = dc_nxn(8, ht, hl, st, sl);
                } else if (ok) {
                    p = pred_nxn(mode, 8, x, y, [&](int i) -> uint32_t { return sh.ftn[i + 1]; },
                                 [&](int i) -> uint32_t { return sh.fln[i + 1]; });
                }
                sh.ty[1 + oy + y][1 + ox + x] = (uint8_t)clip255((int)p + S.res_y[(oy + y) * 16 + ox + x]);
            }
            wave_lds_sync();
        }
    }

    // ---- chroma
    if (pcm) {
        const int comp = l >> 5, r = (l >> 2) & 7, c0 = (l & 3) * 2;
#pragma unroll
        for (int c = 0; c < 2; ++c) sh.tc[comp][1 + r][1 + c0 + c] = (uint8_t)lv[256 + comp * 64 + r * 8 + c0 + c];
    } else {
        const int mode = in.chroma_mode;
        const bool ok = mode <= 3 && (mode != 1 || left_mb) && (mode != 2 || top_mb) && (mode != 3 || (top_mb && left_mb));
        if (!ok && l == 0) atomicOr(status, kDevStatusInvalidData);
        const int comp = l >> 5, r = (l >> 2) & 7, c0 = (l & 3) * 2;
        const uint8_t (*t)[12] = sh.tc[comp];
        uint8_t out[2];
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const int x = c0 + c;
            uint32_t p = 128;
            if (ok) {
                if (mode == 1) {
                    p = t[1 + r][0];
                } else if (mode == 2) {
                    p = t[0][1 + x];
                } else if (mode == 3) {
                    p = plane_sample<12>(t, 8, x, r);
                } else {
                    // predict_chroma_dc: quadrants (0,0) and (1,1) use both lines, (1,0) prefers the
                    // top, (0,1) the left
                    const int qx = x >> 2, qy = r >> 2;
                    const bool ut = top_mb && (!left_mb || qy == 0 || qx == 1);
                    const bool ul = left_mb && (!top_mb || qx == 0 || qy == 1);
                    uint32_t sum = 0, count = 0;
                    if (ut) {
                        for (int i = 0; i < 4; ++i) sum += t[0][1 + qx * 4 + i];
                        count += 4;
                    }
                    if (ul) {
                        for (int i = 0; i < 4; ++i) sum += t[1 + qy * 4 + i][0];
                        count += 4;
                    }
                    p = count ? (sum + count / 2) / count : 128u;
                }
            }
            out[c] = (uint8_t)clip255((int)p + S.res_c[comp][r * 8 + x]);
        }
        wave_lds_sync();
#pragma unroll
        for (int c = 0; c < 2; ++c) sh.tc[comp][1 + r][1 + c0 + c] = out[c];
    }
    wave_lds_sync();

    // ---- store, and keep the right column as the next macroblock's left context
    {
        const int r = l >> 2, c0 = (l & 3) * 4;
        uint32_t word = 0;
#pragma unroll
        for (int c = 0; c < 4; ++c) word |= (uint32_t)sh.ty[1 + r][1 + c0 + c] << (8 * c);
        // (a band's last row: coherent stores, read by the next band's first row)
        MMR_GLOBAL uint32_t* dy = reinterpret_cast<MMR_GLOBAL uint32_t*>(as_global(job.planes[0]) + (size_t)(y0 + r) * W + x0 + c0);
        const int comp = l >> 5, cr = (l >> 2) & 7, cc = (l & 3) * 2;
        const uint16_t half = (uint16_t)((uint32_t)sh.tc[comp][1 + cr][1 + cc] | (uint32_t)sh.tc[comp][1 + cr][2 + cc] << 8);
        MMR_GLOBAL uint16_t* dc = reinterpret_cast<MMR_GLOBAL uint16_t*>(as_global(job.planes[1 + comp]) + (size_t)(cy0 + cr) * CW + cx0 + cc);
        if (agent_out) {
            agent_store(dy, word);
            agent_store(dc, half);
        } else {
            *dy = word;
            *dc = half;
        }
    }
    if (l < 16) sh.left_y[l] = sh.ty[1 + l][16];
    else if (l < 32) sh.left_c[(l - 16) >> 3][(l - 16) & 7] = sh.tc[(l - 16) >> 3][1 + ((l - 16) & 7)][8];
    wave_lds_sync();
}

__global__ __launch_bounds__(kIntraRows * 64) void h264_intra_kernel(const H264Job* __restrict__ jobs, uint32_t job_count,
                                                                      uint32_t total_bands, uint32_t* __restrict__ counters,
                                                                      uint32_t publish, uint32_t* __restrict__ status) {
    __shared__ IntraShared sh_all[kIntraRows];
    __shared__ uint32_t prog[kIntraRows];
    __shared__ uint32_t s_ticket;
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    IntraShared& sh = sh_all[w];
    for (;;) {
        if (threadIdx.x == 0) s_ticket = fetch_add_relaxed(&counters[0], 1);
        if (threadIdx.x < kIntraRows) prog[threadIdx.x] = 0;
        __syncthreads();
        const uint32_t t = s_ticket;
        if (t >= total_bands) break;
        const H264Job& job = jobs[find_job(jobs, job_count, t, [](const H264Job& j) { return j.intra_band_begin; })];
        const uint32_t band = t - job.intra_band_begin;
        const int mby = (int)band * kIntraRows + w, mw = (int)job.width_mbs;
        if (mby < (int)job.height_mbs) {
            const bool first = w == 0, last = w == kIntraRows - 1 || mby == (int)job.height_mbs - 1;
            uint32_t* gprev = first && band > 0 ? &counters[1 + t - 1] : nullptr;   // previous band's last row
            uint32_t* gself = last && mby < (int)job.height_mbs - 1 ? &counters[1 + t] : nullptr;
            uint32_t seen = 0, published = 0, gpublished = 0;
            bool left_in_lds = false;
            auto publish_to = [&](uint32_t x, bool force) {
                if (x > published) {
                    published = x;
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");   // the ring, not the picture
                    __hip_atomic_store(&prog[w], x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                }
                if (gself && x > gpublished && (force || x - gpublished >= publish)) {
                    gpublished = x;
                    publish_agent(gself, x);
                }
            };
            // Records of 64 macroblocks at a time, lane i holding macroblock chunk + i: its record
            // (3 dwords), intra record (5 dwords) and the slice of the macroblock above; read out
            // with readlane.
            uint32_t rec[3] = {0, 0, 0}, irec[5] = {0, 0, 0, 0, 0}, above_slice = 0, above3 = 0;
            const IntraShared* above = first ? nullptr : &sh_all[w - 1];
            uint32_t left_slice = 0;
            // Level record of the next intra macroblock of the chunk, loaded while the current one
            // is reconstructed: lane l holds dwords 3l .. 3l+2 of the 768-byte record.
            uint32_t pf[3] = {0, 0, 0};
            int pf_mb = -1;
            const MMR_GLOBAL uint32_t* levels_dw = reinterpret_cast<const MMR_GLOBAL uint32_t*>(as_global(job.levels));
            auto record_of = [&]() -> bool {   // this lane's macroblock is intra with a level record
                const uint32_t type = irec[0] & 0xffu;
                return (rec[0] & MMR_H264_MB_INTRA) &&
                       (type == MMR_H264_INTRA_PCM || type == MMR_H264_INTRA_16X16 || ((rec[0] >> 16) & 0xffu) != 0);
            };
            for (int mbx = 0; mbx < mw; ++mbx) {
                const uint32_t mb = (uint32_t)(mby * mw + mbx);
                if ((mbx & 63) == 0) {
                    const int xl = mbx + l;
                    if (xl < mw) {
                        const MMR_GLOBAL uint32_t* q = reinterpret_cast<const MMR_GLOBAL uint32_t*>(as_global(job.mbs + mb + l));
                        rec[0] = q[0];
                        rec[1] = q[1];
                        rec[2] = q[2];
                        if (rec[0] & MMR_H264_MB_INTRA) {
                            const MMR_GLOBAL uint32_t* r = reinterpret_cast<const MMR_GLOBAL uint32_t*>(as_global(job.intra + mb + l));
#pragma unroll
                            for (int i = 0; i < 5; ++i) irec[i] = r[i];
                        }
                        above_slice = mby > 0 ? (uint32_t)job.mbs[mb + l - (uint32_t)mw].slice_id : 0u;
                        if (above) {   // intra flags of macroblocks x - 1, x, x + 1 of the row above
                            const mmr_recon_h264_macroblock* up = job.mbs + mb + l - (uint32_t)mw;
                            above3 = (xl > 0 && (up[-1].flags & MMR_H264_MB_INTRA) ? 1u : 0u) |
                                     ((up[0].flags & MMR_H264_MB_INTRA) ? 2u : 0u) |
                                     (xl + 1 < mw && (up[1].flags & MMR_H264_MB_INTRA) ? 4u : 0u);
                        }
                    }
                }
                const int lane = mbx & 63;
                uint32_t wd[3];
#pragma unroll
                for (int i = 0; i < 3; ++i) wd[i] = __builtin_amdgcn_readlane(rec[i], lane);
                mmr_recon_h264_macroblock m;
                __builtin_memcpy(&m, wd, sizeof m);
                const uint32_t slice = m.slice_id, prev_slice = left_slice;
                left_slice = slice;
                if (!(m.flags & MMR_H264_MB_INTRA)) {
                    left_in_lds = false;
                    continue;
                }
                uint32_t iw[5];
#pragma unroll
                for (int i = 0; i < 5; ++i) iw[i] = __builtin_amdgcn_readlane(irec[i], lane);
                mmr_recon_h264_intra in;
                static_assert(sizeof(mmr_recon_h264_intra) == 20, "intra record is 5 dwords");
                __builtin_memcpy(&in, iw, sizeof in);
                publish_to((uint32_t)mbx, false);
                // Stage this macroblock's level record in LDS (prefetched, or loaded now), then start
                // loading the record of the next intra macroblock of the chunk.
                const bool has_record = in.type == MMR_H264_INTRA_PCM || in.type == MMR_H264_INTRA_16X16 || m.cbp != 0;
                if (has_record) {
                    if (pf_mb != mbx) {
                        const MMR_GLOBAL uint32_t* q = levels_dw + (size_t)m.coef_offset * (MMR_H264_LEVELS_PER_MB / 2) + 3 * l;
                        pf[0] = q[0];
                        pf[1] = q[1];
                        pf[2] = q[2];
                    }
                    uint32_t* dst = reinterpret_cast<uint32_t*>(sh.lvs) + 3 * l;
                    dst[0] = pf[0];
                    dst[1] = pf[1];
                    dst[2] = pf[2];
                }
                {
                    const uint64_t later = __ballot(record_of() && mbx - lane + l < mw) & ~((2ull << lane) - 1ull);
                    pf_mb = -1;
                    if (later) {
                        const int nl = __builtin_ctzll(later);
                        const uint32_t off = __builtin_amdgcn_readlane(rec[2], nl);
                        const MMR_GLOBAL uint32_t* q = levels_dw + (size_t)off * (MMR_H264_LEVELS_PER_MB / 2) + 3 * l;
                        pf[0] = q[0];
                        pf[1] = q[1];
                        pf[2] = q[2];
                        pf_mb = mbx - lane + nl;
                    }
                }
                wave_lds_sync();   // staged record
                intra_residual(sh, job, m, in, l, status);
                const uint32_t need = (uint32_t)min(mbx + 2, mw);
                if (gprev) intra_wait_band(gprev, need, seen, status);
                else if (!first) intra_wait_lds(&prog[w - 1], need, status);
                // left_macroblock_available / top_macroblock_available: inside the picture, same slice
                const bool left_mb = mbx > 0 && prev_slice == slice;
                const bool top_mb = mby > 0 && __builtin_amdgcn_readlane(above_slice, lane) == slice;
                intra_macroblock(sh, job, mbx, mby, m, in, left_mb, top_mb, left_in_lds, above,
                                 __builtin_amdgcn_readlane(above3, lane), gself != nullptr, l, status);
                left_in_lds = true;
                if (!last) {
                    // Hand the bottom row to the row below through the ring; its slot was last read
                    // by the row below's macroblock mbx - kIntraRing + 1.
                    const int free_after = mbx - kIntraRing + 2;
                    if (free_after > 0) intra_wait_lds(&prog[w + 1], (uint32_t)free_after, status);
                    const int slot = mbx % kIntraRing;
                    if (l < 16) sh.ring_y[slot][l] = sh.ty[16][1 + l];
                    else if (l < 32) sh.ring_c[slot][(l - 16) >> 3][(l - 16) & 7] = sh.tc[(l - 16) >> 3][8][1 + ((l - 16) & 7)];
                }
            }
            publish_to((uint32_t)mw, true);
        }
        __syncthreads();
    }
}

hipError_t launch_h264_intra(const H264Job* d_jobs, uint32_t job_count, uint32_t total_bands, uint32_t* d_counters, uint32_t groups,
                             uint32_t publish, uint32_t* d_status, hipStream_t s) {
    if (job_count == 0 || total_bands == 0) return hipSuccess;
    if (groups > total_bands) groups = total_bands;
    if (groups == 0) groups = 1;
    if (publish == 0) publish = 1;
    hipLaunchKernelGGL(h264_intra_kernel, dim3(groups), dim3(kIntraRows * 64), 0, s, d_jobs, job_count, total_bands, d_counters,
                       publish, d_status);
    return hipGetLastError();
}

uint32_t h264_intra_band_rows() { return kIntraRows; }

}  // namespace mmr
