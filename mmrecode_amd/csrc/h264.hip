// H.264 inter reconstruction of frame pictures: quarter-sample luma / eighth-sample chroma
// interpolation, explicit / implicit / default (bi-)prediction weights, 4x4 and 8x8 inverse
// transforms with dequantisation, chroma DC Hadamard, residual add.
//
// Reference (crates/codecs/h264/src/decoder.rs): predict_inter_partition :11602-11656,
// predict_bi_inter_partition :11901-12119, predict_weighted_luma/chroma_partition :13640-13768,
// luma_qpel / half_* / chroma_epel / reference_sample / apply_(bi_)prediction_weight :13770-13926,
// transform_residual_4x4 / _8x8 / inverse_transform_8 / level_scale_* :14185-14442,
// transform_chroma_dc / chroma_qp :14102-14140, add_luma_residual_blocks / _8x8_blocks /
// add_chroma_residual :12630-12789.
//
// One wave64 per macroblock, four macroblocks per 256-thread workgroup. Phases, separated by
// workgroup barriers (every wave passes every barrier):
//   1. prediction: lane l predicts row (l & 3) of luma 4x4 block (l >> 2) and one chroma sample of
//      each of Cb / Cr for the same block, into LDS;
//   2. residual front: 4x4 MBs: lanes 0-15 run a whole luma 4x4 block each; 8x8 MBs: all lanes
//      dequantise the 256 luma levels into LDS; lanes 16/17 (4x4) or 0/1 (8x8) run the chroma DC;
//   3. 8x8 row transforms (lanes 0-31) and chroma AC blocks (lanes 32-39);
//   4. 8x8 column transforms;
//   5. every lane adds residual to prediction for 4 luma + 2 chroma samples and stores them.
// Transform arithmetic is 64-bit like the reference, so results match for every level value the
// reference accepts.
#include "device_common.h"
#include "h264_residual.h"
#include "recon_internal.h"

namespace mmr {


MMR_DEV int div4(int v) { return v >> 2; }   // div_euclid(4) for two's complement
MMR_DEV int div8(int v) { return v >> 3; }

MMR_DEV uint32_t avg2(uint32_t a, uint32_t b) { return (a + b + 1) >> 1; }

// Sample pairs (w[k], w[k+1]) of a 9-sample row window held in a0 (samples 0-3), a1 (4-7) and a2
// (sample 8, byte 0), zero-extended to two 16-bit halves: P[k] for k = 0..7.
MMR_DEV void window_pairs(uint32_t a0, uint32_t a1, uint32_t a2, uint32_t P[8]) {
    P[0] = __builtin_amdgcn_perm(a1, a0, 0x0c010c00u);
    P[1] = __builtin_amdgcn_perm(a1, a0, 0x0c020c01u);
    P[2] = __builtin_amdgcn_perm(a1, a0, 0x0c030c02u);
    P[3] = __builtin_amdgcn_perm(a1, a0, 0x0c040c03u);
    P[4] = __builtin_amdgcn_perm(a2, a1, 0x0c010c00u);
    P[5] = __builtin_amdgcn_perm(a2, a1, 0x0c020c01u);
    P[6] = __builtin_amdgcn_perm(a2, a1, 0x0c030c02u);
    P[7] = __builtin_amdgcn_perm(a2, a1, 0x0c040c03u);
}

MMR_DEV mmr_v2i16 as_v2i16(uint32_t v) {
    mmr_v2i16 r;
    __builtin_memcpy(&r, &v, 4);
    return r;
}
MMR_DEV uint32_t as_u32(mmr_v2i16 v) {
    uint32_t r;
    __builtin_memcpy(&r, &v, 4);
    return r;
}

// The six-tap sum a - 5b + 20c + 20d - 5e + f of luma_qpel's half-sample terms, on two 16-bit
// halves at once (sample bytes in: every sum lies in [-2550, 10710], so the 16-bit results are
// exact). Written as the packed instructions themselves: the compiler otherwise splits the packed
// sums whose halves have different uses into 16-bit scalar mads, which measured wrong on gfx950.
MMR_DEV mmr_v2i16 tap6_pairs(uint32_t a, uint32_t b, uint32_t c, uint32_t d, uint32_t e, uint32_t f) {
    uint32_t af, be, cd, r;
    asm("v_pk_add_u16 %0, %1, %2" : "=v"(af) : "v"(a), "v"(f));
    asm("v_pk_add_u16 %0, %1, %2" : "=v"(be) : "v"(b), "v"(e));
    asm("v_pk_add_u16 %0, %1, %2" : "=v"(cd) : "v"(c), "v"(d));
    asm("v_pk_mad_u16 %0, %1, 20, %2 op_sel_hi:[1,0,1]" : "=v"(r) : "v"(cd), "v"(af));
    asm("v_pk_mad_u16 %0, %1, -5, %2 op_sel_hi:[1,0,1]" : "=v"(r) : "v"(be), "v"(r));
    return as_v2i16(r);
}

// clip255((v + 16) >> 5) on both halves.
MMR_DEV uint32_t half_clip_pairs(mmr_v2i16 v) {
    v = (v + (mmr_v2i16){16, 16}) >> (mmr_v2i16){5, 5};
    v = __builtin_elementwise_max(v, (mmr_v2i16){0, 0});
    v = __builtin_elementwise_min(v, (mmr_v2i16){255, 255});
    return as_u32(v);
}

// Low bytes of the halves of (c0, c1) and (c2, c3) as the bytes c0..c3 of a dword.
MMR_DEV uint32_t pack_pairs(uint32_t lo, uint32_t hi) { return __builtin_amdgcn_perm(hi, lo, 0x06040200u); }

// Four consecutive luma samples (x_q4 + 4c, y_q4), c = 0..3, of luma_qpel (decoder.rs:13770-13829)
// with reference_sample's edge clamp, as the bytes of one dword. The 6x9 integer window around
// the row is loaded in one batch (three aligned dwords per row inside the picture, clamped single
// samples at its edges), every half-sample term the reference may use is formed from it without
// branches -- b (rows y, y+1), h (columns x, x+1), j -- exactly as the reference forms it, and the
// quarter position picks the pair that is averaged ((a + b + 1) >> 1; a lone term is averaged
// with itself).
MMR_DEV uint32_t luma_qpel_row4(const uint8_t* pg, int W, int H, int x_q4, int y_q4, bool aligned) {
    const MMR_GLOBAL uint8_t* p = as_global(pg);
    const int x = div4(x_q4), y = div4(y_q4), fx = x_q4 & 3, fy = y_q4 & 3;
    const bool fast = aligned && x >= 2 && x + 6 < W;
    int rows[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) rows[j] = clampi(y - 2 + j, 0, H - 1) * W;
    const int c0 = x - 2;
    uint32_t a0[6], a1[6], a2[6];
    if (fast) {
        // The three dwords end inside the row: c0 & ~3 <= c0 and c0 + 8 < W, a multiple of 4.
        uint32_t d[6][3];
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            const MMR_GLOBAL uint32_t* q = reinterpret_cast<const MMR_GLOBAL uint32_t*>(p + rows[j] + (c0 & ~3));
            d[j][0] = q[0];
            d[j][1] = q[1];
            d[j][2] = q[2];
        }
        const uint32_t o = (uint32_t)(c0 & 3);
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            a0[j] = __builtin_amdgcn_alignbyte(d[j][1], d[j][0], o);
            a1[j] = __builtin_amdgcn_alignbyte(d[j][2], d[j][1], o);
            a2[j] = __builtin_amdgcn_alignbyte(d[j][2], d[j][2], o);
        }
    } else {
        uint32_t s[6][9];
#pragma unroll
        for (int j = 0; j < 6; ++j)
#pragma unroll
            for (int i = 0; i < 9; ++i) s[j][i] = p[rows[j] + clampi(c0 + i, 0, W - 1)];
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            a0[j] = s[j][0] | (s[j][1] << 8) | (s[j][2] << 16) | (s[j][3] << 24);
            a1[j] = s[j][4] | (s[j][5] << 8) | (s[j][6] << 16) | (s[j][7] << 24);
            a2[j] = s[j][8];
        }
    }
    uint32_t P[6][8];
#pragma unroll
    for (int j = 0; j < 6; ++j) window_pairs(a0[j], a1[j], a2[j], P[j]);
    // Horizontal sums of rows y and y + 1 (window rows 2, 3) for columns (0, 1) and (2, 3).
    mmr_v2i16 rh[2][2];
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int q = 0; q < 2; ++q)
            rh[j][q] = tap6_pairs(P[2 + j][2 * q], P[2 + j][2 * q + 1], P[2 + j][2 * q + 2], P[2 + j][2 * q + 3], P[2 + j][2 * q + 4],
                                  P[2 + j][2 * q + 5]);
    // Vertical sums of every window column: pairs (0, 1), (2, 3), (4, 5), (6, 7) and (7, 8).
    mmr_v2i16 rv[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) {
        const int k = q < 4 ? 2 * q : 7;
        rv[q] = tap6_pairs(P[0][k], P[1][k], P[2][k], P[3][k], P[4][k], P[5][k]);
    }
    const uint32_t b0 = pack_pairs(half_clip_pairs(rh[0][0]), half_clip_pairs(rh[0][1]));   // half_horizontal(x + c, y)
    const uint32_t b1 = pack_pairs(half_clip_pairs(rh[1][0]), half_clip_pairs(rh[1][1]));   // half_horizontal(x + c, y + 1)
    const uint32_t v23 = half_clip_pairs(rv[1]), v45 = half_clip_pairs(rv[2]), v67 = half_clip_pairs(rv[3]);
    const uint32_t h0 = pack_pairs(v23, v45);                                                // half_vertical(x + c, y)
    const uint32_t h1 = pack_pairs(__builtin_amdgcn_alignbyte(v45, v23, 2), __builtin_amdgcn_alignbyte(v67, v45, 2));   // (x + c + 1, y)
    // j: the two-dimensional six-tap sum, here as the horizontal six-tap over the vertical sums of
    // window columns c .. c + 5 (the same integer sum as over the horizontal sums of rows y - 2 ..
    // y + 3, in another order), 32-bit.
    int col[9];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        col[2 * q] = rv[q].x;
        col[2 * q + 1] = rv[q].y;
    }
    col[8] = rv[4].y;
    uint32_t jb = 0;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int v = clip255((col[c] - 5 * col[c + 1] + 20 * col[c + 2] + 20 * col[c + 3] - 5 * col[c + 4] + col[c + 5] + 512) >> 10);
        jb |= (uint32_t)v << (8 * c);
    }
    const uint32_t G = __builtin_amdgcn_alignbyte(a1[2], a0[2], 2);    // integer samples x + c, row y
    const uint32_t Gr = __builtin_amdgcn_alignbyte(a1[2], a0[2], 3);   // x + c + 1, row y
    const uint32_t Gd = __builtin_amdgcn_alignbyte(a1[3], a0[3], 2);   // x + c, row y + 1
    // The averaged pair of each quarter position (decoder.rs:13770-13829).
    uint32_t A, B;
    if (fy == 0) {
        A = fx == 0 ? G : (fx == 1 ? G : b0);
        B = fx == 0 ? G : (fx == 3 ? Gr : b0);
    } else if (fx == 0) {
        A = fy == 1 ? G : h0;
        B = fy == 3 ? Gd : h0;
    } else if (fy == 1) {
        A = b0;
        B = fx == 1 ? h0 : (fx == 2 ? jb : h1);
    } else if (fy == 2) {
        A = fx == 1 ? h0 : jb;
        B = fx == 3 ? h1 : jb;
    } else {
        A = fx == 1 ? h0 : (fx == 2 ? jb : h1);
        B = b1;
    }
    return __builtin_amdgcn_lerp(A, B, 0x01010101u);
}

// chroma_epel (decoder.rs:13851-13886).
MMR_DEV uint32_t chroma_epel(const uint8_t* pg, int W, int H, int x_q8, int y_q8) {
    const MMR_GLOBAL uint8_t* p = as_global(pg);
    const int ix = div8(x_q8), iy = div8(y_q8), fx = x_q8 & 7, fy = y_q8 & 7;
    const int x0 = clampi(ix, 0, W - 1), x1 = clampi(ix + 1, 0, W - 1);
    const int y0 = clampi(iy, 0, H - 1) * W, y1 = clampi(iy + 1, 0, H - 1) * W;
    const int A = p[y0 + x0], B = p[y0 + x1], C = p[y1 + x0], D = p[y1 + x1];
    return (uint32_t)clip255(((8 - fx) * (8 - fy) * A + fx * (8 - fy) * B + (8 - fx) * fy * C + fx * fy * D + 32) >> 6);
}

// apply_prediction_weight (decoder.rs:13900-13908).
MMR_DEV uint32_t weight1(uint32_t s, int d, int w, int o) {
    const int weighted = w * (int)s;
    const int scaled = d == 0 ? weighted : (weighted + (1 << (d - 1))) >> d;
    return (uint32_t)clip255(scaled + o);
}
// apply_bi_prediction_weight (decoder.rs:13910-13926) / average_samples (:3573-3576).
MMR_DEV uint32_t weight2(uint32_t s0, uint32_t s1, const MMR_GLOBAL mmr_recon_h264_bi_weight* bw, bool weighted) {
    if (!weighted) return (s0 + s1 + 1) >> 1;
    const int d = bw->log2_denom;
    const int v = (bw->w0 * (int)s0 + bw->w1 * (int)s1 + (1 << d)) >> (d + 1);
    return (uint32_t)clip255(v + ((bw->o0 + bw->o1 + 1) >> 1));
}


// At least 5 waves per SIMD (96 VGPRs, a few bytes spilled): measured 4 % faster than the 4 waves
// the unconstrained allocation (114 VGPRs) gives.
__global__ __launch_bounds__(256, 5) void h264_mb_kernel(const H264Job* __restrict__ jobs, const uint32_t* __restrict__ tile_job,
                                                       uint32_t tile_base, uint32_t tiles, uint32_t per_xcd,
                                                       uint32_t* __restrict__ status) {
    __shared__ MbShared sh[4];
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    MbShared& S = sh[w];
    // Tiles of 4 macroblocks (one per wave) that never straddle pictures; tile -> picture table.
    const uint32_t local = xcd_tile(blockIdx.x, per_xcd);
    if (local >= tiles) return;   // uniform over the workgroup
    const uint32_t tile = tile_base + local;
    const H264Job& job = jobs[tile_job[tile]];
    const uint32_t mbs = job.width_mbs * job.height_mbs;
    const uint32_t mb_raw = (tile - job.tile_begin) * 4 + (uint32_t)w;
    const bool active = mb_raw < mbs;
    const uint32_t mb = active ? mb_raw : 0;
    mmr_recon_h264_macroblock m;
    {
        static_assert(sizeof(mmr_recon_h264_macroblock) == 12, "macroblock record is 3 dwords");
        const MMR_GLOBAL uint32_t* q = reinterpret_cast<const MMR_GLOBAL uint32_t*>(as_global(job.mbs + mb));
        const uint32_t wd[3] = {q[0], q[1], q[2]};
        __builtin_memcpy(&m, wd, sizeof m);
    }
    // Intra macroblocks of a picture with intra records are reconstructed by the intra pass
    // (h264_intra.hip) after this kernel; without records they are reported below.
    if (active && (m.flags & MMR_H264_MB_INTRA) && job.intra) return;   // uniform over the wave
    const int mby = (int)(mb / job.width_mbs), mbx = (int)(mb - (uint32_t)mby * job.width_mbs);
    const int W = (int)job.width_mbs * 16, H = (int)job.height_mbs * 16, CW = W / 2, CH = H / 2;
    const bool weighted = job.weighted == MMR_H264_WEIGHTED_TABLE;

    // ---- phase 1: prediction
    const int k = l >> 2;                               // H.264 block index
    const int bx = ((k >> 2) & 1) * 2 + (k & 1), by = (k >> 3) * 2 + ((k >> 1) & 1);  // luma_block_position
    if (active) {
        mmr_recon_h264_block blk;
        {
            static_assert(sizeof(mmr_recon_h264_block) == 12, "block record is 3 dwords");
            const MMR_GLOBAL uint32_t* q = reinterpret_cast<const MMR_GLOBAL uint32_t*>(as_global(job.blocks + (size_t)mb * 16 + k));
            const uint32_t wd[3] = {q[0], q[1], q[2]};
            __builtin_memcpy(&blk, wd, sizeof blk);
        }
        int r0 = blk.ref_idx[0], r1 = blk.ref_idx[1];
        // Intra macroblocks without intra records are unsupported; out-of-range reference indices
        // are invalid data. Both are reported through the status word and clamped so the kernel
        // still completes without reading outside the reference list.
        if (l == 0 && (m.flags & MMR_H264_MB_INTRA)) atomicOr(status, kDevStatusUnsupported);
        if (r0 >= (int)job.ref_count[0] || r1 >= (int)job.ref_count[1] || (r0 < 0 && r1 < 0)) {
            atomicOr(status, kDevStatusInvalidData);
            r0 = job.ref_count[0] ? min(max(r0, 0), (int)job.ref_count[0] - 1) : -1;
            r1 = (r0 < 0 && job.ref_count[1]) ? min(max(r1, 0), (int)job.ref_count[1] - 1) : -1;
            if (r0 < 0 && r1 < 0) r0 = 0;   // a picture without references: slot 0 holds an in-bounds stand-in
        }
        const int ox = mbx * 16 + bx * 4, oy = mby * 16 + by * 4 + (l & 3);
        const int cxs = mbx * 8 + bx * 2 + (l & 1), cys = mby * 8 + by * 2 + ((l >> 1) & 1);
        // First prediction: list 0 when used, else list 1. Bi-predicted lanes then form the list 1
        // prediction as well; a wave mixing single-list and bi-predicted blocks runs two
        // interpolations, not three.
        const bool bi = r0 >= 0 && r1 >= 0;
        const int la = r0 >= 0 ? 0 : 1, ra = la ? r1 : r0;
        const int mvx = blk.mv[la][0], mvy = blk.mv[la][1];
        uint32_t ys[4], a[4];
        uint32_t cs[2], sa[2];
        {
            const uint32_t q = luma_qpel_row4(job.ref_planes[la][ra][0], W, H, ox * 4 + mvx, oy * 4 + mvy, job.refs_aligned);
#pragma unroll
            for (int c = 0; c < 4; ++c) a[c] = (q >> (8 * c)) & 0xffu;
        }
#pragma unroll
        for (int cc = 0; cc < 2; ++cc) sa[cc] = chroma_epel(job.ref_planes[la][ra][1 + cc], CW, CH, cxs * 8 + mvx, cys * 8 + mvy);
        if (bi) {
            uint32_t b[4];
            const uint32_t q = luma_qpel_row4(job.ref_planes[1][r1][0], W, H, ox * 4 + blk.mv[1][0], oy * 4 + blk.mv[1][1], job.refs_aligned);
#pragma unroll
            for (int c = 0; c < 4; ++c) b[c] = (q >> (8 * c)) & 0xffu;
            const MMR_GLOBAL mmr_recon_h264_bi_weight* bw = weighted ? as_global(job.bi_weights) + ((size_t)r0 * MMR_H264_MAX_REFS + r1) * 3 : nullptr;
#pragma unroll
            for (int c = 0; c < 4; ++c) ys[c] = weight2(a[c], b[c], bw, weighted);
#pragma unroll
            for (int cc = 0; cc < 2; ++cc) {
                const uint32_t s1 = chroma_epel(job.ref_planes[1][r1][1 + cc], CW, CH, cxs * 8 + blk.mv[1][0], cys * 8 + blk.mv[1][1]);
                cs[cc] = weight2(sa[cc], s1, weighted ? bw + 1 + cc : nullptr, weighted);
            }
        } else {
            const MMR_GLOBAL mmr_recon_h264_weight* wt = weighted ? as_global(job.weights) + ((size_t)la * MMR_H264_MAX_REFS + ra) * 3 : nullptr;
#pragma unroll
            for (int c = 0; c < 4; ++c) ys[c] = wt ? weight1(a[c], wt[0].log2_denom, wt[0].weight, wt[0].offset) : a[c];
#pragma unroll
            for (int cc = 0; cc < 2; ++cc) cs[cc] = wt ? weight1(sa[cc], wt[1 + cc].log2_denom, wt[1 + cc].weight, wt[1 + cc].offset) : sa[cc];
        }
        const int py = by * 4 + (l & 3), px = bx * 4;
#pragma unroll
        for (int c = 0; c < 4; ++c) S.pred_y[py * 16 + px + c] = (uint8_t)ys[c];
        const int cpy = by * 2 + ((l >> 1) & 1), cpx = bx * 2 + (l & 1);
        S.pred_c[0][cpy * 8 + cpx] = (uint8_t)cs[0];
        S.pred_c[1][cpy * 8 + cpx] = (uint8_t)cs[1];
    }

    const bool has_res = active && m.cbp != 0;
    const bool t8 = (m.flags & MMR_H264_MB_TRANSFORM_8X8) != 0;
    const MMR_GLOBAL int16_t* lv = as_global(job.levels) + (size_t)m.coef_offset * MMR_H264_LEVELS_PER_MB;
    const int qp = m.qp;
    const int cqp = chroma_qp(qp, job.chroma_qp_offset[0]);   // first offset for both (decoder.rs:1921)

    // ---- phases 2-4: residual (64-bit like the reference, or 32-bit when the levels are small)
    if (has_res) {
        // Largest |level| of the macroblock: below 250 every intermediate of the 4x4 / 8x8 /
        // chroma paths stays below 2^31 (8x8: |v| * 58 * 255 * 4 growing at most 144-fold through
        // the two 8-point passes; 4x4 and chroma less).
        uint32_t mx = 0;
#pragma unroll
        for (int i = 0; i < MMR_H264_LEVELS_PER_MB / 64; ++i) {
            const int v = lv[l + 64 * i];
            mx = max(mx, (uint32_t)(v < 0 ? -v : v));
        }
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) mx = max(mx, (uint32_t)__shfl_xor((int)mx, d, 64));
        if (mx < 250) residual_phases<int32_t>(S, lv, qp, cqp, t8, l, as_global(job.scaling), kInterLists, status);
        else residual_phases<int64_t>(S, lv, qp, cqp, t8, l, as_global(job.scaling), kInterLists, status);
        if (job.transform_bypass && (qp == 0 || cqp == 0)) {
            const Bypass by = {qp == 0 && !t8, cqp == 0, {0x02020202u, 0x02020202u, 0x02020202u, 0x02020202u}, 2};
            bypass_phases(S, lv, l, by);
        }
    }
    wave_lds_sync();

    // ---- phase 5: residual add + store
    if (!active) return;
    {
        const int row = l >> 2, c0 = (l & 3) * 4;
        uint32_t word = 0;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            int v = S.pred_y[row * 16 + c0 + c];
            if (has_res) v = clip255(v + S.res_y[row * 16 + c0 + c]);
            word |= (uint32_t)v << (8 * c);
        }
        *reinterpret_cast<MMR_GLOBAL uint32_t*>(as_global(job.planes[0]) + (size_t)(mby * 16 + row) * W + mbx * 16 + c0) = word;
    }
    {
        const int comp = l >> 5, row = (l >> 2) & 7, c0 = (l & 3) * 2;
        uint32_t half = 0;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            int v = S.pred_c[comp][row * 8 + c0 + c];
            if (has_res) v = clip255(v + S.res_c[comp][row * 8 + c0 + c]);
            half |= (uint32_t)v << (8 * c);
        }
        *reinterpret_cast<MMR_GLOBAL uint16_t*>(as_global(job.planes[1 + comp]) + (size_t)(mby * 8 + row) * CW + mbx * 8 + c0) = (uint16_t)half;
    }
}

hipError_t launch_h264_mc(const H264Job* d_jobs, const uint32_t* d_tile_job, uint32_t tile_base, uint32_t tiles, uint32_t* d_status,
                          hipStream_t s) {
    if (tiles == 0) return hipSuccess;
    const uint32_t per = xcd_per(tiles);
    hipLaunchKernelGGL(h264_mb_kernel, dim3(per * kXcds), dim3(256), 0, s, d_jobs, d_tile_job, tile_base, tiles, per, d_status);
    return hipGetLastError();
}

}  // namespace mmr
