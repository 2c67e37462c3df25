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

// Samples x-2 .. x+6 of reference row y (reference_sample clamps both coordinates into the
// plane). When the 9 samples lie inside the row and the plane is 4-byte aligned they come from the
// three aligned dwords that hold them; otherwise from per-sample clamped loads.
MMR_DEV void load_window_row(const MMR_GLOBAL uint8_t* __restrict__ p, int W, int rowoff, int x, bool fast, int win[9]) {
    if (fast) {
        const int c0 = x - 2;
        const MMR_GLOBAL uint32_t* q = reinterpret_cast<const MMR_GLOBAL uint32_t*>(p + rowoff + (c0 & ~3));
        const uint32_t w0 = q[0], w1 = q[1], w2 = q[2], o = (uint32_t)(c0 & 3);
        const uint32_t a0 = __builtin_amdgcn_alignbyte(w1, w0, o), a1 = __builtin_amdgcn_alignbyte(w2, w1, o);
        const uint32_t a2 = __builtin_amdgcn_alignbyte(w2, w2, o);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            win[i] = (a0 >> (8 * i)) & 0xff;
            win[4 + i] = (a1 >> (8 * i)) & 0xff;
        }
        win[8] = a2 & 0xff;
    } else {
#pragma unroll
        for (int i = 0; i < 9; ++i) win[i] = p[rowoff + clampi(x - 2 + i, 0, W - 1)];
    }
}

// Four consecutive luma samples (x_q4 + 4c, y_q4), c = 0..3, of luma_qpel (decoder.rs:13770-13829)
// with reference_sample's edge clamp. The 6x9 integer window around the row is loaded once and
// the half-sample terms are formed from registers exactly as the reference forms them.
MMR_DEV void luma_qpel_row4(const uint8_t* pg, int W, int H, int x_q4, int y_q4, bool aligned, uint32_t out[4]) {
    const MMR_GLOBAL uint8_t* p = as_global(pg);
    const int x = div4(x_q4), y = div4(y_q4), fx = x_q4 & 3, fy = y_q4 & 3;
    const bool fast = aligned && x >= 2 && x + 6 < W;
    int rows[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) rows[j] = clampi(y - 2 + j, 0, H - 1) * W;
    int win[6][9];
    if (fx == 0 && fy == 0) {
        load_window_row(p, W, rows[2], x, fast, win[2]);
#pragma unroll
        for (int c = 0; c < 4; ++c) out[c] = win[2][c + 2];
        return;
    }
    if (fy == 0) {
        load_window_row(p, W, rows[2], x, fast, win[2]);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int raw = win[2][c] - 5 * win[2][c + 1] + 20 * win[2][c + 2] + 20 * win[2][c + 3] - 5 * win[2][c + 4] + win[2][c + 5];
            const uint32_t b = (uint32_t)clip255((raw + 16) >> 5);
            out[c] = fx == 2 ? b : (fx == 1 ? avg2((uint32_t)win[2][c + 2], b) : avg2(b, (uint32_t)win[2][c + 3]));
        }
        return;
    }
#pragma unroll
    for (int j = 0; j < 6; ++j) load_window_row(p, W, rows[j], x, fast, win[j]);
    if (fx == 0) {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int i = c + 2;
            const int raw = win[0][i] - 5 * win[1][i] + 20 * win[2][i] + 20 * win[3][i] - 5 * win[4][i] + win[5][i];
            const uint32_t h = (uint32_t)clip255((raw + 16) >> 5);
            out[c] = fy == 2 ? h : (fy == 1 ? avg2((uint32_t)win[2][i], h) : avg2(h, (uint32_t)win[3][i]));
        }
        return;
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        int rawh[6];
#pragma unroll
        for (int j = 0; j < 6; ++j)
            rawh[j] = win[j][c] - 5 * win[j][c + 1] + 20 * win[j][c + 2] + 20 * win[j][c + 3] - 5 * win[j][c + 4] + win[j][c + 5];
        const uint32_t b0 = (uint32_t)clip255((rawh[2] + 16) >> 5);   // half_horizontal(x, y)
        const uint32_t b1 = (uint32_t)clip255((rawh[3] + 16) >> 5);   // half_horizontal(x, y + 1)
        const int i0 = c + 2, i1 = c + 3;
        const uint32_t h0 = (uint32_t)clip255((win[0][i0] - 5 * win[1][i0] + 20 * win[2][i0] + 20 * win[3][i0] - 5 * win[4][i0] + win[5][i0] + 16) >> 5);
        const uint32_t h1 = (uint32_t)clip255((win[0][i1] - 5 * win[1][i1] + 20 * win[2][i1] + 20 * win[3][i1] - 5 * win[4][i1] + win[5][i1] + 16) >> 5);
        const uint32_t j = (uint32_t)clip255((rawh[0] - 5 * rawh[1] + 20 * rawh[2] + 20 * rawh[3] - 5 * rawh[4] + rawh[5] + 512) >> 10);
        uint32_t r;
        switch (fy * 4 + fx) {
        case 5: r = avg2(b0, h0); break;
        case 6: r = avg2(b0, j); break;
        case 7: r = avg2(b0, h1); break;
        case 9: r = avg2(h0, j); break;
        case 10: r = j; break;
        case 11: r = avg2(j, h1); break;
        case 13: r = avg2(h0, b1); break;
        case 14: r = avg2(j, b1); break;
        default: r = avg2(h1, b1); break;  // case 15
        }
        out[c] = r;
    }
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
        luma_qpel_row4(job.ref_planes[la][ra][0], W, H, ox * 4 + mvx, oy * 4 + mvy, job.refs_aligned, a);
#pragma unroll
        for (int cc = 0; cc < 2; ++cc) sa[cc] = chroma_epel(job.ref_planes[la][ra][1 + cc], CW, CH, cxs * 8 + mvx, cys * 8 + mvy);
        if (bi) {
            uint32_t b[4];
            luma_qpel_row4(job.ref_planes[1][r1][0], W, H, ox * 4 + blk.mv[1][0], oy * 4 + blk.mv[1][1], job.refs_aligned, b);
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
