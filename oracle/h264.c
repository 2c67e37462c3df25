/* Oracle: H.264 inter reconstruction restated from crates/codecs/h264/src/decoder.rs.
 * Test infrastructure only. */
#include <string.h>

#include "oracle.h"

static uint8_t clip_u8(int32_t v) { return (uint8_t)(v < 0 ? 0 : (v > 255 ? 255 : v)); }

/* ZIG_ZAG_4X4 and ZIG_ZAG_8X8 as (row, column), decoder.rs:13928-13945, :13966-14032. */
static const uint8_t ZZ4[16][2] = {{0, 0}, {0, 1}, {1, 0}, {2, 0}, {1, 1}, {0, 2}, {0, 3}, {1, 2},
                                   {2, 1}, {3, 0}, {3, 1}, {2, 2}, {1, 3}, {2, 3}, {3, 2}, {3, 3}};
static const uint8_t ZZ8[64][2] = {
    {0, 0}, {0, 1}, {1, 0}, {2, 0}, {1, 1}, {0, 2}, {0, 3}, {1, 2}, {2, 1}, {3, 0}, {4, 0}, {3, 1}, {2, 2},
    {1, 3}, {0, 4}, {0, 5}, {1, 4}, {2, 3}, {3, 2}, {4, 1}, {5, 0}, {6, 0}, {5, 1}, {4, 2}, {3, 3}, {2, 4},
    {1, 5}, {0, 6}, {0, 7}, {1, 6}, {2, 5}, {3, 4}, {4, 3}, {5, 2}, {6, 1}, {7, 0}, {7, 1}, {6, 2}, {5, 3},
    {4, 4}, {3, 5}, {2, 6}, {1, 7}, {2, 7}, {3, 6}, {4, 5}, {5, 4}, {6, 3}, {7, 2}, {7, 3}, {6, 4}, {5, 5},
    {4, 6}, {3, 7}, {4, 7}, {5, 6}, {6, 5}, {7, 4}, {7, 5}, {6, 6}, {5, 7}, {6, 7}, {7, 6}, {7, 7}};

/* level_scale_4x4, decoder.rs:14424-14442. */
static int32_t level_scale_4x4(int qp, int row, int column, const uint8_t scaling[16]) {
    static const int32_t NORM[6][3] = {{10, 16, 13}, {11, 18, 14}, {13, 20, 16}, {14, 23, 18}, {16, 25, 20}, {18, 29, 23}};
    int category = (row % 2 == 0 && column % 2 == 0) ? 0 : ((row % 2 != 0 && column % 2 != 0) ? 1 : 2);
    return NORM[qp % 6][category] * (int32_t)scaling[row * 4 + column];
}

/* level_scale_8x8, decoder.rs:14310-14337. */
static int32_t level_scale_8x8(int qp, int row, int column, const uint8_t scaling[64]) {
    static const int32_t NORM[6][6] = {{20, 18, 32, 19, 25, 24}, {22, 19, 35, 21, 28, 26}, {26, 23, 42, 24, 33, 31},
                                       {28, 25, 45, 26, 35, 33}, {32, 28, 51, 30, 40, 38}, {36, 32, 58, 34, 46, 43}};
    int rm = row % 4, cm = column % 4, category;
    int re = rm % 2 == 0, ce = cm % 2 == 0;
    if (!re && !ce) category = 1;
    else if (re != ce) category = (rm == 2 || cm == 2) ? 5 : 3;
    else category = (rm == 0 && cm == 0) ? 0 : ((rm == 2 && cm == 2) ? 2 : 4);
    return NORM[qp % 6][category] * (int32_t)scaling[row * 8 + column];
}

/* chroma_qp, decoder.rs:14102-14112. */
int oracle_h264_chroma_qp(int luma_qp, int offset) {
    static const int T[22] = {29, 30, 31, 32, 32, 33, 34, 34, 35, 35, 36, 36, 37, 37, 37, 38, 38, 38, 39, 39, 39, 39};
    int index = luma_qp + offset;
    if (index < 0) index = 0;
    if (index > 51) index = 51;
    return index < 30 ? index : T[index - 30];
}

/* transform_residual_4x4, decoder.rs:14185-14233 (i64 arithmetic). */
int oracle_h264_transform_residual_4x4(const int32_t levels[16], int qp, int dc_already_scaled,
                                       const uint8_t scaling[16], int32_t out[16]) {
    if (qp < 0 || qp > 51) return -1;
    int64_t s[16], h[16];
    for (int i = 0; i < 16; ++i) {
        int row = ZZ4[i][0], column = ZZ4[i][1];
        int64_t value = levels[i];
        if (i == 0 && dc_already_scaled) {
            s[row * 4 + column] = value;
        } else {
            int64_t scale = level_scale_4x4(qp, row, column, scaling);
            if (qp >= 24) s[row * 4 + column] = (value * scale) << (qp / 6 - 4);
            else s[row * 4 + column] = (value * scale + ((int64_t)1 << (3 - qp / 6))) >> (4 - qp / 6);
        }
    }
    for (int row = 0; row < 4; ++row) {
        int b = row * 4;
        int64_t e0 = s[b] + s[b + 2], e1 = s[b] - s[b + 2];
        int64_t e2 = (s[b + 1] >> 1) - s[b + 3], e3 = s[b + 1] + (s[b + 3] >> 1);
        h[b] = e0 + e3;
        h[b + 1] = e1 + e2;
        h[b + 2] = e1 - e2;
        h[b + 3] = e0 - e3;
    }
    for (int column = 0; column < 4; ++column) {
        int64_t g0 = h[column] + h[8 + column], g1 = h[column] - h[8 + column];
        int64_t g2 = (h[4 + column] >> 1) - h[12 + column], g3 = h[4 + column] + (h[12 + column] >> 1);
        int64_t v[4] = {g0 + g3, g1 + g2, g1 - g2, g0 - g3};
        for (int row = 0; row < 4; ++row) {
            int64_t r = (v[row] + 32) >> 6;
            if (r < INT32_MIN || r > INT32_MAX) return -2;
            out[row * 4 + column] = (int32_t)r;
        }
    }
    return 0;
}

/* inverse_transform_8, decoder.rs:14280-14308. */
static void inverse_transform_8(const int64_t in[8], int64_t out[8]) {
    int64_t a0 = in[0] + in[4], a2 = in[0] - in[4];
    int64_t a4 = (in[2] >> 1) - in[6], a6 = in[2] + (in[6] >> 1);
    int64_t b0 = a0 + a6, b2 = a2 + a4, b4 = a2 - a4, b6 = a0 - a6;
    int64_t a1 = -in[3] + in[5] - in[7] - (in[7] >> 1);
    int64_t a3 = in[1] + in[7] - in[3] - (in[3] >> 1);
    int64_t a5 = -in[1] + in[7] + in[5] + (in[5] >> 1);
    int64_t a7 = in[3] + in[5] + in[1] + (in[1] >> 1);
    int64_t b1 = a1 + (a7 >> 2), b7 = a7 - (a1 >> 2), b3 = a3 + (a5 >> 2), b5 = (a3 >> 2) - a5;
    out[0] = b0 + b7;
    out[1] = b2 + b5;
    out[2] = b4 + b3;
    out[3] = b6 + b1;
    out[4] = b6 - b1;
    out[5] = b4 - b3;
    out[6] = b2 - b5;
    out[7] = b0 - b7;
}

/* transform_residual_8x8, decoder.rs:14235-14278. */
int oracle_h264_transform_residual_8x8(const int32_t levels[64], int qp, const uint8_t scaling[64],
                                       int32_t out[64]) {
    if (qp < 0 || qp > 51) return -1;
    int64_t s[64], h[64], t[64];
    for (int i = 0; i < 64; ++i) {
        int row = ZZ8[i][0], column = ZZ8[i][1];
        int64_t value = (int64_t)levels[i] * level_scale_8x8(qp, row, column, scaling);
        int shift = qp / 6;
        s[row * 8 + column] = shift >= 6 ? value << (shift - 6) : (value + ((int64_t)1 << (5 - shift))) >> (6 - shift);
    }
    s[0] += 32;
    for (int row = 0; row < 8; ++row) inverse_transform_8(s + row * 8, h + row * 8);
    for (int column = 0; column < 8; ++column) {
        int64_t in[8], o[8];
        for (int row = 0; row < 8; ++row) in[row] = h[row * 8 + column];
        inverse_transform_8(in, o);
        for (int row = 0; row < 8; ++row) t[row * 8 + column] = o[row];
    }
    for (int i = 0; i < 64; ++i) {
        int64_t r = t[i] >> 6;
        if (r < INT32_MIN || r > INT32_MAX) return -2;
        out[i] = (int32_t)r;
    }
    return 0;
}

/* transform_chroma_dc, decoder.rs:14114-14140. */
int oracle_h264_transform_chroma_dc(const int32_t levels[4], int qp, const uint8_t scaling[16], int32_t out[4]) {
    if (qp < 0 || qp > 39) return -1;
    int64_t c0 = levels[0], c1 = levels[1], c2 = levels[2], c3 = levels[3];
    int64_t f[4] = {c0 + c1 + c2 + c3, c0 - c1 + c2 - c3, c0 + c1 - c2 - c3, c0 - c1 - c2 + c3};
    int64_t scale = level_scale_4x4(qp, 0, 0, scaling);
    for (int i = 0; i < 4; ++i) {
        int64_t scaled = ((f[i] * scale) << (qp / 6)) >> 5;
        if (scaled < INT32_MIN || scaled > INT32_MAX) return -2;
        out[i] = (int32_t)scaled;
    }
    return 0;
}

static int div_euclid(int v, int d) { int q = v / d; if (v % d < 0) q -= 1; return q; }
static int rem_euclid(int v, int d) { int r = v % d; return r < 0 ? r + d : r; }

/* reference_sample, decoder.rs:13888-13894: clamp to the coded plane. */
static uint8_t reference_sample(const uint8_t *plane, int width, int height, int x, int y) {
    if (x < 0) x = 0;
    if (x > width - 1) x = width - 1;
    if (y < 0) y = 0;
    if (y > height - 1) y = height - 1;
    return plane[(size_t)y * width + x];
}

static uint8_t average(uint8_t a, uint8_t b) { return (uint8_t)((a + b + 1) >> 1); }

/* half_horizontal_raw / half_horizontal / half_vertical / half_diagonal, decoder.rs:13831-13849. */
static int32_t half_h_raw(const uint8_t *p, int w, int h, int x, int y) {
    int32_t t[6];
    for (int k = 0; k < 6; ++k) t[k] = reference_sample(p, w, h, x + k - 2, y);
    return t[0] - 5 * t[1] + 20 * t[2] + 20 * t[3] - 5 * t[4] + t[5];
}
static uint8_t half_h(const uint8_t *p, int w, int h, int x, int y) { return clip_u8((half_h_raw(p, w, h, x, y) + 16) >> 5); }
static uint8_t half_v(const uint8_t *p, int w, int h, int x, int y) {
    int32_t t[6];
    for (int k = 0; k < 6; ++k) t[k] = reference_sample(p, w, h, x, y + k - 2);
    return clip_u8((t[0] - 5 * t[1] + 20 * t[2] + 20 * t[3] - 5 * t[4] + t[5] + 16) >> 5);
}
static uint8_t half_d(const uint8_t *p, int w, int h, int x, int y) {
    int32_t t[6];
    for (int k = 0; k < 6; ++k) t[k] = half_h_raw(p, w, h, x, y + k - 2);
    return clip_u8((t[0] - 5 * t[1] + 20 * t[2] + 20 * t[3] - 5 * t[4] + t[5] + 512) >> 10);
}

/* luma_qpel, decoder.rs:13770-13829: the 16 fractional cases. */
uint8_t oracle_h264_luma_qpel(const uint8_t *p, int w, int h, int x_q4, int y_q4) {
    int x = div_euclid(x_q4, 4), y = div_euclid(y_q4, 4);
    int fx = rem_euclid(x_q4, 4), fy = rem_euclid(y_q4, 4);
    uint8_t G = reference_sample(p, w, h, x, y);
    switch (fy * 4 + fx) {
    case 0: return G;
    case 1: return average(G, half_h(p, w, h, x, y));
    case 2: return half_h(p, w, h, x, y);
    case 3: return average(half_h(p, w, h, x, y), reference_sample(p, w, h, x + 1, y));
    case 4: return average(G, half_v(p, w, h, x, y));
    case 8: return half_v(p, w, h, x, y);
    case 12: return average(half_v(p, w, h, x, y), reference_sample(p, w, h, x, y + 1));
    case 5: return average(half_h(p, w, h, x, y), half_v(p, w, h, x, y));
    case 6: return average(half_h(p, w, h, x, y), half_d(p, w, h, x, y));
    case 7: return average(half_h(p, w, h, x, y), half_v(p, w, h, x + 1, y));
    case 9: return average(half_v(p, w, h, x, y), half_d(p, w, h, x, y));
    case 10: return half_d(p, w, h, x, y);
    case 11: return average(half_d(p, w, h, x, y), half_v(p, w, h, x + 1, y));
    case 13: return average(half_v(p, w, h, x, y), half_h(p, w, h, x, y + 1));
    case 14: return average(half_d(p, w, h, x, y), half_h(p, w, h, x, y + 1));
    default: return average(half_v(p, w, h, x + 1, y), half_h(p, w, h, x, y + 1));
    }
}

/* chroma_epel, decoder.rs:13851-13886. */
uint8_t oracle_h264_chroma_epel(const uint8_t *p, int w, int h, int x_q8, int y_q8) {
    int ix = div_euclid(x_q8, 8), iy = div_euclid(y_q8, 8);
    int fx = rem_euclid(x_q8, 8), fy = rem_euclid(y_q8, 8);
    int32_t A = reference_sample(p, w, h, ix, iy), B = reference_sample(p, w, h, ix + 1, iy);
    int32_t C = reference_sample(p, w, h, ix, iy + 1), D = reference_sample(p, w, h, ix + 1, iy + 1);
    return clip_u8(((8 - fx) * (8 - fy) * A + fx * (8 - fy) * B + (8 - fx) * fy * C + fx * fy * D + 32) >> 6);
}

/* apply_prediction_weight, decoder.rs:13900-13908. */
uint8_t oracle_h264_apply_weight(uint8_t sample, int d, int w, int o) {
    int32_t weighted = w * (int32_t)sample;
    int32_t scaled = d == 0 ? weighted : (weighted + (1 << (d - 1))) >> d;
    return clip_u8(scaled + o);
}

/* apply_bi_prediction_weight + average_samples, decoder.rs:13910-13926, :3573-3576. */
uint8_t oracle_h264_apply_bi_weight(uint8_t s0, uint8_t s1, int d, int w0, int w1, int o0, int o1,
                                    int explicit_weights) {
    if (!explicit_weights) return (uint8_t)((s0 + s1 + 1) >> 1);
    int32_t weighted = w0 * (int32_t)s0 + w1 * (int32_t)s1;
    int32_t scaled = (weighted + (1 << d)) >> (d + 1);
    int32_t offset = (o0 + o1 + 1) >> 1;
    return clip_u8(scaled + offset);
}

/* implicit_b_prediction_weights, decoder.rs:3577-3620. */
void oracle_h264_implicit_weights(int32_t poc0, int32_t poc1, int32_t cur, int lt0, int lt1, int32_t *w0, int32_t *w1) {
    int64_t td = (int64_t)poc1 - poc0;
    if (td < -128) td = -128;
    if (td > 127) td = 127;
    int32_t weight1;
    if (lt0 || lt1 || td == 0) {
        weight1 = 32;
    } else {
        int64_t tb = (int64_t)cur - poc0;
        if (tb < -128) tb = -128;
        if (tb > 127) tb = 127;
        int64_t tx = (16384 + (td < 0 ? -td : td) / 2) / td;
        int64_t ds = (tb * tx + 32) >> 6;
        if (ds < -1024) ds = -1024;
        if (ds > 1023) ds = 1023;
        int64_t candidate = ds >> 2;
        weight1 = (candidate >= -64 && candidate <= 128) ? (int32_t)candidate : 32;
    }
    *w0 = 64 - weight1;
    *w1 = weight1;
}

static const uint8_t FLAT16[64] = {16, 16, 16, 16, 16, 16, 16, 16, 16, 16, 16, 16, 16, 16, 16, 16,
                                   16, 16, 16, 16, 16, 16, 16, 16, 16, 16, 16, 16, 16, 16, 16, 16,
                                   16, 16, 16, 16, 16, 16, 16, 16, 16, 16, 16, 16, 16, 16, 16, 16,
                                   16, 16, 16, 16, 16, 16, 16, 16, 16, 16, 16, 16, 16, 16, 16, 16};
static const uint8_t *scaling4(const mmr_recon_h264_picture *p, int i) { return p->scaling_lists ? p->scaling_lists + 16 * i : FLAT16; }
static const uint8_t *scaling8(const mmr_recon_h264_picture *p, int i) { return p->scaling_lists ? p->scaling_lists + 96 + 64 * i : FLAT16; }

/* luma_block_position, decoder.rs:13015-13022. */
static void luma_block_position(int index, int *bx, int *by) {
    int group = index / 4, within = index % 4;
    *bx = (group % 2) * 2 + within % 2;
    *by = (group / 2) * 2 + within / 2;
}

static void add_residual(uint8_t *plane, int stride, int ox, int oy, int n, const int32_t *res) {
    for (int y = 0; y < n; ++y)
        for (int x = 0; x < n; ++x) {
            size_t i = (size_t)(oy + y) * stride + ox + x;
            plane[i] = clip_u8((int32_t)plane[i] + res[y * n + x]);
        }
}

/* One 4x4 partition: predict_inter_partition (decoder.rs:11602-11656) with
 * predict_weighted_luma/chroma_partition (:13640-13768), or predict_bi_inter_partition
 * (:11901-12119). The reference predicts whole partitions; every predicted sample depends only
 * on its own position and the partition's vector and weights, so predicting each 4x4 block of a
 * partition with the partition's motion gives the same samples. */
static int predict_block(const mmr_recon_h264_picture *p, int mb, int k) {
    const mmr_recon_h264_block *blk = &p->blocks[(size_t)mb * 16 + k];
    int W = (int)p->width_mbs * 16, H = (int)p->height_mbs * 16, CW = W / 2, CH = H / 2;
    int bx, by;
    luma_block_position(k, &bx, &by);
    int mbx = mb % (int)p->width_mbs, mby = mb / (int)p->width_mbs;
    int ox = mbx * 16 + bx * 4, oy = mby * 16 + by * 4, cx = mbx * 8 + bx * 2, cy = mby * 8 + by * 2;
    int r0 = blk->ref_idx[0], r1 = blk->ref_idx[1];
    if (r0 < 0 && r1 < 0) return -1;
    if ((r0 >= 0 && (uint32_t)r0 >= p->ref_count[0]) || (r1 >= 0 && (uint32_t)r1 >= p->ref_count[1])) return -1;
    if (r0 >= 0 && r1 >= 0) {
        const mmr_recon_h264_reference *R0 = &p->refs[0][r0], *R1 = &p->refs[1][r1];
        int expl = p->weighted_mode == MMR_H264_WEIGHTED_TABLE;
        for (int c = 0; c < 3; ++c) {
            const mmr_recon_h264_bi_weight *bw = p->bi_weights ? &p->bi_weights[((size_t)r0 * MMR_H264_MAX_REFS + r1) * 3 + c] : 0;
            int d = bw ? bw->log2_denom : 0, w0 = bw ? bw->w0 : 1, w1 = bw ? bw->w1 : 1, o0 = bw ? bw->o0 : 0, o1 = bw ? bw->o1 : 0;
            if (c == 0) {
                for (int y = 0; y < 4; ++y)
                    for (int x = 0; x < 4; ++x) {
                        uint8_t s0 = oracle_h264_luma_qpel(R0->planes[0], W, H, (ox + x) * 4 + blk->mv[0][0], (oy + y) * 4 + blk->mv[0][1]);
                        uint8_t s1 = oracle_h264_luma_qpel(R1->planes[0], W, H, (ox + x) * 4 + blk->mv[1][0], (oy + y) * 4 + blk->mv[1][1]);
                        p->planes[0][(size_t)(oy + y) * W + ox + x] = oracle_h264_apply_bi_weight(s0, s1, d, w0, w1, o0, o1, expl);
                    }
            } else {
                for (int y = 0; y < 2; ++y)
                    for (int x = 0; x < 2; ++x) {
                        uint8_t s0 = oracle_h264_chroma_epel(R0->planes[c], CW, CH, (cx + x) * 8 + blk->mv[0][0], (cy + y) * 8 + blk->mv[0][1]);
                        uint8_t s1 = oracle_h264_chroma_epel(R1->planes[c], CW, CH, (cx + x) * 8 + blk->mv[1][0], (cy + y) * 8 + blk->mv[1][1]);
                        p->planes[c][(size_t)(cy + y) * CW + cx + x] = oracle_h264_apply_bi_weight(s0, s1, d, w0, w1, o0, o1, expl);
                    }
            }
        }
        return 0;
    }
    int list = r0 >= 0 ? 0 : 1, r = r0 >= 0 ? r0 : r1;
    const mmr_recon_h264_reference *R = &p->refs[list][r];
    for (int c = 0; c < 3; ++c) {
        const mmr_recon_h264_weight *wt = (p->weighted_mode == MMR_H264_WEIGHTED_TABLE && p->weights)
                                              ? &p->weights[((size_t)list * MMR_H264_MAX_REFS + r) * 3 + c] : 0;
        int d = wt ? wt->log2_denom : 0, w = wt ? wt->weight : 1, o = wt ? wt->offset : 0;
        if (c == 0) {
            for (int y = 0; y < 4; ++y)
                for (int x = 0; x < 4; ++x) {
                    uint8_t s = oracle_h264_luma_qpel(R->planes[0], W, H, (ox + x) * 4 + blk->mv[list][0], (oy + y) * 4 + blk->mv[list][1]);
                    p->planes[0][(size_t)(oy + y) * W + ox + x] = oracle_h264_apply_weight(s, d, w, o);
                }
        } else {
            for (int y = 0; y < 2; ++y)
                for (int x = 0; x < 2; ++x) {
                    uint8_t s = oracle_h264_chroma_epel(R->planes[c], CW, CH, (cx + x) * 8 + blk->mv[list][0], (cy + y) * 8 + blk->mv[list][1]);
                    p->planes[c][(size_t)(cy + y) * CW + cx + x] = oracle_h264_apply_weight(s, d, w, o);
                }
        }
    }
    return 0;
}

void oracle_h264_bypass_residual_4x4(const int32_t levels[16], int intra_mode, int32_t out[16]);
static void bypass_chroma(const int16_t *lv, int comp, int mode, int32_t res[4][16]);

/* decode_inter_residuals' reconstruction half, decoder.rs:8097-8173: luma 4x4 blocks
 * (add_luma_residual_blocks :12630-12682, scaling list 3) or 8x8 blocks
 * (add_luma_residual_8x8_blocks :12684-12714, 8x8 list 1), then chroma
 * (add_chroma_residual :12716-12789, inter lists 4 and 5) at
 * chroma_qp(qp, chroma_qp_index_offset); the reference passes the first chroma offset for
 * both components here (decoder.rs:1921), so the oracle does too. */
static int add_residuals(const mmr_recon_h264_picture *p, int mb) {
    const mmr_recon_h264_macroblock *m = &p->macroblocks[mb];
    if (m->cbp == 0) return 0;
    const int16_t *lv = p->levels + (size_t)m->coef_offset * MMR_H264_LEVELS_PER_MB;
    int W = (int)p->width_mbs * 16, CW = W / 2;
    int mbx = mb % (int)p->width_mbs, mby = mb / (int)p->width_mbs;
    int qp = m->qp;
    int32_t levels[64], res[64];
    if (m->flags & MMR_H264_MB_TRANSFORM_8X8) {
        for (int g = 0; g < 4; ++g) {
            for (int i = 0; i < 64; ++i) levels[i] = lv[g * 64 + i];
            if (oracle_h264_transform_residual_8x8(levels, qp, scaling8(p, 1), res)) return -2;
            add_residual(p->planes[0], W, mbx * 16 + (g % 2) * 8, mby * 16 + (g / 2) * 8, 8, res);
        }
    } else {
        for (int k = 0; k < 16; ++k) {
            for (int i = 0; i < 16; ++i) levels[i] = lv[k * 16 + i];
            if (p->transform_bypass && qp == 0) oracle_h264_bypass_residual_4x4(levels, 2, res);
            else if (oracle_h264_transform_residual_4x4(levels, qp, 0, scaling4(p, 3), res)) return -2;
            int bx, by;
            luma_block_position(k, &bx, &by);
            add_residual(p->planes[0], W, mbx * 16 + bx * 4, mby * 16 + by * 4, 4, res);
        }
    }
    int cqp = oracle_h264_chroma_qp(qp, p->chroma_qp_offset[0]);
    if (p->transform_bypass && cqp == 0) {
        for (int comp = 0; comp < 2; ++comp) {
            int32_t br[4][16];
            bypass_chroma(lv, comp, 2, br);
            for (int b = 0; b < 4; ++b) add_residual(p->planes[1 + comp], CW, mbx * 8 + (b % 2) * 4, mby * 8 + (b / 2) * 4, 4, br[b]);
        }
        return 0;
    }
    for (int comp = 0; comp < 2; ++comp) {
        const uint8_t *sl = scaling4(p, 3 + comp + 1);
        int32_t dc_levels[4], dc[4];
        for (int i = 0; i < 4; ++i) dc_levels[i] = lv[256 + comp * 4 + i];
        if (oracle_h264_transform_chroma_dc(dc_levels, cqp, sl, dc)) return -2;
        for (int b = 0; b < 4; ++b) {
            int32_t co[16];
            co[0] = dc[b];
            for (int i = 1; i < 16; ++i) co[i] = lv[264 + (comp * 4 + b) * 15 + (i - 1)];
            if (oracle_h264_transform_residual_4x4(co, cqp, 1, sl, res)) return -2;
            add_residual(p->planes[1 + comp], CW, mbx * 8 + (b % 2) * 4, mby * 8 + (b / 2) * 4, 4, res);
        }
    }
    return 0;
}

/* transform_bypass_residual_4x4, decoder.rs:14339-14380: levels placed by the zig-zag scan, then
 * for intra_mode 0 (vertical) each row adds the one above, for 1 (horizontal) each column the one
 * to its left; 2: placement only. */
void oracle_h264_bypass_residual_4x4(const int32_t levels[16], int intra_mode, int32_t out[16]) {
    for (int i = 0; i < 16; ++i) out[ZZ4[i][0] * 4 + ZZ4[i][1]] = levels[i];
    if (intra_mode == 0)
        for (int r = 1; r < 4; ++r)
            for (int c = 0; c < 4; ++c) out[r * 4 + c] += out[(r - 1) * 4 + c];
    if (intra_mode == 1)
        for (int r = 0; r < 4; ++r)
            for (int c = 1; c < 4; ++c) out[r * 4 + c] += out[r * 4 + c - 1];
}

/* The chroma residual of one component with the transform bypassed (add_chroma_residual's bypass
 * branch, decoder.rs:12739-12764): DC levels taken as they are, each 4x4 block through
 * oracle_h264_bypass_residual_4x4 with the mode of the chroma intra prediction (1 horizontal -> 1,
 * 2 vertical -> 0, otherwise and for inter 2), then continue_transform_bypass_chroma
 * (:14382-14420): the lower blocks add the upper block's last row (vertical), the right blocks
 * the left block's last column (horizontal). res[b] in raster order per block. */
static void bypass_chroma(const int16_t *lv, int comp, int mode, int32_t res[4][16]) {
    for (int b = 0; b < 4; ++b) {
        int32_t co[16];
        co[0] = lv[256 + comp * 4 + b];
        for (int i = 1; i < 16; ++i) co[i] = lv[264 + (comp * 4 + b) * 15 + (i - 1)];
        oracle_h264_bypass_residual_4x4(co, mode, res[b]);
    }
    if (mode == 0)
        for (int b = 2; b < 4; ++b)
            for (int r = 0; r < 4; ++r)
                for (int c = 0; c < 4; ++c) res[b][r * 4 + c] += res[b - 2][12 + c];
    if (mode == 1)
        for (int b = 1; b < 4; b += 2)
            for (int r = 0; r < 4; ++r) {
                const int32_t carry = res[b - 1][r * 4 + 3];
                for (int c = 0; c < 4; ++c) res[b][r * 4 + c] += carry;
            }
}

static void place(uint8_t *plane, int stride, int ox, int oy, int n, const uint8_t *v) {
    for (int y = 0; y < n; ++y) memcpy(plane + (size_t)(oy + y) * stride + ox, v + y * n, (size_t)n);
}

/* luma_block_index: block index of the 4x4 block at (bx, by). */
static int luma_block_index(int bx, int by) { return (by / 2) * 8 + (bx / 2) * 4 + (by % 2) * 2 + bx % 2; }

/* One intra macroblock, as the reference's intra paths do it (decoder.rs:5690-5722, :5835-5848,
 * :5955-5961): I_PCM samples (place_pcm_macroblock :10982-11004); Intra4x4 blocks in block order,
 * each predicted then given its residual (reconstruct_intra4_luma_block :12377-12480, list 0);
 * Intra8x8 likewise (:12482-12578, 8x8 list 0); Intra16x16 predicted whole, then the DC
 * transform and 16 residual blocks (predict_intra16_macroblock :12224-12279,
 * add_intra16_luma_residual :12580-12628); then chroma prediction (predict_chroma_macroblock
 * :12281-12330) and the chroma residual with the intra lists 1 and 2 (add_chroma_residual
 * :12716-12789). Neighbour availability: same slice, inside the picture (:10801-10814). As for
 * inter macroblocks, a level record (present when cbp != 0, and always for Intra16x16 / I_PCM)
 * holds zeros outside the coded block pattern, so every block of it is transformed. */
static int reconstruct_intra(const mmr_recon_h264_picture *p, int mb) {
    const mmr_recon_h264_macroblock *m = &p->macroblocks[mb];
    const mmr_recon_h264_intra *in = &p->intra[mb];
    const int wm = (int)p->width_mbs, W = wm * 16, CW = W / 2;
    const int mbx = mb % wm, mby = mb / wm, qp = m->qp;
    const int left_mb = mbx > 0 && p->macroblocks[mb - 1].slice_id == m->slice_id;
    const int top_mb = mby > 0 && p->macroblocks[mb - wm].slice_id == m->slice_id;
    const int has_record = m->cbp != 0 || in->type == MMR_H264_INTRA_16X16 || in->type == MMR_H264_INTRA_PCM;
    const int16_t *lv = has_record ? p->levels + (size_t)m->coef_offset * MMR_H264_LEVELS_PER_MB : NULL;
    int32_t levels[64], res[64];
    uint8_t pred[256];
    if (in->type == MMR_H264_INTRA_PCM) {
        for (int y = 0; y < 16; ++y)
            for (int x = 0; x < 16; ++x) p->planes[0][(size_t)(mby * 16 + y) * W + mbx * 16 + x] = (uint8_t)lv[y * 16 + x];
        for (int c = 0; c < 2; ++c)
            for (int y = 0; y < 8; ++y)
                for (int x = 0; x < 8; ++x)
                    p->planes[1 + c][(size_t)(mby * 8 + y) * CW + mbx * 8 + x] = (uint8_t)lv[256 + c * 64 + y * 8 + x];
        return 0;
    }
    if (in->type == MMR_H264_INTRA_4X4) {
        for (int k = 0; k < 16; ++k) {
            int bx, by;
            luma_block_position(k, &bx, &by);
            const int ox = mbx * 16 + bx * 4, oy = mby * 16 + by * 4;
            if (oracle_h264_intra4_predict(p->planes[0], W, ox, oy, k, in->modes[k], top_mb, left_mb, pred)) return -1;
            place(p->planes[0], W, ox, oy, 4, pred);
            if (lv) {
                for (int i = 0; i < 16; ++i) levels[i] = lv[k * 16 + i];
                if (p->transform_bypass && qp == 0) oracle_h264_bypass_residual_4x4(levels, in->modes[k], res);
                else if (oracle_h264_transform_residual_4x4(levels, qp, 0, scaling4(p, 0), res)) return -2;
                add_residual(p->planes[0], W, ox, oy, 4, res);
            }
        }
    } else if (in->type == MMR_H264_INTRA_8X8) {
        if (!(m->flags & MMR_H264_MB_TRANSFORM_8X8)) return -1;
        for (int g = 0; g < 4; ++g) {
            const int ox = mbx * 16 + (g % 2) * 8, oy = mby * 16 + (g / 2) * 8;
            if (oracle_h264_intra8_predict(p->planes[0], W, ox, oy, g, in->modes[g], top_mb, left_mb, pred)) return -1;
            place(p->planes[0], W, ox, oy, 8, pred);
            if (lv) {
                for (int i = 0; i < 64; ++i) levels[i] = lv[g * 64 + i];
                if (oracle_h264_transform_residual_8x8(levels, qp, scaling8(p, 0), res)) return -2;
                add_residual(p->planes[0], W, ox, oy, 8, res);
            }
        }
    } else if (in->type == MMR_H264_INTRA_16X16) {
        if (oracle_h264_intra_block_predict(p->planes[0], W, mbx * 16, mby * 16, 16, in->luma_mode, top_mb, left_mb, pred))
            return -1;
        place(p->planes[0], W, mbx * 16, mby * 16, 16, pred);
        int32_t dcl[16], dc[16];
        for (int k = 0; k < 16; ++k) dcl[k] = lv[16 * k];
        if (oracle_h264_intra16_dc(dcl, qp, scaling4(p, 0), dc)) return -2;
        for (int by = 0; by < 4; ++by)
            for (int bx = 0; bx < 4; ++bx) {
                const int k = luma_block_index(bx, by);
                levels[0] = dc[by * 4 + bx];
                for (int i = 1; i < 16; ++i) levels[i] = lv[k * 16 + i];
                if (oracle_h264_transform_residual_4x4(levels, qp, 1, scaling4(p, 0), res)) return -2;
                add_residual(p->planes[0], W, mbx * 16 + bx * 4, mby * 16 + by * 4, 4, res);
            }
    } else {
        return -1;
    }
    for (int c = 0; c < 2; ++c) {
        if (oracle_h264_chroma_predict(p->planes[1 + c], CW, mbx * 8, mby * 8, in->chroma_mode, top_mb, left_mb, pred))
            return -1;
        place(p->planes[1 + c], CW, mbx * 8, mby * 8, 8, pred);
    }
    if (lv) {
        const int cqp = oracle_h264_chroma_qp(qp, p->chroma_qp_offset[0]);
        if (p->transform_bypass && cqp == 0) {
            const int mode = in->chroma_mode == 1 ? 1 : (in->chroma_mode == 2 ? 0 : 2);
            for (int comp = 0; comp < 2; ++comp) {
                int32_t br[4][16];
                bypass_chroma(lv, comp, mode, br);
                for (int b = 0; b < 4; ++b)
                    add_residual(p->planes[1 + comp], CW, mbx * 8 + (b % 2) * 4, mby * 8 + (b / 2) * 4, 4, br[b]);
            }
            return 0;
        }
        for (int comp = 0; comp < 2; ++comp) {
            const uint8_t *sl = scaling4(p, comp + 1);
            int32_t dc_levels[4], dc[4];
            for (int i = 0; i < 4; ++i) dc_levels[i] = lv[256 + comp * 4 + i];
            if (oracle_h264_transform_chroma_dc(dc_levels, cqp, sl, dc)) return -2;
            for (int b = 0; b < 4; ++b) {
                int32_t co[16];
                co[0] = dc[b];
                for (int i = 1; i < 16; ++i) co[i] = lv[264 + (comp * 4 + b) * 15 + (i - 1)];
                if (oracle_h264_transform_residual_4x4(co, cqp, 1, sl, res)) return -2;
                add_residual(p->planes[1 + comp], CW, mbx * 8 + (b % 2) * 4, mby * 8 + (b / 2) * 4, 4, res);
            }
        }
    }
    return 0;
}

/* Reconstruction of one frame picture in macroblock order (inter macroblocks read only the
 * reference pictures, intra ones the reconstructed neighbours) followed by the loop filter when
 * enabled (decode_*_picture -> deblock, decoder.rs:1158, :1930). Returns -3 for an intra
 * macroblock without intra records. */
int oracle_h264_reconstruct(const mmr_recon_h264_picture *p) {
    int mbs = (int)(p->width_mbs * p->height_mbs);
    for (int mb = 0; mb < mbs; ++mb) {
        if (p->macroblocks[mb].flags & MMR_H264_MB_INTRA) {
            if (!p->intra) return -3;
            if (reconstruct_intra(p, mb)) return -4;
            continue;
        }
        for (int k = 0; k < 16; ++k)
            if (predict_block(p, mb, k)) return -1;
        if (add_residuals(p, mb)) return -2;
    }
    if (p->deblock) return oracle_h264_deblock(p);
    return 0;
}
