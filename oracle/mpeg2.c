/* Oracle: MPEG-2 reconstruction restated from crates/codecs/mpeg2/src. Test infrastructure only. */
#include <math.h>
#include <string.h>

#include "oracle.h"
#include "tables.h"

/* inverse_dct, crates/codecs/mpeg2/src/transform.rs:90-110: same passes as the MJPEG IDCT,
 * output sum.round().clamp(-256, 255) with no level shift. */
void oracle_mpeg2_inverse_dct(const int32_t coefficients[64], int16_t out[64]) {
    double temporary[64];
    for (int i = 0; i < 64; ++i) temporary[i] = 0.0;
    for (int v = 0; v < 8; ++v)
        for (int x = 0; x < 8; ++x)
            for (int u = 0; u < 8; ++u) {
                double product = ORACLE_IDCT_MATRIX[x][u] * (double)coefficients[v * 8 + u];
                temporary[v * 8 + x] = temporary[v * 8 + x] + product;
            }
    for (int y = 0; y < 8; ++y)
        for (int x = 0; x < 8; ++x) {
            double sum = 0.0;
            for (int v = 0; v < 8; ++v) {
                double product = ORACLE_IDCT_MATRIX[y][v] * temporary[v * 8 + x];
                sum = sum + product;
            }
            double r = round(sum);
            if (r < -256.0) r = -256.0;
            if (r > 255.0) r = 255.0;
            out[y * 8 + x] = (int16_t)r;
        }
}

static int32_t copysign_i32(int32_t magnitude, int32_t sign) { return sign < 0 ? -magnitude : magnitude; }
static int32_t clamp_i32(int32_t v, int32_t lo, int32_t hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* decode_intra_block dequantisation, crates/codecs/mpeg2/src/decoder.rs:734-792.
 * levels[p] is the level at scan position p; levels[0] is the DC predictor value. */
void oracle_mpeg2_dequantize_intra(const int16_t levels[64], int component, int qscale,
                                   int intra_dc_precision, int alternate_scan,
                                   const uint8_t matrix[64], int32_t c[64]) {
    (void)component;
    const int *scan = alternate_scan ? ORACLE_ALTERNATE_SCAN : ORACLE_ZIGZAG;
    memset(c, 0, 64 * sizeof(int32_t));
    c[0] = (int32_t)levels[0] << (3 - intra_dc_precision);
    int32_t mismatch = c[0] ^ 1;
    for (int p = 1; p < 64; ++p) {
        int32_t level = levels[p];
        if (level == 0) continue;
        int position = scan[p];
        int32_t magnitude = (level < 0 ? -level : level) * qscale * (int32_t)matrix[position] / 16;
        int32_t value = clamp_i32(copysign_i32(magnitude, level), -2048, 2047);
        c[position] = value;
        mismatch ^= value;
    }
    c[63] ^= mismatch & 1;
}

/* decode_non_intra_block dequantisation, decoder.rs:794-846. The first-coefficient
 * special case (:813-823) is the same expression with abs(level) = 1. */
void oracle_mpeg2_dequantize_non_intra(const int16_t levels[64], int qscale, int alternate_scan,
                                       const uint8_t matrix[64], int32_t c[64]) {
    const int *scan = alternate_scan ? ORACLE_ALTERNATE_SCAN : ORACLE_ZIGZAG;
    memset(c, 0, 64 * sizeof(int32_t));
    int32_t mismatch = 1;
    for (int p = 0; p < 64; ++p) {
        int32_t level = levels[p];
        if (level == 0) continue;
        int position = scan[p];
        int32_t magnitude = (2 * (level < 0 ? -level : level) + 1) * qscale * (int32_t)matrix[position] / 32;
        int32_t value = clamp_i32(copysign_i32(magnitude, level), -2048, 2047);
        c[position] = value;
        mismatch ^= value;
    }
    c[63] ^= mismatch & 1;
}

static int div_euclid2(int v) { return v >= 0 ? v / 2 : -((-v + 1) / 2); }

/* clamped_sample, decoder.rs:1208-1222: clamp to the coded plane (height = len / stride). */
static uint16_t clamped_sample(const uint8_t *reference, int stride, int height, int x, int y) {
    if (x < 0) x = 0;
    if (x > stride - 1) x = stride - 1;
    if (y < 0) y = 0;
    if (y > height - 1) y = height - 1;
    return reference[(size_t)y * stride + x];
}

/* halfpel_sample, decoder.rs:1181-1206: integer part div_euclid(2), fraction rem_euclid(2),
 * 2-tap (a+b).div_ceil(2), 4-tap (a+b+c+d+2)/4. */
uint8_t oracle_mpeg2_halfpel_sample(const uint8_t *reference, int stride, int height, int x, int y,
                                    int mv_x, int mv_y) {
    int base_x = x + div_euclid2(mv_x), base_y = y + div_euclid2(mv_y);
    int fx = (mv_x - 2 * div_euclid2(mv_x)) != 0, fy = (mv_y - 2 * div_euclid2(mv_y)) != 0;
    uint16_t a = clamped_sample(reference, stride, height, base_x, base_y);
    if (!fx && !fy) return (uint8_t)a;
    if (fx && !fy) {
        uint16_t b = clamped_sample(reference, stride, height, base_x + 1, base_y);
        return (uint8_t)((a + b + 1) / 2);
    }
    if (!fx && fy) {
        uint16_t c = clamped_sample(reference, stride, height, base_x, base_y + 1);
        return (uint8_t)((a + c + 1) / 2);
    }
    uint16_t b = clamped_sample(reference, stride, height, base_x + 1, base_y);
    uint16_t c = clamped_sample(reference, stride, height, base_x, base_y + 1);
    uint16_t d = clamped_sample(reference, stride, height, base_x + 1, base_y + 1);
    return (uint8_t)((a + b + c + d + 2) / 4);
}

/* predict_plane, decoder.rs:1150-1179: fwd+bwd combine (a+b).div_ceil(2). */
static void predict_plane(uint8_t *dst, int stride, int height, int x, int y, int w, int h,
                          const int *fwd_mv, const int *bwd_mv, const uint8_t *fwd, const uint8_t *bwd) {
    for (int row = 0; row < h; ++row)
        for (int col = 0; col < w; ++col) {
            int has_f = fwd_mv && fwd, has_b = bwd_mv && bwd;
            uint16_t a = 0, b = 0;
            if (has_f) a = oracle_mpeg2_halfpel_sample(fwd, stride, height, x + col, y + row, fwd_mv[0], fwd_mv[1]);
            if (has_b) b = oracle_mpeg2_halfpel_sample(bwd, stride, height, x + col, y + row, bwd_mv[0], bwd_mv[1]);
            uint8_t sample;
            if (has_f && has_b) sample = (uint8_t)((a + b + 1) / 2);
            else if (has_f) sample = (uint8_t)a;
            else if (has_b) sample = (uint8_t)b;
            else sample = 0;
            dst[(size_t)(y + row) * stride + x + col] = sample;
        }
}

/* halfpel_field_sample + clamped_field_sample, decoder.rs:1099-1148: the destination row's field
 * line is dest_y / 2; samples come from the field of parity `select` (rows select, select + 2, ...),
 * clamped to that field's last line (height - 1 - select) / 2 and to the row width `stride`. */
static uint16_t clamped_field_sample(const uint8_t *reference, int stride, int height, int x, int field_y, int select) {
    int max_row = height > 1 + select ? (height - 1 - select) / 2 : 0;
    if (field_y < 0) field_y = 0;
    if (field_y > max_row) field_y = max_row;
    if (x < 0) x = 0;
    if (x > stride - 1) x = stride - 1;
    return reference[(size_t)(field_y * 2 + select) * stride + x];
}

uint8_t oracle_mpeg2_halfpel_field_sample(const uint8_t *reference, int stride, int height, int x, int dest_y,
                                          int mv_x, int mv_y, int select) {
    int base_x = x + div_euclid2(mv_x), base_y = dest_y / 2 + div_euclid2(mv_y);
    int fx = (mv_x - 2 * div_euclid2(mv_x)) != 0, fy = (mv_y - 2 * div_euclid2(mv_y)) != 0;
    uint16_t a = clamped_field_sample(reference, stride, height, base_x, base_y, select);
    if (!fx && !fy) return (uint8_t)a;
    if (fx && !fy) return (uint8_t)((a + clamped_field_sample(reference, stride, height, base_x + 1, base_y, select) + 1) / 2);
    if (!fx && fy) return (uint8_t)((a + clamped_field_sample(reference, stride, height, base_x, base_y + 1, select) + 1) / 2);
    uint16_t b = clamped_field_sample(reference, stride, height, base_x + 1, base_y, select);
    uint16_t c = clamped_field_sample(reference, stride, height, base_x, base_y + 1, select);
    uint16_t d = clamped_field_sample(reference, stride, height, base_x + 1, base_y + 1, select);
    return (uint8_t)((a + b + c + d + 2) / 4);
}

/* predict_field_plane, decoder.rs:1053-1097: destination rows of parity f use the field-f
 * vector and select of each direction; fwd+bwd combine (a+b).div_ceil(2). mv[f] / sel[f]. */
static void predict_field_plane(uint8_t *dst, int stride, int height, int x, int y, int w, int h,
                                const int (*fwd_mv)[2], const int *fwd_sel, const int (*bwd_mv)[2], const int *bwd_sel,
                                const uint8_t *fwd, const uint8_t *bwd) {
    for (int row = 0; row < h; ++row) {
        int f = row & 1, py = y + row;
        for (int col = 0; col < w; ++col) {
            int has_f = fwd_mv && fwd, has_b = bwd_mv && bwd;
            uint16_t a = 0, b = 0;
            if (has_f)
                a = oracle_mpeg2_halfpel_field_sample(fwd, stride, height, x + col, py, fwd_mv[f][0], fwd_mv[f][1], fwd_sel[f]);
            if (has_b)
                b = oracle_mpeg2_halfpel_field_sample(bwd, stride, height, x + col, py, bwd_mv[f][0], bwd_mv[f][1], bwd_sel[f]);
            uint8_t sample;
            if (has_f && has_b) sample = (uint8_t)((a + b + 1) / 2);
            else if (has_f) sample = (uint8_t)a;
            else if (has_b) sample = (uint8_t)b;
            else sample = 0;
            dst[(size_t)py * stride + x + col] = sample;
        }
    }
}

/* put_block, decoder.rs:1231-1298. */
static void put_block(const mmr_recon_mpeg2_picture *p, int stride_y, int mb_x, int mb_y, int block,
                      const int32_t c[64], int intra, int interlaced) {
    int16_t samples[64];
    oracle_mpeg2_inverse_dct(c, samples);
    uint8_t *plane;
    int stride, x, y, step;
    int stride_c = stride_y / 2;
    switch (block) {
    case 0: plane = p->planes[0]; stride = stride_y; x = mb_x * 16; y = mb_y * 16; step = interlaced ? 2 : 1; break;
    case 1: plane = p->planes[0]; stride = stride_y; x = mb_x * 16 + 8; y = mb_y * 16; step = interlaced ? 2 : 1; break;
    case 2: plane = p->planes[0]; stride = stride_y; x = mb_x * 16; y = mb_y * 16 + (interlaced ? 1 : 8); step = interlaced ? 2 : 1; break;
    case 3: plane = p->planes[0]; stride = stride_y; x = mb_x * 16 + 8; y = mb_y * 16 + (interlaced ? 1 : 8); step = interlaced ? 2 : 1; break;
    case 4: plane = p->planes[1]; stride = stride_c; x = mb_x * 8; y = mb_y * 8; step = 1; break;
    default: plane = p->planes[2]; stride = stride_c; x = mb_x * 8; y = mb_y * 8; step = 1; break;
    }
    for (int row = 0; row < 8; ++row)
        for (int col = 0; col < 8; ++col) {
            size_t index = (size_t)(y + row * step) * stride + x + col;
            int32_t residual = samples[row * 8 + col];
            int32_t value = intra ? residual : (int32_t)plane[index] + residual;
            plane[index] = (uint8_t)clamp_i32(value, 0, 255);
        }
}

/* One frame picture: the per-macroblock body of decode_picture_data (decoder.rs:340-605)
 * with VLC decoding replaced by the macroblock records and level buffer. Macroblocks are
 * independent inside a picture, so raster order equals slice order. Prediction follows
 * predict_macroblock (decoder.rs:944-1046): P uses the newer anchor as forward reference,
 * B uses older/newer as forward/backward; chroma MV = v / 2 truncating (chroma_motion), per
 * field for field prediction (chroma_field_motion, decoder.rs:1224-1229, selects unchanged). */
int oracle_mpeg2_reconstruct(const mmr_recon_mpeg2_picture *p) {
    int mb_w = (int)((p->width + 15) / 16), mb_h = (int)((p->height + 15) / 16);
    int stride_y = mb_w * 16, height_y = mb_h * 16;
    int stride_c = stride_y / 2, height_c = height_y / 2;
    int32_t c[64];
    for (int mb = 0; mb < mb_w * mb_h; ++mb) {
        const mmr_recon_mpeg2_macroblock *m = &p->macroblocks[mb];
        int mb_x = mb % mb_w, mb_y = mb / mb_w;
        int interlaced = (m->flags & MMR_MPEG2_MB_INTERLACED_DCT) != 0;
        const int16_t *levels = p->levels + (size_t)m->block_offset * 64;
        if (m->flags & MMR_MPEG2_MB_INTRA) {
            for (int b = 0; b < 6; ++b) {
                int component = b < 4 ? 0 : b - 3;
                const uint8_t *matrix = p->matrices + (component == 0 ? 0 : 2) * 64;
                oracle_mpeg2_dequantize_intra(levels + b * 64, component, m->qscale, (int)p->intra_dc_precision,
                                              (int)p->alternate_scan, matrix, c);
                put_block(p, stride_y, mb_x, mb_y, b, c, 1, interlaced);
            }
            continue;
        }
        int fwd[2] = {m->mv[0][0], m->mv[0][1]}, bwd[2] = {m->mv[1][0], m->mv[1][1]};
        int has_f = (m->flags & MMR_MPEG2_MB_FORWARD) != 0, has_b = (m->flags & MMR_MPEG2_MB_BACKWARD) != 0;
        if (!has_f && !has_b) return -2;
        if (has_f && !p->forward[0]) return -2;
        if (has_b && !p->backward[0]) return -2;
        if (m->flags & MMR_MPEG2_MB_FIELD_MOTION) {
            if (!p->field_motion) return -2;
            const mmr_recon_mpeg2_field_motion *fm = &p->field_motion[mb];
            int lf[2][2] = {{fwd[0], fwd[1]}, {fm->mv[0][0], fm->mv[0][1]}};
            int lb[2][2] = {{bwd[0], bwd[1]}, {fm->mv[1][0], fm->mv[1][1]}};
            int sf[2] = {fm->field_select[0][0], fm->field_select[0][1]};
            int sb[2] = {fm->field_select[1][0], fm->field_select[1][1]};
            int cfv[2][2], cbv[2][2];
            for (int f = 0; f < 2; ++f)
                for (int k = 0; k < 2; ++k) {
                    cfv[f][k] = lf[f][k] / 2;
                    cbv[f][k] = lb[f][k] / 2;
                }
            predict_field_plane(p->planes[0], stride_y, height_y, mb_x * 16, mb_y * 16, 16, 16, has_f ? lf : 0, sf,
                                has_b ? lb : 0, sb, p->forward[0], p->backward[0]);
            for (int k = 1; k < 3; ++k)
                predict_field_plane(p->planes[k], stride_c, height_c, mb_x * 8, mb_y * 8, 8, 8, has_f ? cfv : 0, sf,
                                    has_b ? cbv : 0, sb, p->forward[k], p->backward[k]);
        } else {
            int cf[2] = {fwd[0] / 2, fwd[1] / 2}, cb[2] = {bwd[0] / 2, bwd[1] / 2};
            predict_plane(p->planes[0], stride_y, height_y, mb_x * 16, mb_y * 16, 16, 16, has_f ? fwd : 0,
                          has_b ? bwd : 0, p->forward[0], p->backward[0]);
            for (int k = 1; k < 3; ++k)
                predict_plane(p->planes[k], stride_c, height_c, mb_x * 8, mb_y * 8, 8, 8, has_f ? cf : 0,
                              has_b ? cb : 0, p->forward[k], p->backward[k]);
        }
        int coded = 0;
        for (int b = 0; b < 6; ++b) {
            if (!(m->cbp & (1 << (5 - b)))) continue;
            const uint8_t *matrix = p->matrices + (b < 4 ? 1 : 3) * 64;
            oracle_mpeg2_dequantize_non_intra(levels + coded * 64, m->qscale, (int)p->alternate_scan, matrix, c);
            ++coded;
            put_block(p, stride_y, mb_x, mb_y, b, c, 0, interlaced);
        }
    }
    return 0;
}
