/* Oracle: DV25 reconstruction restated from crates/codecs/dv/src. Test infrastructure only. */
#include <math.h>
#include <string.h>

#include "oracle.h"
#include "tables.h"

/* ZIGZAG_248, QUANT_SHIFTS, QUANT_OFFSET, IWEIGHT_88, IWEIGHT_248:
 * crates/codecs/dv/src/tables.rs:6-50 (normative DV tables). */
static const int DV_ZIGZAG_248[64] = {
    0, 8, 1, 9, 16, 24, 2, 10, 17, 25, 32, 40, 48, 56, 33, 41, 18, 26, 3, 11, 4, 12, 19, 27, 34,
    42, 49, 57, 50, 58, 35, 43, 20, 28, 5, 13, 6, 14, 21, 29, 36, 44, 51, 59, 52, 60, 37, 45, 22,
    30, 7, 15, 23, 31, 38, 46, 53, 61, 54, 62, 39, 47, 55, 63};
static const uint8_t DV_QUANT_SHIFTS[22][4] = {
    {3, 3, 4, 4}, {3, 3, 4, 4}, {2, 3, 3, 4}, {2, 3, 3, 4}, {2, 2, 3, 3}, {2, 2, 3, 3},
    {1, 2, 2, 3}, {1, 2, 2, 3}, {1, 1, 2, 2}, {1, 1, 2, 2}, {0, 1, 1, 2}, {0, 1, 1, 2},
    {0, 0, 1, 1}, {0, 0, 1, 1}, {0, 0, 0, 1}, {0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0},
    {0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}};
static const int DV_QUANT_OFFSET[4] = {6, 3, 0, 1};
static const int32_t DV_IWEIGHT_88[64] = {
    32768, 16705, 16705, 17734, 17032, 17734, 18205, 18081, 18081, 18205, 18725, 18562, 19195,
    18562, 18725, 19266, 19091, 19705, 19705, 19091, 19266, 21407, 19643, 20267, 20228, 20267,
    19643, 21407, 22725, 21826, 20853, 20806, 20806, 20853, 21826, 22725, 23170, 23170, 21407,
    21400, 21407, 23170, 23170, 24598, 23786, 22018, 22018, 23786, 24598, 25251, 24465, 22654,
    24465, 25251, 25972, 25172, 25172, 25972, 26722, 27969, 26722, 29692, 29692, 31521};
static const int32_t DV_IWEIGHT_248[64] = {
    32768, 16384, 16705, 16705, 17734, 17734, 17734, 17734, 18081, 18081, 18725, 18725, 21407,
    21407, 19091, 19091, 19195, 19195, 18205, 18205, 18725, 18725, 19705, 19705, 20267, 20267,
    21826, 21826, 23170, 23170, 20806, 20806, 20267, 20267, 19266, 19266, 21407, 21407, 20853,
    20853, 21400, 21400, 23786, 23786, 24465, 24465, 22018, 22018, 23170, 23170, 22725, 22725,
    24598, 24598, 24465, 24465, 25172, 25172, 27969, 27969, 25972, 25972, 29692, 29692};

static double COS16[64], COS8[32];
static int cos_ready;

/* The cosine arguments of inverse_88 / inverse_248 (crates/codecs/dv/src/decoder.rs:535-536,
 * :576, :592), evaluated left to right in f64 and passed to the C library cos, the same libm
 * routine Rust's f64::cos calls on linux-gnu. */
static void init_cos(void) {
    if (cos_ready) return;
    for (int x = 0; x < 8; ++x)
        for (int u = 0; u < 8; ++u) COS16[x * 8 + u] = cos((double)(2 * x + 1) * (double)u * M_PI / 16.0);
    for (int y = 0; y < 8; ++y)
        for (int v = 0; v < 4; ++v) COS8[y * 4 + v] = cos((double)(2 * y + 1) * (double)v * M_PI / 8.0);
    cos_ready = 1;
}
const double *oracle_dv_cos16_table(void) { init_cos(); return COS16; }
const double *oracle_dv_cos8_table(void) { init_cos(); return COS8; }

static uint8_t round_clamp_u8(double v) {
    double r = round(v);
    if (r < 0.0) r = 0.0;
    if (r > 255.0) r = 255.0;
    return (uint8_t)r;
}

/* inverse_88, crates/codecs/dv/src/decoder.rs:520-548: direct 2-D sum, v outer, u inner,
 * term = cu * cv * c * cos_x * cos_y evaluated left to right, then (sum / 4).round().clamp. */
void oracle_dv_inverse_88(const int32_t c[64], uint8_t out[64]) {
    init_cos();
    for (int y = 0; y < 8; ++y)
        for (int x = 0; x < 8; ++x) {
            double sum = 0.0;
            for (int v = 0; v < 8; ++v)
                for (int u = 0; u < 8; ++u) {
                    double cu = u == 0 ? M_SQRT1_2 : 1.0;
                    double cv = v == 0 ? M_SQRT1_2 : 1.0;
                    double term = cu * cv;
                    term = term * (double)c[v * 8 + u];
                    term = term * COS16[x * 8 + u];
                    term = term * COS16[y * 8 + v];
                    sum = sum + term;
                }
            out[y * 8 + x] = round_clamp_u8(sum / 4.0);
        }
}

/* inverse_248, decoder.rs:555-602: row-pair butterfly, horizontal 8-point pass (/2),
 * per-field vertical 4-point pass (/2), round, clamp. */
void oracle_dv_inverse_248(const int32_t c[64], uint8_t out[64]) {
    init_cos();
    double butterfly[64], horizontal[64];
    for (int pair = 0; pair < 4; ++pair)
        for (int u = 0; u < 8; ++u) {
            double first = (double)c[pair * 16 + u], second = (double)c[pair * 16 + 8 + u];
            butterfly[pair * 16 + u] = first + second;
            butterfly[pair * 16 + 8 + u] = first - second;
        }
    for (int row = 0; row < 8; ++row)
        for (int x = 0; x < 8; ++x) {
            double sum = 0.0;
            for (int u = 0; u < 8; ++u) {
                double cu = u == 0 ? M_SQRT1_2 : 1.0;
                double term = cu * butterfly[row * 8 + u];
                term = term * COS16[x * 8 + u];
                sum = sum + term;
            }
            horizontal[row * 8 + x] = sum / 2.0;
        }
    for (int field = 0; field < 2; ++field)
        for (int y = 0; y < 4; ++y)
            for (int x = 0; x < 8; ++x) {
                double sum = 0.0;
                for (int v = 0; v < 4; ++v) {
                    double cv = v == 0 ? M_SQRT1_2 : 1.0;
                    double term = cv * horizontal[(v * 2 + field) * 8 + x];
                    term = term * COS8[y * 4 + v];
                    sum = sum + term;
                }
                out[(y * 2 + field) * 8 + x] = round_clamp_u8(sum / 2.0);
            }
}

/* inverse_factor_for_encoding, decoder.rs:290-310. */
int32_t oracle_dv_inverse_factor(int quantization, int class_number, int dct_mode, int position) {
    int q = quantization + DV_QUANT_OFFSET[class_number];
    int area = position <= 5 ? 0 : (position <= 20 ? 1 : (position <= 42 ? 2 : 3));
    int32_t weight = dct_mode ? DV_IWEIGHT_248[position] : DV_IWEIGHT_88[position];
    int class_shift = class_number == 3;
    return weight << (DV_QUANT_SHIFTS[q][area] + 1 + class_shift);
}

/* DC (decode_macroblock, decoder.rs:214: dc * 4 + 1024) and AC dequantisation
 * (consume_ac, decoder.rs:262-271: (level * factor + 8192) >> 14 at scan[position]). */
void oracle_dv_dequantize(const int16_t levels[64], uint8_t info, int32_t c[64]) {
    int dct_mode = info & 1, class_number = (info >> 1) & 3, qno = (info >> 3) & 15;
    const int *scan = dct_mode ? DV_ZIGZAG_248 : ORACLE_ZIGZAG;
    memset(c, 0, 64 * sizeof(int32_t));
    c[0] = (int32_t)levels[0] * 4 + 1024;
    for (int p = 1; p < 64; ++p) {
        if (levels[p] == 0) continue;
        int32_t factor = oracle_dv_inverse_factor(qno, class_number, dct_mode, p);
        c[scan[p]] = ((int32_t)levels[p] * factor + 8192) >> 14;
    }
}

static int dif_sequences(uint32_t system) { return system == MMR_DV_SYSTEM_625_50 ? 12 : 10; }
uint32_t oracle_dv_block_count(uint32_t system) { return (uint32_t)dif_sequences(system) * 27u * 5u * 6u; }

/* macroblock_coordinates, decoder.rs:395-422 (in units of 8 pixels). */
void oracle_dv_macroblock_coordinates(uint32_t system, int sequence, int slot, int out[10]) {
    static const int OFF[5] = {2, 6, 8, 0, 4};
    static const int SHUF3[5] = {18, 9, 27, 0, 36};
    static const int L_START_SHUFFLED[5] = {9, 4, 13, 0, 18};
    static const int SERPENT1[27] = {0, 1, 2, 2, 1, 0, 0, 1, 2, 2, 1, 0, 0, 1, 2, 2, 1, 0, 0, 1, 2, 2, 1, 0, 0, 1, 2};
    static const int SERPENT2[30] = {0, 1, 2, 3, 4, 5, 5, 4, 3, 2, 1, 0, 0, 1, 2, 3, 4, 5, 5, 4, 3, 2, 1, 0, 0, 1, 2, 3, 4, 5};
    int seqs = dif_sequences(system);
    for (int mb = 0; mb < 5; ++mb) {
        if (system == MMR_DV_SYSTEM_625_50) {
            int x = SHUF3[mb] + slot / 3;
            int y = SERPENT1[slot] + ((sequence + OFF[mb]) % seqs) * 3;
            out[mb * 2] = x * 2;
            out[mb * 2 + 1] = y * 2;
        } else {
            int row = (sequence + OFF[mb]) % seqs;
            int k = slot + ((mb == 1 || mb == 2) ? 3 : 0);
            int x = L_START_SHUFFLED[mb] + k / 6;
            int y = SERPENT2[k] + row * 6;
            if (x > 21) y = y * 2 - row * 6;
            out[mb * 2] = x * 4;
            out[mb * 2 + 1] = y;
        }
    }
}

static void put8(uint8_t *plane, int stride, int x, int y, const uint8_t px[64]) {
    for (int r = 0; r < 8; ++r) memcpy(plane + (size_t)(y + r) * stride + x, px + r * 8, 8);
}

/* place_macroblock / put_block / put_split_chroma, decoder.rs:432-505. Block 4 goes to
 * planes[2] (Cr) and block 5 to planes[1] (Cb). */
static void place_macroblock(const mmr_recon_dv_frame *f, uint8_t px[6][64], int mb_x, int mb_y) {
    int x = mb_x * 8, y = mb_y * 8;
    int width = 720, cw = f->system == MMR_DV_SYSTEM_625_50 ? 360 : 180;
    if (f->system == MMR_DV_SYSTEM_625_50) {
        put8(f->planes[0], width, x, y, px[0]);
        put8(f->planes[0], width, x + 8, y, px[1]);
        put8(f->planes[0], width, x, y + 8, px[2]);
        put8(f->planes[0], width, x + 8, y + 8, px[3]);
        put8(f->planes[2], cw, x / 2, y / 2, px[4]);
        put8(f->planes[1], cw, x / 2, y / 2, px[5]);
    } else if (mb_x < 88) {
        for (int i = 0; i < 4; ++i) put8(f->planes[0], width, x + i * 8, y, px[i]);
        put8(f->planes[2], cw, x / 4, y, px[4]);
        put8(f->planes[1], cw, x / 4, y, px[5]);
    } else {
        put8(f->planes[0], width, x, y, px[0]);
        put8(f->planes[0], width, x + 8, y, px[1]);
        put8(f->planes[0], width, x, y + 8, px[2]);
        put8(f->planes[0], width, x + 8, y + 8, px[3]);
        for (int k = 0; k < 2; ++k) {
            uint8_t *plane = f->planes[k == 0 ? 2 : 1];
            const uint8_t *p = px[4 + k];
            for (int r = 0; r < 8; ++r) {
                memcpy(plane + (size_t)(y + r) * cw + x / 4, p + r * 8, 4);
                memcpy(plane + (size_t)(y + 8 + r) * cw + x / 4, p + r * 8 + 4, 4);
            }
        }
    }
}

/* decode_video_with_options (decoder.rs:87-131) + decode_segment (:152-191) with the VLC and
 * spill stages replaced by the level buffer: planes start at Y 16 / C 128, then every
 * segment's five macroblocks are transformed and placed. */
int oracle_dv_reconstruct(const mmr_recon_dv_frame *f) {
    int seqs = dif_sequences(f->system);
    int height = f->system == MMR_DV_SYSTEM_625_50 ? 576 : 480;
    int cw = f->system == MMR_DV_SYSTEM_625_50 ? 360 : 180, ch = f->system == MMR_DV_SYSTEM_625_50 ? 288 : 480;
    memset(f->planes[0], 16, (size_t)720 * height);
    memset(f->planes[1], 128, (size_t)cw * ch);
    memset(f->planes[2], 128, (size_t)cw * ch);
    int32_t c[64];
    uint8_t px[6][64];
    size_t block = 0;
    for (int seq = 0; seq < seqs; ++seq)
        for (int slot = 0; slot < 27; ++slot) {
            int xy[10];
            oracle_dv_macroblock_coordinates(f->system, seq, slot, xy);
            for (int mb = 0; mb < 5; ++mb) {
                for (int b = 0; b < 6; ++b, ++block) {
                    uint8_t info = f->block_info[block];
                    oracle_dv_dequantize(f->levels + block * 64, info, c);
                    if (info & 1) oracle_dv_inverse_248(c, px[b]);
                    else oracle_dv_inverse_88(c, px[b]);
                }
                place_macroblock(f, px, xy[mb * 2], xy[mb * 2 + 1]);
            }
        }
    return 0;
}
