/* Oracle: H.264 intra prediction restated from crates/codecs/h264/src/decoder.rs -- TEST
 * INFRASTRUCTURE (the product never links it).
 *
 *   predict_intra4_block + vertical-right / horizontal-down   :13311-13532
 *   predict_intra8_block + filter_intra8_neighbors + VR / HD  :13030-13309
 *   predict_block / dc_prediction / plane_prediction           :14583-14698
 *   predict_chroma_block / predict_chroma_dc / dc_value        :14499-14581
 *   average / filter_1_2_1 / average_1_3                       :13553-13564
 *   transform_intra16_luma_dc                                  :14142-14183
 *
 * Availability: the left macroblock when it is not in the first column and its address is at
 * or after the slice start (:10801-10814), the top one likewise; the corner (top-left) sample is
 * read whenever both are available, and top-right samples follow the per-block rules below. */
#include <stdint.h>
#include <string.h>

#include "oracle.h"

static uint8_t avg2(uint8_t a, uint8_t b) { return (uint8_t)((a + b + 1) >> 1); }
static uint8_t f121(uint8_t a, uint8_t b, uint8_t c) { return (uint8_t)((a + 2 * b + c + 2) >> 2); }
static uint8_t avg13(uint8_t a, uint8_t b) { return (uint8_t)((a + 3 * b + 2) >> 2); }

/* top_or_corner / left_or_corner: index -1 is the corner. */
static uint8_t tc(const uint8_t *top, uint8_t corner, int i) { return i < 0 ? corner : top[i]; }

/* ---------------------------------------------------------------- Intra4x4 */

int oracle_h264_intra4_predict(const uint8_t *plane, int stride, int ox, int oy, int block_index, int mode,
                               int top_mb, int left_mb, uint8_t out[16]) {
    const int bx = ox % 16, by = oy % 16;
    const int has_top = by > 0 || top_mb, has_left = bx > 0 || left_mb;
    uint8_t top[8], left[4], corner = 0;
    if (has_top) {
        for (int x = 0; x < 4; ++x) top[x] = plane[(size_t)(oy - 1) * stride + ox + x];
        for (int x = 4; x < 8; ++x) {
            /* top-right: not for blocks 3 and 11, not into the next macroblock below the first
             * block row, not past the picture's right edge */
            const int future = (oy % 16) != 0 && (ox % 16) + x >= 16;
            top[x] = (block_index == 3 || block_index == 11 || future || ox + x >= stride)
                         ? top[3]
                         : plane[(size_t)(oy - 1) * stride + ox + x];
        }
    }
    if (has_left)
        for (int y = 0; y < 4; ++y) left[y] = plane[(size_t)(oy + y) * stride + ox - 1];
    const int has_corner = has_top && has_left;
    if (has_corner) corner = plane[(size_t)(oy - 1) * stride + ox - 1];
    switch (mode) {
    case 0:   /* vertical */
        if (!has_top) return -1;
        for (int y = 0; y < 4; ++y) memcpy(out + 4 * y, top, 4);
        return 0;
    case 1:   /* horizontal */
        if (!has_left) return -1;
        for (int y = 0; y < 4; ++y) memset(out + 4 * y, left[y], 4);
        return 0;
    case 2: { /* DC */
        uint32_t sum = 0, v;
        if (has_top && has_left) {
            for (int i = 0; i < 4; ++i) sum += top[i] + left[i];
            v = (sum + 4) / 8;
        } else if (has_top || has_left) {
            const uint8_t *s = has_top ? top : left;
            for (int i = 0; i < 4; ++i) sum += s[i];
            v = (sum + 2) / 4;
        } else {
            v = 128;
        }
        memset(out, (int)v, 16);
        return 0;
    }
    case 3:   /* diagonal down left */
        if (!has_top) return -1;
        for (int y = 0; y < 4; ++y)
            for (int x = 0; x < 4; ++x)
                out[y * 4 + x] = (x == 3 && y == 3) ? avg13(top[6], top[7]) : f121(top[x + y], top[x + y + 1], top[x + y + 2]);
        return 0;
    case 4:   /* diagonal down right */
        if (!has_top || !has_left || !has_corner) return -1;
        for (int y = 0; y < 4; ++y)
            for (int x = 0; x < 4; ++x) {
                uint8_t v;
                if (x > y) v = f121(tc(top, corner, x - y - 2), tc(top, corner, x - y - 1), tc(top, corner, x - y));
                else if (x < y) v = f121(tc(left, corner, y - x - 2), tc(left, corner, y - x - 1), tc(left, corner, y - x));
                else v = f121(top[0], corner, left[0]);
                out[y * 4 + x] = v;
            }
        return 0;
    case 5:   /* vertical right */
        if (!has_top || !has_left || !has_corner) return -1;
        for (int y = 0; y < 4; ++y)
            for (int x = 0; x < 4; ++x) {
                const int z = 2 * x - y;
                uint8_t v;
                if (z >= 0 && (z % 2) == 0) v = avg2(tc(top, corner, x - y / 2 - 1), tc(top, corner, x - y / 2));
                else if (z >= 0) v = f121(tc(top, corner, x - y / 2 - 2), tc(top, corner, x - y / 2 - 1), tc(top, corner, x - y / 2));
                else if (z == -1) v = f121(left[0], corner, top[0]);
                else v = f121(tc(left, corner, y - 1), tc(left, corner, y - 2), tc(left, corner, y - 3));
                out[y * 4 + x] = v;
            }
        return 0;
    case 6:   /* horizontal down */
        if (!has_top || !has_left || !has_corner) return -1;
        for (int y = 0; y < 4; ++y)
            for (int x = 0; x < 4; ++x) {
                const int z = 2 * y - x;
                uint8_t v;
                if (z >= 0 && (z % 2) == 0) v = avg2(tc(left, corner, y - x / 2 - 1), tc(left, corner, y - x / 2));
                else if (z >= 0) v = f121(tc(left, corner, y - x / 2 - 2), tc(left, corner, y - x / 2 - 1), tc(left, corner, y - x / 2));
                else if (z == -1) v = f121(top[0], corner, left[0]);
                else v = f121(tc(top, corner, x - 1), tc(top, corner, x - 2), tc(top, corner, x - 3));
                out[y * 4 + x] = v;
            }
        return 0;
    case 7:   /* vertical left */
        if (!has_top) return -1;
        for (int y = 0; y < 4; ++y)
            for (int x = 0; x < 4; ++x) {
                const int s = x + y / 2;
                out[y * 4 + x] = (y % 2) == 0 ? avg2(top[s], top[s + 1]) : f121(top[s], top[s + 1], top[s + 2]);
            }
        return 0;
    case 8:   /* horizontal up */
        if (!has_left) return -1;
        for (int y = 0; y < 4; ++y)
            for (int x = 0; x < 4; ++x) {
                const int z = x + 2 * y, i = y + x / 2;
                uint8_t v;
                if (z == 0 || z == 2 || z == 4) v = avg2(left[i], left[i + 1]);
                else if (z == 1 || z == 3) v = f121(left[i], left[i + 1], left[i + 2]);
                else if (z == 5) v = avg13(left[2], left[3]);
                else v = left[3];
                out[y * 4 + x] = v;
            }
        return 0;
    default:
        return -1;
    }
}

/* ---------------------------------------------------------------- Intra8x8 */

int oracle_h264_intra8_predict(const uint8_t *plane, int stride, int ox, int oy, int group, int mode, int top_mb,
                               int left_mb, uint8_t out[64]) {
    const int has_top = group / 2 > 0 || top_mb, has_left = (group % 2) != 0 || left_mb;
    uint8_t t[16], l[8], c = 0;
    if (has_top) {
        for (int x = 0; x < 8; ++x) t[x] = plane[(size_t)(oy - 1) * stride + ox + x];
        for (int x = 8; x < 16; ++x)   /* top-right: not for the last 8x8 block, not past the right edge */
            t[x] = (group == 3 || ox + x >= stride) ? t[7] : plane[(size_t)(oy - 1) * stride + ox + x];
    }
    if (has_left)
        for (int y = 0; y < 8; ++y) l[y] = plane[(size_t)(oy + y) * stride + ox - 1];
    const int has_corner = has_top && has_left;
    if (has_corner) c = plane[(size_t)(oy - 1) * stride + ox - 1];
    /* filter_intra8_neighbors: the corner from the unfiltered samples, each line [1 2 1] with
     * its ends from the corner (or [1 3] without it) and [1 3] at the far end */
    uint8_t top[16], left[8], corner = c;
    if (has_corner) corner = f121(l[0], c, t[0]);
    if (has_top) {
        top[0] = has_corner ? f121(c, t[0], t[1]) : avg13(t[1], t[0]);
        for (int i = 1; i < 15; ++i) top[i] = f121(t[i - 1], t[i], t[i + 1]);
        top[15] = avg13(t[14], t[15]);
    }
    if (has_left) {
        left[0] = has_corner ? f121(c, l[0], l[1]) : avg13(l[1], l[0]);
        for (int i = 1; i < 7; ++i) left[i] = f121(l[i - 1], l[i], l[i + 1]);
        left[7] = avg13(l[6], l[7]);
    }
    switch (mode) {
    case 0:
        if (!has_top) return -1;
        for (int y = 0; y < 8; ++y) memcpy(out + 8 * y, top, 8);
        return 0;
    case 1:
        if (!has_left) return -1;
        for (int y = 0; y < 8; ++y) memset(out + 8 * y, left[y], 8);
        return 0;
    case 2: {
        uint32_t sum = 0, v;
        if (has_top && has_left) {
            for (int i = 0; i < 8; ++i) sum += top[i] + left[i];
            v = (sum + 8) / 16;
        } else if (has_top || has_left) {
            const uint8_t *s = has_top ? top : left;
            for (int i = 0; i < 8; ++i) sum += s[i];
            v = (sum + 4) / 8;
        } else {
            v = 128;
        }
        memset(out, (int)v, 64);
        return 0;
    }
    case 3:
        if (!has_top) return -1;
        for (int y = 0; y < 8; ++y)
            for (int x = 0; x < 8; ++x)
                out[y * 8 + x] = (x == 7 && y == 7) ? avg13(top[14], top[15]) : f121(top[x + y], top[x + y + 1], top[x + y + 2]);
        return 0;
    case 4:
        if (!has_top || !has_left || !has_corner) return -1;
        for (int y = 0; y < 8; ++y)
            for (int x = 0; x < 8; ++x) {
                uint8_t v;
                if (x > y) v = f121(tc(top, corner, x - y - 2), tc(top, corner, x - y - 1), tc(top, corner, x - y));
                else if (x < y) v = f121(tc(left, corner, y - x - 2), tc(left, corner, y - x - 1), tc(left, corner, y - x));
                else v = f121(top[0], corner, left[0]);
                out[y * 8 + x] = v;
            }
        return 0;
    case 5:
        if (!has_top || !has_left || !has_corner) return -1;
        for (int y = 0; y < 8; ++y)
            for (int x = 0; x < 8; ++x) {
                const int z = 2 * x - y;
                uint8_t v;
                if (z >= 0 && (z % 2) == 0) v = avg2(tc(top, corner, x - y / 2 - 1), tc(top, corner, x - y / 2));
                else if (z >= 0) v = f121(tc(top, corner, x - y / 2 - 2), tc(top, corner, x - y / 2 - 1), tc(top, corner, x - y / 2));
                else if (z == -1) v = f121(left[0], corner, top[0]);
                else v = f121(tc(left, corner, y - 2 * x - 1), tc(left, corner, y - 2 * x - 2), tc(left, corner, y - 2 * x - 3));
                out[y * 8 + x] = v;
            }
        return 0;
    case 6:
        if (!has_top || !has_left || !has_corner) return -1;
        for (int y = 0; y < 8; ++y)
            for (int x = 0; x < 8; ++x) {
                const int z = 2 * y - x;
                uint8_t v;
                if (z >= 0 && (z % 2) == 0) v = avg2(tc(left, corner, y - x / 2 - 1), tc(left, corner, y - x / 2));
                else if (z >= 0) v = f121(tc(left, corner, y - x / 2 - 2), tc(left, corner, y - x / 2 - 1), tc(left, corner, y - x / 2));
                else if (z == -1) v = f121(top[0], corner, left[0]);
                else v = f121(tc(top, corner, x - 2 * y - 1), tc(top, corner, x - 2 * y - 2), tc(top, corner, x - 2 * y - 3));
                out[y * 8 + x] = v;
            }
        return 0;
    case 7:
        if (!has_top) return -1;
        for (int y = 0; y < 8; ++y)
            for (int x = 0; x < 8; ++x) {
                const int s = x + y / 2;
                out[y * 8 + x] = (y % 2) == 0 ? avg2(top[s], top[s + 1]) : f121(top[s], top[s + 1], top[s + 2]);
            }
        return 0;
    case 8:
        if (!has_left) return -1;
        for (int y = 0; y < 8; ++y)
            for (int x = 0; x < 8; ++x) {
                const int z = x + 2 * y;
                uint8_t v;
                if (z < 13 && (z % 2) == 0) v = avg2(left[z / 2], left[z / 2 + 1]);
                else if (z < 13) v = f121(left[z / 2], left[z / 2 + 1], left[z / 2 + 2]);
                else if (z == 13) v = f121(left[6], left[7], left[7]);
                else v = left[7];
                out[y * 8 + x] = v;
            }
        return 0;
    default:
        return -1;
    }
}

/* ---------------------------------------------------------------- 16x16 / chroma 8x8 */

/* predict_block: size 16 (Intra16x16 modes 0 vertical, 1 horizontal, 2 DC, 3 plane) or 8. */
int oracle_h264_intra_block_predict(const uint8_t *plane, int stride, int ox, int oy, int size, int mode, int has_top,
                                    int has_left, uint8_t *out) {
    uint8_t top[16], left[16];
    if (has_top)
        for (int x = 0; x < size; ++x) top[x] = plane[(size_t)(oy - 1) * stride + ox + x];
    if (has_left)
        for (int y = 0; y < size; ++y) left[y] = plane[(size_t)(oy + y) * stride + ox - 1];
    switch (mode) {
    case 0:
        if (!has_top) return -1;
        for (int y = 0; y < size; ++y) memcpy(out + size * y, top, (size_t)size);
        return 0;
    case 1:
        if (!has_left) return -1;
        for (int y = 0; y < size; ++y) memset(out + size * y, left[y], (size_t)size);
        return 0;
    case 2: {   /* dc_prediction */
        uint32_t sum = 0, v;
        if (has_top && has_left) {
            for (int i = 0; i < size; ++i) sum += top[i] + left[i];
            v = (sum + (uint32_t)size) / (uint32_t)(2 * size);
        } else if (has_top || has_left) {
            const uint8_t *s = has_top ? top : left;
            for (int i = 0; i < size; ++i) sum += s[i];
            v = (sum + (uint32_t)size / 2) / (uint32_t)size;
        } else {
            v = 128;
        }
        memset(out, (int)v, (size_t)(size * size));
        return 0;
    }
    case 3: {   /* plane_prediction */
        if (!has_top || !has_left) return -1;
        const int corner = plane[(size_t)(oy - 1) * stride + ox - 1], half = size / 2;
        int32_t h = 0, v = 0;
        for (int i = 1; i <= half; ++i) {
            const int et = i == half ? corner : top[half - 1 - i];
            const int el = i == half ? corner : left[half - 1 - i];
            h += i * (top[half - 1 + i] - et);
            v += i * (left[half - 1 + i] - el);
        }
        const int scale = size == 16 ? 5 : 17, shift = size == 16 ? 6 : 5;
        const int32_t a = 16 * (top[size - 1] + left[size - 1]);
        const int32_t b = (scale * h + (1 << (shift - 1))) >> shift;
        const int32_t c = (scale * v + (1 << (shift - 1))) >> shift;
        const int center = half - 1;
        for (int y = 0; y < size; ++y)
            for (int x = 0; x < size; ++x) {
                int32_t p = (a + b * (x - center) + c * (y - center) + 16) >> 5;
                out[y * size + x] = (uint8_t)(p < 0 ? 0 : (p > 255 ? 255 : p));
            }
        return 0;
    }
    default:
        return -1;
    }
}

/* predict_chroma_block: mode 0 DC (per 4x4 quadrant), 1 horizontal, 2 vertical, 3 plane. */
int oracle_h264_chroma_predict(const uint8_t *plane, int stride, int ox, int oy, int mode, int has_top, int has_left,
                               uint8_t out[64]) {
    if (mode != 0) return oracle_h264_intra_block_predict(plane, stride, ox, oy, 8, mode == 2 ? 0 : mode, has_top, has_left, out);
    uint8_t top[8], left[8];
    if (has_top)
        for (int x = 0; x < 8; ++x) top[x] = plane[(size_t)(oy - 1) * stride + ox + x];
    if (has_left)
        for (int y = 0; y < 8; ++y) left[y] = plane[(size_t)(oy + y) * stride + ox - 1];
    for (int qy = 0; qy < 2; ++qy)
        for (int qx = 0; qx < 2; ++qx) {
            /* quadrants (0,0) and (1,1) use both lines; (1,0) the top, (0,1) the left, each
             * falling back to the other line when theirs is missing */
            const int use_top = has_top && (!has_left || qy == 0 || qx == 1);
            const int use_left = has_left && (!has_top || qx == 0 || qy == 1);
            uint32_t sum = 0, count = 0;
            if (use_top)
                for (int i = 0; i < 4; ++i, ++count) sum += top[qx * 4 + i];
            if (use_left)
                for (int i = 0; i < 4; ++i, ++count) sum += left[qy * 4 + i];
            const uint8_t v = count ? (uint8_t)((sum + count / 2) / count) : 128;
            for (int y = qy * 4; y < qy * 4 + 4; ++y) memset(out + y * 8 + qx * 4, v, 4);
        }
    return 0;
}

/* transform_intra16_luma_dc: levels by scan index (zig-zag of the 4x4 DC array); out[row * 4 +
 * column] is the scaled DC of the 4x4 block at (column, row). */
int oracle_h264_intra16_dc(const int32_t levels[16], int qp, const uint8_t scaling[16], int32_t out[16]) {
    static const uint8_t ZZ[16][2] = {{0, 0}, {0, 1}, {1, 0}, {2, 0}, {1, 1}, {0, 2}, {0, 3}, {1, 2},
                                      {2, 1}, {3, 0}, {3, 1}, {2, 2}, {1, 3}, {2, 3}, {3, 2}, {3, 3}};
    static const int H[4][4] = {{1, 1, 1, 1}, {1, 1, -1, -1}, {1, -1, -1, 1}, {1, -1, 1, -1}};
    static const int NORM[6] = {10, 11, 13, 14, 16, 18};
    if (qp < 0 || qp > 51) return -1;
    int64_t c[16] = {0}, h[16];
    for (int i = 0; i < 16; ++i) c[ZZ[i][0] * 4 + ZZ[i][1]] = levels[i];
    for (int r = 0; r < 4; ++r)
        for (int k = 0; k < 4; ++k) {
            int64_t s = 0;
            for (int i = 0; i < 4; ++i) s += (int64_t)H[r][i] * c[i * 4 + k];
            h[r * 4 + k] = s;
        }
    const int64_t scale = (int64_t)NORM[qp % 6] * scaling[0];   /* level_scale_4x4(qp, 0, 0) */
    for (int r = 0; r < 4; ++r)
        for (int k = 0; k < 4; ++k) {
            int64_t t = 0;
            for (int i = 0; i < 4; ++i) t += h[r * 4 + i] * H[i][k];
            const int64_t s = qp >= 36 ? (t * scale) << (qp / 6 - 6) : (t * scale + ((int64_t)1 << (5 - qp / 6))) >> (6 - qp / 6);
            if (s < INT32_MIN || s > INT32_MAX) return -2;
            out[r * 4 + k] = (int32_t)s;
        }
    return 0;
}
