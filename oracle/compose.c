/* Oracle: colour conversion and blending restated from crates/cli, crates/render, crates/mmfx.
 * Test infrastructure only. */
#include <string.h>

#include "oracle.h"

static uint8_t clamp_u8(int v) { return (uint8_t)(v < 0 ? 0 : (v > 255 ? 255 : v)); }

/* ycbcr_to_rgb, crates/cli/src/terminal_preview.rs:7308-7331 (BT.601 limited / full).
 * BT.709 has no reference implementation; it uses the same integer form with the
 * BT.709 coefficients scaled by 256: limited 298/459/55/136/541, full 403/48/120/475. */
void oracle_ycbcr_to_rgb(uint8_t y, uint8_t cb, uint8_t cr, uint32_t range, uint32_t matrix,
                         uint8_t rgb[3]) {
    int b = (int)cb - 128, r = (int)cr - 128;
    int red, green, blue;
    if (range == MMR_RANGE_LIMITED) {
        int yy = (int)y - 16;
        if (yy < 0) yy = 0;
        if (matrix == MMR_MATRIX_BT709) {
            red = (298 * yy + 459 * r + 128) >> 8;
            green = (298 * yy - 55 * b - 136 * r + 128) >> 8;
            blue = (298 * yy + 541 * b + 128) >> 8;
        } else {
            red = (298 * yy + 409 * r + 128) >> 8;
            green = (298 * yy - 100 * b - 208 * r + 128) >> 8;
            blue = (298 * yy + 516 * b + 128) >> 8;
        }
    } else {
        int yy = (int)y;
        if (matrix == MMR_MATRIX_BT709) {
            red = yy + ((403 * r + 128) >> 8);
            green = yy - ((48 * b + 120 * r + 128) >> 8);
            blue = yy + ((475 * b + 128) >> 8);
        } else {
            red = yy + ((359 * r + 128) >> 8);
            green = yy - ((88 * b + 183 * r + 128) >> 8);
            blue = yy + ((454 * b + 128) >> 8);
        }
    }
    rgb[0] = clamp_u8(red);
    rgb[1] = clamp_u8(green);
    rgb[2] = clamp_u8(blue);
}

/* video_frame_image at native size, crates/cli/src/terminal_preview.rs:7159-7220:
 * chroma column = min(x * cw / W, cw - 1), chroma row = min(y * ch / H, ch - 1). */
int oracle_yuv_to_rgb(const mmr_recon_yuv_to_rgb_job *j) {
    if (j->width == 0 || j->height == 0 || j->chroma_width == 0 || j->chroma_height == 0) return -1;
    for (uint32_t y = 0; y < j->height; ++y) {
        uint32_t cy = (uint32_t)((uint64_t)y * j->chroma_height / j->height);
        if (cy > j->chroma_height - 1) cy = j->chroma_height - 1;
        for (uint32_t x = 0; x < j->width; ++x) {
            uint32_t cx = (uint32_t)((uint64_t)x * j->chroma_width / j->width);
            if (cx > j->chroma_width - 1) cx = j->chroma_width - 1;
            uint8_t rgb[3];
            oracle_ycbcr_to_rgb(j->planes[0][(size_t)y * j->luma_stride + x],
                                j->planes[1][(size_t)cy * j->chroma_stride + cx],
                                j->planes[2][(size_t)cy * j->chroma_stride + cx], j->range, j->matrix, rgb);
            size_t p = (size_t)y * j->width + x;
            if (j->rgb) memcpy(j->rgb + p * 3, rgb, 3);
            if (j->rgba) {
                memcpy(j->rgba + p * 4, rgb, 3);
                j->rgba[p * 4 + 3] = 255;
            }
        }
    }
    return 0;
}

/* blend_rgba_pixel, crates/render/src/frame_graph.rs:638-657 (straight-alpha source-over). */
void oracle_blend_rgba_pixel(uint8_t d[4], const uint8_t s[4]) {
    uint32_t sa = s[3];
    if (sa == 0) return;
    if (sa == 255 || d[3] == 0) {
        memcpy(d, s, 4);
        return;
    }
    uint32_t da = d[3];
    uint32_t inverse = 255 - sa;
    uint32_t oa = sa + (da * inverse + 127) / 255;
    for (int c = 0; c < 3; ++c) {
        uint32_t pre = (uint32_t)s[c] * sa + ((uint32_t)d[c] * da * inverse + 127) / 255;
        d[c] = (uint8_t)((pre + oa / 2) / oa);
    }
    d[3] = (uint8_t)oa;
}

/* LinearPixel::over with multiply_channel / add_saturating, crates/mmfx/src/render.rs:1826-1854,
 * with CHANNEL_MAX = 255 (8-bit premultiplied channels) instead of 65535. */
static uint32_t mul255(uint32_t a, uint32_t b) { return (a * b + 127) / 255; }
static uint8_t sat_add(uint32_t a, uint32_t b) { return (uint8_t)(a + b > 255 ? 255 : a + b); }
void oracle_premultiplied_over_pixel(uint8_t d[4], const uint8_t s[4], uint32_t opacity) {
    uint32_t sa = mul255(s[3], opacity);
    uint32_t inverse = 255 - sa;
    for (int c = 0; c < 3; ++c) d[c] = sat_add(mul255(s[c], opacity), mul255(d[c], inverse));
    d[3] = sat_add(sa, mul255(d[3], inverse));
}

/* blend_channel, crates/render/src/frame_graph.rs:659-671. */
uint8_t oracle_blend_channel(uint8_t d, uint8_t s, uint8_t a) {
    if (a == 0) return d;
    if (a == 255) return s;
    return (uint8_t)(((uint32_t)s * a + (uint32_t)d * (255u - a) + 127) / 255);
}

/* blend_rgba8, crates/render/src/frame_graph.rs:504-545, layers applied in pass order. */
int oracle_composite(const mmr_recon_composite_job *j) {
    for (uint32_t l = 0; l < j->layer_count; ++l) {
        const mmr_recon_layer *L = &j->layers[l];
        if (L->x + L->width > j->width || L->y + L->height > j->height) return -1;
        for (uint32_t r = 0; r < L->height; ++r)
            for (uint32_t c = 0; c < L->width; ++c) {
                uint8_t *d = j->target + (size_t)(L->y + r) * j->stride + (size_t)(L->x + c) * 4;
                const uint8_t *s = L->rgba + (size_t)r * L->stride + (size_t)c * 4;
                if (L->op == MMR_BLEND_PREMULTIPLIED_OVER)
                    oracle_premultiplied_over_pixel(d, s, L->opacity);
                else
                    oracle_blend_rgba_pixel(d, s);
            }
    }
    return 0;
}

/* blend_yuv420, crates/render/src/frame_graph.rs:548-635. */
int oracle_yuv_blend(const mmr_recon_yuv_blend_job *j) {
    if ((j->width | j->height | j->x | j->y) & 1) return -1;
    if (j->x + j->width > j->target_width || j->y + j->height > j->target_height) return -1;
    for (uint32_t r = 0; r < j->height; ++r)
        for (uint32_t c = 0; c < j->width; ++c) {
            size_t s = (size_t)r * j->width + c;
            uint8_t *d = j->target[0] + (size_t)(j->y + r) * j->target_stride[0] + j->x + c;
            *d = oracle_blend_channel(*d, j->overlay[0][s], j->rgba[s * 4 + 3]);
        }
    uint32_t cw = j->width / 2, ch = j->height / 2;
    for (uint32_t r = 0; r < ch; ++r)
        for (uint32_t c = 0; c < cw; ++c) {
            size_t s = (size_t)r * cw + c;
            uint8_t a = j->chroma_alpha[s];
            uint8_t *u = j->target[1] + (size_t)(j->y / 2 + r) * j->target_stride[1] + j->x / 2 + c;
            uint8_t *v = j->target[2] + (size_t)(j->y / 2 + r) * j->target_stride[2] + j->x / 2 + c;
            *u = oracle_blend_channel(*u, j->overlay[1][s], a);
            *v = oracle_blend_channel(*v, j->overlay[2][s], a);
        }
    return 0;
}
