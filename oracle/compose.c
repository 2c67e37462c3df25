/* Oracle: colour conversion and blending restated from crates/cli, crates/render, crates/mmfx.
 * Test infrastructure only. */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
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

/* rgb_to_yuv, crates/render/src/project_compositor.rs:1163-1185: BT.709 limited and full range. */
void oracle_rgb_to_yuv(uint8_t red, uint8_t green, uint8_t blue, uint8_t limited[3], uint8_t full[3]) {
    const int r = red, g = green, b = blue;
    int ly = 16 + ((47 * r + 157 * g + 16 * b + 128) >> 8);
    int lu = 128 + ((-26 * r - 87 * g + 112 * b + 128) >> 8);
    int lv = 128 + ((112 * r - 102 * g - 10 * b + 128) >> 8);
    int fy = (54 * r + 183 * g + 18 * b + 128) >> 8;
    int fu = 128 + ((-29 * r - 99 * g + 128 * b + 128) >> 8);
    int fv = 128 + ((128 * r - 116 * g - 12 * b + 128) >> 8);
    limited[0] = (uint8_t)(ly < 16 ? 16 : (ly > 235 ? 235 : ly));
    limited[1] = (uint8_t)(lu < 16 ? 16 : (lu > 240 ? 240 : lu));
    limited[2] = (uint8_t)(lv < 16 ? 16 : (lv > 240 ? 240 : lv));
    full[0] = clamp_u8(fy);
    full[1] = clamp_u8(fu);
    full[2] = clamp_u8(fv);
}

/* PreparedOverlay::prepare_yuv, project_compositor.rs:1098-1142. */
int oracle_overlay_prepare(const mmr_recon_overlay_prepare_job *j) {
    const uint32_t w = j->width, h = j->height, cw = (w + 1) / 2;
    for (uint32_t i = 0; i < w * h; ++i) {
        uint8_t lim[3], full[3];
        oracle_rgb_to_yuv(j->rgba[4 * i], j->rgba[4 * i + 1], j->rgba[4 * i + 2], lim, full);
        if (j->y_limited) j->y_limited[i] = lim[0];
        if (j->y_full) j->y_full[i] = full[0];
    }
    for (uint32_t y = 0; y < h; y += 2)
        for (uint32_t x = 0; x < w; x += 2) {
            uint32_t alpha_sum = 0, lu = 0, lv = 0, fu = 0, fv = 0;
            const uint32_t nh = h - y < 2 ? h - y : 2, nw = w - x < 2 ? w - x : 2;
            for (uint32_t dy = 0; dy < nh; ++dy)
                for (uint32_t dx = 0; dx < nw; ++dx) {
                    const uint8_t *p = j->rgba + ((size_t)(y + dy) * w + x + dx) * 4;
                    uint8_t lim[3], full[3];
                    oracle_rgb_to_yuv(p[0], p[1], p[2], lim, full);
                    alpha_sum += p[3];
                    lu += (uint32_t)lim[1] * p[3];
                    lv += (uint32_t)lim[2] * p[3];
                    fu += (uint32_t)full[1] * p[3];
                    fv += (uint32_t)full[2] * p[3];
                }
            const uint32_t n = nh * nw, c = (y / 2) * cw + x / 2;
#define WEIGHTED(sum) (uint8_t)(alpha_sum ? ((sum) + alpha_sum / 2) / alpha_sum : 128)
            if (j->chroma_alpha) j->chroma_alpha[c] = (uint8_t)((alpha_sum + n / 2) / n);
            if (j->u_limited) j->u_limited[c] = WEIGHTED(lu);
            if (j->v_limited) j->v_limited[c] = WEIGHTED(lv);
            if (j->u_full) j->u_full[c] = WEIGHTED(fu);
            if (j->v_full) j->v_full[c] = WEIGHTED(fv);
#undef WEIGHTED
        }
    return 0;
}

/* ---------------------------------------------------------------- Lanczos3 scaling
 * image::imageops::resize with FilterType::Lanczos3 or ::Triangle (the image crate 0.25.6, a
 * dependency of the reference that it does not vendor), as called by scale_yuv420_to_canvas_cpu
 * (crates/render/src/compositor.rs:89-141): a vertical pass into f32, then a horizontal pass; for
 * output coordinate o of a pass from n to m samples, ratio = n / m, sratio = max(ratio, 1),
 * support = 3 sratio (1 sratio for Triangle), centre c = (o + 0.5) ratio, taps left = clamp(floor(c - support), 0, n - 1)
 * .. right = clamp(ceil(c + support), left + 1, n), weight of tap i = lanczos3((i - (c - 0.5)) /
 * sratio) divided by the sum of the weights, lanczos3(x) = sinc(x) sinc(x / 3) for |x| < 3,
 * sinc(t) = sin(pi t) / (pi t), triangle(x) = 1 - |x| for |x| < 1; accumulation in f32 in tap order; the last pass rounds
 * (half away from zero) the value clamped to [0, 255]. All f32, no fused multiply-add. */
static float lz_sinc(float t) {
    const float a = t * 3.14159265358979323846f;
    return t == 0.0f ? 1.0f : sinf(a) / a;
}
static float lz_kernel(float x) { return fabsf(x) < 3.0f ? lz_sinc(x) * lz_sinc(x / 3.0f) : 0.0f; }
/* FilterType::Triangle: 1 - |x| for |x| < 1, support 1. */
static float tri_kernel(float x) { return fabsf(x) < 1.0f ? 1.0f - fabsf(x) : 0.0f; }

/* Taps of output o of filter f (MMR_SCALE_FILTER_*): *left and the count; weights into w
 * (capacity cap). Returns the count, -1 when it exceeds cap. */
int oracle_scale_taps(uint32_t filter, uint32_t n, uint32_t m, uint32_t o, int32_t *left, float *w, int cap) {
    const float ratio = (float)n / (float)m;
    const float sratio = ratio < 1.0f ? 1.0f : ratio;
    const float support = (filter == MMR_SCALE_FILTER_TRIANGLE ? 1.0f : 3.0f) * sratio;
    float c = ((float)o + 0.5f) * ratio;
    int64_t l = (int64_t)floorf(c - support);
    if (l < 0) l = 0;
    if (l > (int64_t)n - 1) l = (int64_t)n - 1;
    int64_t r = (int64_t)ceilf(c + support);
    if (r < l + 1) r = l + 1;
    if (r > (int64_t)n) r = (int64_t)n;
    c = c - 0.5f;
    const int count = (int)(r - l);
    if (count > cap) return -1;
    float sum = 0.0f;
    for (int i = 0; i < count; ++i) {
        const float x = ((float)(l + i) - c) / sratio;
        w[i] = filter == MMR_SCALE_FILTER_TRIANGLE ? tri_kernel(x) : lz_kernel(x);
        sum += w[i];
    }
    for (int i = 0; i < count; ++i) w[i] /= sum;
    *left = (int32_t)l;
    return count;
}

static float *lz_resize_vertical(uint32_t filter, const uint8_t *src, uint32_t stride, uint32_t w, uint32_t h, uint32_t nh) {
    float *out = (float *)malloc((size_t)w * nh * sizeof(float));
    float *wt = (float *)malloc(((size_t)h + 1) * sizeof(float)); /* a pass has at most h taps */
    if (!out || !wt) {
        free(out);
        free(wt);
        return NULL;
    }
    for (uint32_t oy = 0; oy < nh; ++oy) {
        int32_t left;
        const int count = oracle_scale_taps(filter, h, nh, oy, &left, wt, (int)h + 1);
        if (count < 0) {
            free(out);
            free(wt);
            return NULL;
        }
        for (uint32_t x = 0; x < w; ++x) {
            float t = 0.0f;
            for (int i = 0; i < count; ++i) t += (float)src[(size_t)(left + i) * stride + x] * wt[i];
            out[(size_t)oy * w + x] = t;
        }
    }
    free(wt);
    return out;
}

/* scaled_even / scaled_dimensions, compositor.rs:211-254. */
static uint64_t lz_scaled_even(uint64_t source, uint64_t target, uint64_t denominator) {
    uint64_t units = (source * target + denominator) / (2 * denominator);
    return units * 2 < 2 ? 2 : units * 2;
}
void oracle_scaled_dimensions(uint32_t sw, uint32_t sh, uint32_t cw, uint32_t ch, uint32_t mode, uint32_t *ow, uint32_t *oh) {
    if (mode == MMR_SCALE_STRETCH) {
        *ow = cw;
        *oh = ch;
        return;
    }
    if (mode == MMR_SCALE_NATIVE) {
        *ow = sw;
        *oh = sh;
        return;
    }
    const int width_limited = (uint64_t)cw * sh <= (uint64_t)ch * sw;
    const int use_width = mode == MMR_SCALE_FILL ? !width_limited : width_limited;
    if (use_width) {
        *ow = cw;
        *oh = (uint32_t)lz_scaled_even(sh, cw, sw);
    } else {
        *ow = (uint32_t)lz_scaled_even(sw, ch, sh);
        *oh = ch;
    }
}

/* scale_yuv420_to_canvas_cpu + center_crop_or_pad, compositor.rs:89-141, :277-305. */
int oracle_scale_yuv420(const mmr_recon_scale_job *j) {
    uint32_t sw, sh;
    oracle_scaled_dimensions(j->width, j->height, j->canvas_width, j->canvas_height, j->mode, &sw, &sh);
    const uint8_t luma_black = j->range == MMR_RANGE_LIMITED ? 16 : 0;
    for (int k = 0; k < 3; ++k) {
        const uint32_t div = k ? 2 : 1;
        const uint32_t pw = j->width / div, ph = j->height / div, nw = sw / div, nh = sh / div;
        const uint32_t W = j->canvas_width / div, H = j->canvas_height / div;
        const uint8_t bg = k ? 128 : luma_black;
        /* the scaled plane, as u8 */
        uint8_t *scaled = (uint8_t *)malloc((size_t)nw * nh);
        if (!scaled) return -1;
        if (nw == pw && nh == ph) {
            for (uint32_t y = 0; y < ph; ++y) memcpy(scaled + (size_t)y * pw, j->src[k] + (size_t)y * j->src_stride[k], pw);
        } else {
            float *tmp = lz_resize_vertical(j->filter, j->src[k], j->src_stride[k], pw, ph, nh);
            float *wt = (float *)malloc(((size_t)pw + 1) * sizeof(float));
            if (!tmp || !wt) {
                free(tmp);
                free(wt);
                free(scaled);
                return -1;
            }
            for (uint32_t ox = 0; ox < nw; ++ox) {
                int32_t left;
                const int count = oracle_scale_taps(j->filter, pw, nw, ox, &left, wt, (int)pw + 1);
                for (uint32_t y = 0; y < nh; ++y) {
                    float t = 0.0f;
                    for (int i = 0; i < count; ++i) t += tmp[(size_t)y * pw + left + i] * wt[i];
                    t = t < 0.0f ? 0.0f : (t > 255.0f ? 255.0f : t);
                    scaled[(size_t)y * nw + ox] = (uint8_t)roundf(t);
                }
            }
            free(tmp);
            free(wt);
        }
        const uint32_t copy_w = nw < W ? nw : W, copy_h = nh < H ? nh : H;
        const uint32_t sx = (nw - copy_w) / 2, sy = (nh - copy_h) / 2, dx = (W - copy_w) / 2, dy = (H - copy_h) / 2;
        for (uint32_t y = 0; y < H; ++y) memset(j->dst[k] + (size_t)y * j->dst_stride[k], bg, W);
        for (uint32_t y = 0; y < copy_h; ++y)
            memcpy(j->dst[k] + (size_t)(dy + y) * j->dst_stride[k] + dx, scaled + (size_t)(sy + y) * nw + sx, copy_w);
        free(scaled);
    }
    return 0;
}

/* ---------------------------------------------------------------- MMFX linear light (u16)
 * LinearPixel (crates/mmfx/src/render.rs:1780-1880): premultiplied linear-light RGBA with
 * CHANNEL_MAX = 65535 (:23). */
#define LINEAR_MAX 65535u

/* multiply_channel (:1847-1850): (left * right + CHANNEL_MAX / 2) / CHANNEL_MAX, saturated to u16. */
uint16_t oracle_linear_multiply_channel(uint16_t left, uint16_t right) {
    const uint32_t q = ((uint32_t)left * right + LINEAR_MAX / 2) / LINEAR_MAX;
    return (uint16_t)(q > 65535u ? 65535u : q);
}

/* add_saturating (:1852-1854). */
static uint16_t linear_add_saturating(uint16_t a, uint16_t b) {
    const uint32_t s = (uint32_t)a + b;
    return (uint16_t)(s > 65535u ? 65535u : s);
}

/* srgb_to_linear_u16 (:1856-1867): f64, the C library's pow (Rust's powf on Linux), round half away. */
uint16_t oracle_srgb_to_linear_u16(uint8_t value) {
    const double encoded = (double)value / 255.0;
    const double linear = encoded <= 0.04045 ? encoded / 12.92 : pow((encoded + 0.055) / 1.055, 2.4);
    return (uint16_t)round(linear * 65535.0);
}

/* linear_u16_to_srgb (:1869-1880). */
uint8_t oracle_linear_u16_to_srgb(uint16_t value) {
    const double linear = (double)value / 65535.0;
    const double encoded = linear <= 0.0031308 ? linear * 12.92 : 1.055 * pow(linear, 1.0 / 2.4) - 0.055;
    double r = round(encoded * 255.0);
    r = r < 0.0 ? 0.0 : (r > 255.0 ? 255.0 : r);
    return (uint8_t)r;
}

/* LinearPixel::from_color / from_rgba (:1796-1808): alpha = a * 257, channels premultiplied. */
void oracle_linear_from_rgba(const uint8_t rgba[4], uint16_t out[4]) {
    const uint16_t alpha = (uint16_t)(rgba[3] * 257u);
    for (int c = 0; c < 3; ++c) out[c] = oracle_linear_multiply_channel(oracle_srgb_to_linear_u16(rgba[c]), alpha);
    out[3] = alpha;
}

/* LinearPixel::over (:1826-1845): source over destination with an extra opacity. */
void oracle_linear_over(const uint16_t src[4], uint16_t dst[4], uint16_t opacity) {
    const uint16_t source_alpha = oracle_linear_multiply_channel(src[3], opacity);
    const uint16_t inverse = (uint16_t)(65535u - source_alpha);
    for (int c = 0; c < 3; ++c)
        dst[c] = linear_add_saturating(oracle_linear_multiply_channel(src[c], opacity),
                                       oracle_linear_multiply_channel(dst[c], inverse));
    dst[3] = linear_add_saturating(source_alpha, oracle_linear_multiply_channel(dst[3], inverse));
}

/* Surface::to_rgba8 (:233-254) for one pixel. */
void oracle_linear_to_rgba8(const uint16_t px[4], uint8_t out[4]) {
    const uint32_t alpha = px[3];
    if (alpha == 0) {
        out[0] = out[1] = out[2] = out[3] = 0;
        return;
    }
    for (int c = 0; c < 3; ++c) {
        uint32_t linear = ((uint32_t)px[c] * LINEAR_MAX + alpha / 2) / alpha;
        if (linear > LINEAR_MAX) linear = LINEAR_MAX;
        out[c] = oracle_linear_u16_to_srgb((uint16_t)linear);
    }
    const uint32_t a8 = (alpha + 128) / 257;
    out[3] = (uint8_t)(a8 > 255 ? 255 : a8);
}

/* Layers in order: Surface::blend_surface (:269-277) for surface layers, execute_draw_image
 * (:1410-1444) for sRGBA8 layers at integer destinations of the image's size (every target
 * pixel in the clip and the layer takes source.over(target, opacity)). */
int oracle_linear_composite(const mmr_recon_linear_composite_job *j) {
    for (uint32_t l = 0; l < j->layer_count; ++l) {
        const mmr_recon_linear_layer *L = &j->layers[l];
        const int64_t top = L->clip_top > 0 ? L->clip_top : 0;
        const int64_t left = L->clip_left > 0 ? L->clip_left : 0;
        const int64_t bottom = L->clip_bottom < (int64_t)j->height ? L->clip_bottom : (int64_t)j->height;
        const int64_t right = L->clip_right < (int64_t)j->width ? L->clip_right : (int64_t)j->width;
        for (int64_t y = top; y < bottom; ++y)
            for (int64_t x = left; x < right; ++x) {
                const int64_t sx = x - L->x, sy = y - L->y;
                if (sx < 0 || sy < 0 || sx >= (int64_t)L->width || sy >= (int64_t)L->height) continue;
                uint16_t src[4];
                if (L->format == MMR_LINEAR_SOURCE_SRGBA8) {
                    oracle_linear_from_rgba((const uint8_t *)L->pixels + ((size_t)sy * L->stride + (size_t)sx) * 4, src);
                } else if (L->format == MMR_LINEAR_SOURCE_SURFACE) {
                    memcpy(src, (const uint16_t *)L->pixels + ((size_t)sy * L->stride + (size_t)sx) * 4, sizeof src);
                } else {
                    return -1;
                }
                oracle_linear_over(src, j->target + ((size_t)y * j->stride + (size_t)x) * 4, (uint16_t)L->opacity);
            }
    }
    return 0;
}

int oracle_linear_convert(const mmr_recon_linear_convert_job *j) {
    for (uint32_t y = 0; y < j->height; ++y)
        for (uint32_t x = 0; x < j->width; ++x) {
            uint16_t *lin = j->linear + ((size_t)y * j->linear_stride + x) * 4;
            uint8_t *rgba = j->rgba + (size_t)y * j->rgba_stride + (size_t)x * 4;
            if (j->direction == MMR_LINEAR_TO_SRGBA8) oracle_linear_to_rgba8(lin, rgba);
            else if (j->direction == MMR_LINEAR_FROM_SRGBA8) oracle_linear_from_rgba(rgba, lin);
            else return -1;
        }
    return 0;
}
