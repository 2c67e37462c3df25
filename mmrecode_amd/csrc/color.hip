// YCbCr -> RGB8, RGBA8 compositing and YUV 4:2:0 overlay blending.
//
// Reference: ycbcr_to_rgb + video_frame_image's chroma mapping
// (crates/cli/src/terminal_preview.rs:7159-7238, :7308-7331); blend_rgba8 / blend_rgba_pixel
// (crates/render/src/frame_graph.rs:504-545, :638-657); LinearPixel::over / multiply_channel /
// add_saturating (crates/mmfx/src/render.rs:1826-1854, here with 8-bit channels); blend_yuv420 /
// blend_channel (crates/render/src/frame_graph.rs:548-671).
#include "device_common.h"
#include "recon_internal.h"

namespace mmr {

// floor(n / d) for n < 2^32, d >= 1, from inv = 1.0 / d (rounded): n * inv is within
// 2^-51 * n / d < 1 of n / d, so the truncated estimate is floor(n / d) or one off; the remainder
// settles which.
MMR_DEV uint32_t div_floor(uint32_t n, uint32_t d, double inv) {
    uint32_t q = (uint32_t)((double)n * inv);
    const int64_t r = (int64_t)n - (int64_t)q * d;
    if (r < 0) --q;
    else if (r >= (int64_t)d) ++q;
    return q;
}

// Chroma terms of one (Cb, Cr) sample in binary32 (exact; RgbJob::kcb / kcr / k0); full range
// floors them here, limited range after adding the luma term.
struct ChromaTerms {
    float r, g, b;
};

template <bool kLimited>
MMR_DEV ChromaTerms chroma_terms(const RgbJob& j, uint32_t cb, uint32_t cr) {
    const float fb = (float)cb, fr = (float)cr;
    ChromaTerms t;
    t.r = __builtin_fmaf(j.kcr[0], fr, j.k0[0]);
    t.g = __builtin_fmaf(j.kcb[1], fb, __builtin_fmaf(j.kcr[1], fr, j.k0[1]));
    t.b = __builtin_fmaf(j.kcb[2], fb, j.k0[2]);
    if (!kLimited) {
        t.r = __builtin_floorf(t.r);
        t.g = __builtin_floorf(t.g);
        t.b = __builtin_floorf(t.b);
    }
    return t;
}

MMR_DEV float clamp_u8f(float v) { return __builtin_fminf(__builtin_fmaxf(v, 0.0f), 255.0f); }

// R, G, B of luma sample y under chroma terms t, each an integer in [0, 255] held as binary32.
template <bool kLimited>
MMR_DEV void rgb_of(uint32_t y, const ChromaTerms& t, float out[3]) {
    const float fy = (float)y;
    if (kLimited) {
        const float yy = __builtin_fmaxf(fy - 16.0f, 0.0f);
        out[0] = __builtin_floorf(__builtin_fmaf(298.0f / 256.0f, yy, t.r));
        out[1] = __builtin_floorf(__builtin_fmaf(298.0f / 256.0f, yy, t.g));
        out[2] = __builtin_floorf(__builtin_fmaf(298.0f / 256.0f, yy, t.b));
    } else {
        out[0] = fy + t.r;
        out[1] = fy - t.g;
        out[2] = fy + t.b;
    }
    out[0] = clamp_u8f(out[0]);
    out[1] = clamp_u8f(out[1]);
    out[2] = clamp_u8f(out[2]);
}

// Byte k (0..3) of dword d replaced by the integer v in [0, 255] held as binary32.
MMR_DEV uint32_t put_byte(uint32_t d, float v, uint32_t k) { return __builtin_amdgcn_cvt_pk_u8_f32(v, k, d); }

// Four horizontally adjacent pixels (x0 .. x0 + 3 of row y, clipped to the width; returns how many)
// per thread as R, G, B in [0, 255] held as binary32, from one luma dword. The chroma row is floor(y ch / height) and the chroma
// sample of pixel x floor(x cw / width) (video_frame_image): when 2 cw == width that is x / 2 and
// pixels 2k, 2k + 1 share one set of chroma terms; otherwise the first from div_floor, the next
// three by carrying the remainder (cw <= width, so each step moves at most one chroma sample).
template <bool kLimited>
MMR_DEV uint32_t yuv_px4(const RgbJob& j, uint32_t y, uint32_t x0, float px[4][3]) {
    const uint32_t cy = min(div_floor(y * j.ch, j.height, j.inv_height), j.ch - 1);
    const MMR_GLOBAL uint8_t* Y = as_global(j.planes[0]) + (size_t)y * j.luma_stride;
    const MMR_GLOBAL uint8_t* U = as_global(j.planes[1]) + (size_t)cy * j.chroma_stride;
    const MMR_GLOBAL uint8_t* V = as_global(j.planes[2]) + (size_t)cy * j.chroma_stride;
    const uint32_t n = min(4u, j.width - x0);
    uint32_t ys;
    if (n == 4 && ((reinterpret_cast<uintptr_t>(Y + x0) & 3) == 0)) {
        ys = *reinterpret_cast<const MMR_GLOBAL uint32_t*>(Y + x0);
    } else {
        ys = 0;
        for (uint32_t i = 0; i < n; ++i) ys |= (uint32_t)Y[x0 + i] << (8 * i);
    }
    if (j.half_chroma) {
        // x0 is a multiple of 4: chroma samples x0 / 2 and x0 / 2 + 1 (the second only when pixel
        // x0 + 2 exists, so it never passes cw - 1).
        const uint32_t c0 = x0 >> 1;
        const uint32_t c1 = n > 2 ? c0 + 1 : c0;
        uint32_t u0, u1, v0, v1;
        if (n > 2 && ((reinterpret_cast<uintptr_t>(U + c0) | reinterpret_cast<uintptr_t>(V + c0)) & 1) == 0) {   // both samples at once
            const uint32_t uu = *reinterpret_cast<const MMR_GLOBAL uint16_t*>(U + c0), vv = *reinterpret_cast<const MMR_GLOBAL uint16_t*>(V + c0);
            u0 = uu & 0xffu;
            u1 = uu >> 8;
            v0 = vv & 0xffu;
            v1 = vv >> 8;
        } else {
            u0 = U[c0];
            u1 = U[c1];
            v0 = V[c0];
            v1 = V[c1];
        }
        const ChromaTerms t0 = chroma_terms<kLimited>(j, u0, v0);
        const ChromaTerms t1 = chroma_terms<kLimited>(j, u1, v1);
        rgb_of<kLimited>(ys & 0xffu, t0, px[0]);
        rgb_of<kLimited>((ys >> 8) & 0xffu, t0, px[1]);
        rgb_of<kLimited>((ys >> 16) & 0xffu, t1, px[2]);
        rgb_of<kLimited>(ys >> 24, t1, px[3]);
    } else {
        uint32_t cx = div_floor(x0 * j.cw, j.width, j.inv_width);
        uint32_t rem = x0 * j.cw - cx * j.width;
#pragma unroll
        for (uint32_t i = 0; i < 4; ++i) {
            const uint32_t c = min(cx, j.cw - 1);
            rgb_of<kLimited>((ys >> (8 * i)) & 0xffu, chroma_terms<kLimited>(j, U[c], V[c]), px[i]);
            rem += j.cw;
            if (rem >= j.width) {
                rem -= j.width;
                ++cx;
            }
        }
    }
    return n;
}

// One lane's four pixels (yuv_px4) stored as RGB8 and / or RGBA8 (alpha 255), rows packed.
template <bool kLimited>
MMR_DEV void yuv_to_rgb_px4(const RgbJob& j, uint32_t y, uint32_t x0) {
    float px[4][3];
    const uint32_t n = yuv_px4<kLimited>(j, y, x0, px);
    const size_t p = (size_t)y * j.width + x0;
    if (j.rgb) {
        MMR_GLOBAL uint8_t* o = as_global(j.rgb) + p * 3;
        if (n == 4 && (((uintptr_t)o) & 3) == 0) {   // 12 bytes as three dwords
            uint32_t d0 = put_byte(0u, px[0][0], 0), d1 = put_byte(0u, px[1][1], 0), d2 = put_byte(0u, px[2][2], 0);
            d0 = put_byte(d0, px[0][1], 1);
            d0 = put_byte(d0, px[0][2], 2);
            d0 = put_byte(d0, px[1][0], 3);
            d1 = put_byte(d1, px[1][2], 1);
            d1 = put_byte(d1, px[2][0], 2);
            d1 = put_byte(d1, px[2][1], 3);
            d2 = put_byte(d2, px[3][0], 1);
            d2 = put_byte(d2, px[3][1], 2);
            d2 = put_byte(d2, px[3][2], 3);
            mmr_v3u32 v;   // one 12-byte store: a wave's stores are contiguous
            v.x = d0;
            v.y = d1;
            v.z = d2;
            *reinterpret_cast<MMR_GLOBAL mmr_v3u32*>(o) = v;
        } else {
            for (uint32_t i = 0; i < n; ++i) {
                o[i * 3 + 0] = (uint8_t)px[i][0];
                o[i * 3 + 1] = (uint8_t)px[i][1];
                o[i * 3 + 2] = (uint8_t)px[i][2];
            }
        }
    }
    if (j.rgba) {
        MMR_GLOBAL uint8_t* o = as_global(j.rgba) + p * 4;
        if (n == 4 && (((uintptr_t)o) & 15) == 0) {
            mmr_v4i32 v;
            for (int i = 0; i < 4; ++i) {
                uint32_t d = put_byte(0xff000000u, px[i][0], 0);
                d = put_byte(d, px[i][1], 1);
                v[i] = (int)put_byte(d, px[i][2], 2);
            }
            *reinterpret_cast<MMR_GLOBAL mmr_v4i32*>(o) = v;
        } else {
            for (uint32_t i = 0; i < n; ++i) {
                o[i * 4 + 0] = (uint8_t)px[i][0];
                o[i * 4 + 1] = (uint8_t)px[i][1];
                o[i * 4 + 2] = (uint8_t)px[i][2];
                o[i * 4 + 3] = 255;
            }
        }
    }
}

__global__ __launch_bounds__(256) void yuv_to_rgb_kernel(const RgbJob* __restrict__ jobs, const uint32_t* __restrict__ tile_job,
                                                         uint32_t tile_base, uint32_t tiles, uint32_t per_xcd) {
    const uint32_t local = xcd_tile(blockIdx.x, per_xcd);
    if (local >= tiles) return;
    const uint32_t tile = tile_base + local;
    const RgbJob& j = jobs[tile_job[tile]];
    const uint32_t wave = (uint32_t)__builtin_amdgcn_readfirstlane((int)((tile - j.tile_begin) * 4 + (threadIdx.x >> 6)));
    const uint32_t lane = threadIdx.x & 63;
    // The wave's segments (see set_rgb_layout); the job's last tile may hold fewer.
    for (uint32_t k = 0; k < kRgbSegmentsPerWave; ++k) {
        const uint32_t seg = wave * kRgbSegmentsPerWave + k;
        if (seg >= j.segments) break;   // uniform over the wave
        uint32_t y, x0;
        if (j.flat) {   // raster order: a lane's 4 pixels lie in one row (width % 4 == 0)
            const uint32_t p = seg * 256 + lane * 4;
            if (p >= j.width * j.height) continue;
            y = div_floor(p, j.width, j.inv_width);
            x0 = p - y * j.width;
        } else {
            y = seg / j.waves_per_row;
            x0 = (seg - y * j.waves_per_row) * 256 + lane * 4;
            if (x0 >= j.width) continue;
        }
        if (j.limited) yuv_to_rgb_px4<true>(j, y, x0);
        else yuv_to_rgb_px4<false>(j, y, x0);
    }
}

// ---------------------------------------------------------------- MMFX linear light (u16)
// LinearPixel (crates/mmfx/src/render.rs:1780-1880): premultiplied linear RGBA, CHANNEL_MAX 65535.

// multiply_channel (:1847-1850): (l * r + 32767) / 65535; the product is below 2^32 and the
// quotient at most 65535, so it needs no saturation.
MMR_DEV uint32_t linear_mul(uint32_t l, uint32_t r) { return (l * r + 32767u) / 65535u; }

// LinearPixel::over (:1826-1845) with add_saturating (:1852-1854), in place on d.
MMR_DEV void linear_over(const uint32_t s[4], uint32_t d[4], uint32_t opacity) {
    const uint32_t sa = linear_mul(s[3], opacity);
    const uint32_t inverse = 65535u - sa;
#pragma unroll
    for (int c = 0; c < 3; ++c) d[c] = min(linear_mul(s[c], opacity) + linear_mul(d[c], inverse), 65535u);
    d[3] = min(sa + linear_mul(d[3], inverse), 65535u);
}

// LinearPixel::from_rgba (:1796-1808) through the srgb_to_linear_u16 table.
MMR_DEV void linear_from_rgba(uint32_t rgba, const uint16_t* __restrict__ to_linear, uint32_t out[4]) {
    const uint32_t alpha = (rgba >> 24) * 257u;
#pragma unroll
    for (int c = 0; c < 3; ++c) out[c] = linear_mul(to_linear[(rgba >> (8 * c)) & 0xffu], alpha);
    out[3] = alpha;
}

MMR_DEV void unpack_linear(mmr_v2u32 v, uint32_t p[4]) {
    p[0] = v.x & 0xffffu;
    p[1] = v.x >> 16;
    p[2] = v.y & 0xffffu;
    p[3] = v.y >> 16;
}

MMR_DEV mmr_v2u32 pack_linear(const uint32_t p[4]) {
    mmr_v2u32 v;
    v.x = p[0] | (p[1] << 16);
    v.y = p[2] | (p[3] << 16);
    return v;
}

// The wave's row and this lane's pixel of a 64-pixel-per-wave linear job; false past the job.
MMR_DEV bool linear_position(uint32_t tile, uint32_t tile_begin, uint32_t waves_per_row, uint32_t width, uint32_t height,
                             uint32_t* y, uint32_t* x) {
    const uint32_t wave = (uint32_t)__builtin_amdgcn_readfirstlane((int)((tile - tile_begin) * 4 + (threadIdx.x >> 6)));
    *y = wave / waves_per_row;
    *x = (wave - *y * waves_per_row) * 64 + (threadIdx.x & 63);
    return *y < height && *x < width;
}

// One lane per target pixel: the layers in order, each where the pixel lies inside its clip and
// its extent (Surface::blend_surface :269-277; execute_draw_image :1410-1444 at integer
// destinations of the image's size).
__global__ __launch_bounds__(256) void linear_composite_kernel(const LinearCompositeDev* __restrict__ jobs,
                                                               const uint32_t* __restrict__ tile_job, uint32_t tile_base,
                                                               uint32_t tiles, const uint16_t* __restrict__ to_linear,
                                                               uint32_t per_xcd) {
    const uint32_t local = xcd_tile(blockIdx.x, per_xcd);
    if (local >= tiles) return;
    const uint32_t tile = tile_base + local;
    const LinearCompositeDev& j = jobs[tile_job[tile]];
    uint32_t y, x;
    if (!linear_position(tile, j.tile_begin, j.waves_per_row, j.width, j.height, &y, &x)) return;
    MMR_GLOBAL mmr_v2u32* t = reinterpret_cast<MMR_GLOBAL mmr_v2u32*>(as_global(j.target) + ((size_t)y * j.stride + x) * 4);
    uint32_t d[4];
    unpack_linear(*t, d);
    for (uint32_t l = 0; l < j.layer_count; ++l) {
        const mmr_recon_linear_layer& L = j.layers[l];
        if ((int64_t)y < L.clip_top || (int64_t)y >= L.clip_bottom || (int64_t)x < L.clip_left || (int64_t)x >= L.clip_right)
            continue;
        const int64_t sx = (int64_t)x - L.x, sy = (int64_t)y - L.y;
        if (sx < 0 || sy < 0 || sx >= (int64_t)L.width || sy >= (int64_t)L.height) continue;
        const size_t at = (size_t)sy * L.stride + (size_t)sx;
        uint32_t src[4];
        if (L.format == MMR_LINEAR_SOURCE_SRGBA8) {
            linear_from_rgba(as_global(reinterpret_cast<const uint32_t*>(L.pixels))[at], to_linear, src);
        } else {
            unpack_linear(as_global(reinterpret_cast<const mmr_v2u32*>(L.pixels))[at], src);
        }
        linear_over(src, d, L.opacity);
    }
    *t = pack_linear(d);
}

// Surface::to_rgba8 (:233-254) or LinearPixel::from_rgba per pixel; tables: srgb_to_linear_u16
// (256 x u16) then linear_u16_to_srgb (65536 x u8).
__global__ __launch_bounds__(256) void linear_convert_kernel(const LinearConvertDev* __restrict__ jobs,
                                                             const uint32_t* __restrict__ tile_job, uint32_t tile_base,
                                                             uint32_t tiles, const uint8_t* __restrict__ tables,
                                                             uint32_t per_xcd) {
    const uint32_t local = xcd_tile(blockIdx.x, per_xcd);
    if (local >= tiles) return;
    const uint32_t tile = tile_base + local;
    const LinearConvertDev& j = jobs[tile_job[tile]];
    uint32_t y, x;
    if (!linear_position(tile, j.tile_begin, j.waves_per_row, j.width, j.height, &y, &x)) return;
    const uint16_t* to_linear = reinterpret_cast<const uint16_t*>(tables);
    const MMR_GLOBAL uint8_t* to_srgb = as_global(tables + 512);
    MMR_GLOBAL mmr_v2u32* lin = reinterpret_cast<MMR_GLOBAL mmr_v2u32*>(as_global(j.linear) + ((size_t)y * j.linear_stride + x) * 4);
    MMR_GLOBAL uint32_t* rgba = reinterpret_cast<MMR_GLOBAL uint32_t*>(as_global(j.rgba) + (size_t)y * j.rgba_stride) + x;
    if (j.direction == MMR_LINEAR_FROM_SRGBA8) {
        uint32_t p[4];
        linear_from_rgba(*rgba, to_linear, p);
        *lin = pack_linear(p);
        return;
    }
    uint32_t p[4];
    unpack_linear(*lin, p);
    const uint32_t alpha = p[3];
    uint32_t out = 0;
    if (alpha != 0) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const uint32_t linear = min((p[c] * 65535u + alpha / 2) / alpha, 65535u);
            out |= (uint32_t)to_srgb[linear] << (8 * c);
        }
        out |= min((alpha + 128u) / 257u, 255u) << 24;
    }
    *rgba = out;
}

MMR_DEV void blend_rgba_pixel(uint32_t d[4], const uint32_t s[4]) {
    const uint32_t sa = s[3];
    if (sa == 0) return;
    if (sa == 255 || d[3] == 0) {
        d[0] = s[0]; d[1] = s[1]; d[2] = s[2]; d[3] = s[3];
        return;
    }
    const uint32_t da = d[3], inverse = 255 - sa;
    const uint32_t oa = sa + (da * inverse + 127) / 255;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const uint32_t pre = s[c] * sa + (d[c] * da * inverse + 127) / 255;
        d[c] = (pre + oa / 2) / oa;
    }
    d[3] = oa;
}

// (a b + 127) / 255 for a, b <= 255: t / 255 == (t + 1 + (t >> 8)) >> 8 for every t <= 65534.
MMR_DEV uint32_t mul255(uint32_t a, uint32_t b) {
    const uint32_t t = a * b + 127;
    return (t + 1 + (t >> 8)) >> 8;
}
MMR_DEV uint32_t sat255(uint32_t v) { return v > 255 ? 255 : v; }

// `opaque`: opacity 255, where mul255(v, 255) == v.
MMR_DEV void premultiplied_over(uint32_t d[4], const uint32_t s[4], uint32_t opacity, bool opaque) {
    const uint32_t sa = opaque ? s[3] : mul255(s[3], opacity), inverse = 255 - sa;
#pragma unroll
    for (int c = 0; c < 3; ++c) d[c] = sat255((opaque ? s[c] : mul255(s[c], opacity)) + mul255(d[c], inverse));
    d[3] = sat255(sa + mul255(d[3], inverse));
}

// The layers of job j applied in order to this lane's pixels x0 .. x0 + n - 1 of row y (px, RGBA8
// dwords): a layer covers a pixel inside its placement; four covered pixels at a 16-byte aligned
// layer address move as one 16-byte load.
MMR_DEV void composite_layers(const CompositeJobDev& j, uint32_t y, uint32_t x0, uint32_t n, uint32_t px[4]) {
    for (uint32_t li = 0; li < j.layer_count; ++li) {
        const mmr_recon_layer& L = j.layers[li];
        if (y < L.y || y >= L.y + L.height) continue;   // uniform over the wave (one row)
        const MMR_GLOBAL uint8_t* row = as_global(L.rgba) + (size_t)(y - L.y) * L.stride;
        uint32_t sp[4] = {0, 0, 0, 0};
        uint32_t covered = 0;
        if (n == 4 && x0 >= L.x && x0 + 4 <= L.x + L.width && (((uintptr_t)(row + (size_t)(x0 - L.x) * 4)) & 15) == 0) {
            const mmr_v4i32 v = *reinterpret_cast<const MMR_GLOBAL mmr_v4i32*>(row + (size_t)(x0 - L.x) * 4);
#pragma unroll
            for (int i = 0; i < 4; ++i) sp[i] = (uint32_t)v[i];
            covered = 15;
        } else {
            for (uint32_t i = 0; i < n; ++i) {
                const uint32_t x = x0 + i;
                if (x >= L.x && x < L.x + L.width) {
                    sp[i] = *reinterpret_cast<const MMR_GLOBAL uint32_t*>(row + (size_t)(x - L.x) * 4);
                    covered |= 1u << i;
                }
            }
        }
        const bool premultiplied = L.op == MMR_BLEND_PREMULTIPLIED_OVER, opaque = L.opacity == 255;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            if (!((covered >> i) & 1u)) continue;
            uint32_t d[4] = {px[i] & 255, (px[i] >> 8) & 255, (px[i] >> 16) & 255, px[i] >> 24};
            const uint32_t s[4] = {sp[i] & 255, (sp[i] >> 8) & 255, (sp[i] >> 16) & 255, sp[i] >> 24};
            if (premultiplied) premultiplied_over(d, s, L.opacity, opaque);
            else blend_rgba_pixel(d, s);
            px[i] = d[0] | (d[1] << 8) | (d[2] << 16) | (d[3] << 24);
        }
    }
}

// One lane per 4 consecutive target pixels, a wave per 256 pixels of one row (tile -> job table);
// the layers covering each pixel are applied in order. A group of 4 pixels inside a layer, at
// 16-byte aligned addresses, moves as one 16-byte load; the target as one 16-byte load and store.
__global__ __launch_bounds__(256) void composite_kernel(const CompositeJobDev* __restrict__ jobs, const uint32_t* __restrict__ tile_job,
                                                        uint32_t tile_base, uint32_t tiles, uint32_t per_xcd) {
    const uint32_t local = xcd_tile(blockIdx.x, per_xcd);
    if (local >= tiles) return;
    const uint32_t tile = tile_base + local;
    const CompositeJobDev& j = jobs[tile_job[tile]];
    const uint32_t wave = (uint32_t)__builtin_amdgcn_readfirstlane((int)((tile - j.tile_begin) * 4 + (threadIdx.x >> 6)));
    const uint32_t y = wave / j.waves_per_row;
    const uint32_t x0 = (wave - y * j.waves_per_row) * 256 + (threadIdx.x & 63) * 4;
    if (y >= j.height || x0 >= j.width) return;
    const uint32_t n = min(4u, j.width - x0);
    MMR_GLOBAL uint32_t* t = reinterpret_cast<MMR_GLOBAL uint32_t*>(as_global(j.target) + (size_t)y * j.stride + (size_t)x0 * 4);
    const bool tvec = n == 4 && ((uintptr_t)t & 15) == 0;
    uint32_t px[4] = {0, 0, 0, 0};
    if (tvec) {
        const mmr_v4i32 v = *reinterpret_cast<const MMR_GLOBAL mmr_v4i32*>(t);
#pragma unroll
        for (int i = 0; i < 4; ++i) px[i] = (uint32_t)v[i];
    } else {
        for (uint32_t i = 0; i < n; ++i) px[i] = t[i];
    }
    composite_layers(j, y, x0, n, px);
    if (tvec) {
        mmr_v4i32 v;
#pragma unroll
        for (int i = 0; i < 4; ++i) v[i] = (int)px[i];
        *reinterpret_cast<MMR_GLOBAL mmr_v4i32*>(t) = v;
    } else {
        for (uint32_t i = 0; i < n; ++i) t[i] = px[i];
    }
}

// Conversion and composite in one pass: job i converts rjobs[i]'s planes (yuv_px4, alpha 255)
// into the pixels of cjobs[i]'s target (the conversion's RGBA output, rows packed), applies the
// layers (composite_layers) and stores each target pixel once -- the result of the conversion
// followed by the composite, without writing the converted frame and reading it back. Launch
// shape of composite_kernel (a lane per 4 pixels, a wave per 256 pixels of one row).
__global__ __launch_bounds__(256) void yuv_composite_kernel(const RgbJob* __restrict__ rjobs, const CompositeJobDev* __restrict__ cjobs,
                                                            const uint32_t* __restrict__ tile_job, uint32_t tile_base, uint32_t tiles,
                                                            uint32_t per_xcd) {
    const uint32_t local = xcd_tile(blockIdx.x, per_xcd);
    if (local >= tiles) return;
    const uint32_t tile = tile_base + local;
    const uint32_t ji = tile_job[tile];
    const CompositeJobDev& j = cjobs[ji];
    const RgbJob& r = rjobs[ji];
    const uint32_t wave = (uint32_t)__builtin_amdgcn_readfirstlane((int)((tile - j.tile_begin) * 4 + (threadIdx.x >> 6)));
    const uint32_t y = wave / j.waves_per_row;
    const uint32_t x0 = (wave - y * j.waves_per_row) * 256 + (threadIdx.x & 63) * 4;
    if (y >= j.height || x0 >= j.width) return;
    float f[4][3];
    const uint32_t n = r.limited ? yuv_px4<true>(r, y, x0, f) : yuv_px4<false>(r, y, x0, f);
    uint32_t px[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) px[i] = put_byte(put_byte(put_byte(0xff000000u, f[i][0], 0), f[i][1], 1), f[i][2], 2);
    composite_layers(j, y, x0, n, px);
    MMR_GLOBAL uint32_t* t = reinterpret_cast<MMR_GLOBAL uint32_t*>(as_global(j.target) + (size_t)y * j.stride + (size_t)x0 * 4);
    if (n == 4 && ((uintptr_t)t & 15) == 0) {
        mmr_v4i32 v;
#pragma unroll
        for (int i = 0; i < 4; ++i) v[i] = (int)px[i];
        *reinterpret_cast<MMR_GLOBAL mmr_v4i32*>(t) = v;
    } else {
        for (uint32_t i = 0; i < n; ++i) t[i] = px[i];
    }
}

MMR_DEV uint8_t blend_channel(uint32_t d, uint32_t s, uint32_t a) {
    if (a == 0) return (uint8_t)d;
    if (a == 255) return (uint8_t)s;
    return (uint8_t)((s * a + d * (255 - a) + 127) / 255);
}

// grid.y = job; one thread per overlay luma pixel; the pixel at even (x, y) also blends the
// co-located chroma sample.
__global__ __launch_bounds__(256) void yuv_blend_kernel(const mmr_recon_yuv_blend_job* __restrict__ jobs) {
    const mmr_recon_yuv_blend_job& j = jobs[blockIdx.y];
    const uint32_t p = blockIdx.x * 256 + threadIdx.x;
    if (p >= j.width * j.height) return;
    const uint32_t y = p / j.width, x = p % j.width;
    uint8_t* dy = j.target[0] + (size_t)(j.y + y) * j.target_stride[0] + j.x + x;
    *dy = blend_channel(*dy, j.overlay[0][p], j.rgba[(size_t)p * 4 + 3]);
    if (((x | y) & 1) == 0) {
        const uint32_t cw = j.width / 2, s = (y / 2) * cw + x / 2;
        const uint32_t a = j.chroma_alpha[s];
        uint8_t* du = j.target[1] + (size_t)(j.y / 2 + y / 2) * j.target_stride[1] + j.x / 2 + x / 2;
        uint8_t* dv = j.target[2] + (size_t)(j.y / 2 + y / 2) * j.target_stride[2] + j.x / 2 + x / 2;
        *du = blend_channel(*du, j.overlay[1][s], a);
        *dv = blend_channel(*dv, j.overlay[2][s], a);
    }
}

// rgb_to_yuv (crates/render/src/project_compositor.rs:1163-1185): BT.709 limited (ly, lu, lv) and
// full range (fy, fu, fv).
MMR_DEV void rgb_to_yuv(int r, int g, int b, uint32_t lim[3], uint32_t full[3]) {
    lim[0] = (uint32_t)clampi(16 + ((47 * r + 157 * g + 16 * b + 128) >> 8), 16, 235);
    lim[1] = (uint32_t)clampi(128 + ((-26 * r - 87 * g + 112 * b + 128) >> 8), 16, 240);
    lim[2] = (uint32_t)clampi(128 + ((112 * r - 102 * g - 10 * b + 128) >> 8), 16, 240);
    full[0] = (uint32_t)clampi((54 * r + 183 * g + 18 * b + 128) >> 8, 0, 255);
    full[1] = (uint32_t)clampi(128 + ((-29 * r - 99 * g + 128 * b + 128) >> 8), 0, 255);
    full[2] = (uint32_t)clampi(128 + ((128 * r - 116 * g - 12 * b + 128) >> 8), 0, 255);
}

// PreparedOverlay::prepare_yuv (project_compositor.rs:1098-1142): one thread per 2x2 cell (the
// last row / column of an odd size alone): luma of its pixels, then the rounded mean alpha and the
// alpha-weighted mean chroma of both ranges (128 for a transparent cell).
__global__ __launch_bounds__(256) void overlay_prepare_kernel(const mmr_recon_overlay_prepare_job* __restrict__ jobs) {
    const mmr_recon_overlay_prepare_job& j = jobs[blockIdx.y];
    const uint32_t cw = (j.width + 1) / 2, ch = (j.height + 1) / 2;
    const uint32_t c = blockIdx.x * 256 + threadIdx.x;
    if (c >= cw * ch) return;
    const uint32_t cy = c / cw, cx = c - cy * cw, x0 = cx * 2, y0 = cy * 2;
    const uint32_t nw = min(2u, j.width - x0), nh = min(2u, j.height - y0);
    uint32_t alpha_sum = 0, lu = 0, lv = 0, fu = 0, fv = 0;
    for (uint32_t dy = 0; dy < nh; ++dy)
        for (uint32_t dx = 0; dx < nw; ++dx) {
            const size_t p = (size_t)(y0 + dy) * j.width + x0 + dx;
            const uint32_t px = *reinterpret_cast<const uint32_t*>(j.rgba + p * 4);   // rgba is 4-byte aligned
            uint32_t lim[3], full[3];
            rgb_to_yuv((int)(px & 0xff), (int)((px >> 8) & 0xff), (int)((px >> 16) & 0xff), lim, full);
            const uint32_t a = px >> 24;
            if (j.y_limited) j.y_limited[p] = (uint8_t)lim[0];
            if (j.y_full) j.y_full[p] = (uint8_t)full[0];
            alpha_sum += a;
            lu += lim[1] * a;
            lv += lim[2] * a;
            fu += full[1] * a;
            fv += full[2] * a;
        }
    const uint32_t n = nw * nh;
    auto weighted = [&](uint32_t sum) -> uint8_t { return (uint8_t)(alpha_sum ? (sum + alpha_sum / 2) / alpha_sum : 128u); };
    if (j.chroma_alpha) j.chroma_alpha[c] = (uint8_t)((alpha_sum + n / 2) / n);
    if (j.u_limited) j.u_limited[c] = weighted(lu);
    if (j.v_limited) j.v_limited[c] = weighted(lv);
    if (j.u_full) j.u_full[c] = weighted(fu);
    if (j.v_full) j.v_full[c] = weighted(fv);
}

// Lanczos3 resize (image::imageops::resize of the image crate 0.25.6, as scale_yuv420_to_canvas_cpu
// calls it, crates/render/src/compositor.rs:89-141), vertical pass: f32 sums, in tap order, of the
// source samples times the row's normalised weights (taps from the host), for the scaled rows the
// canvas window keeps. Thread (x, r) of plane blockIdx.y.
__global__ __launch_bounds__(256) void scale_vertical_kernel(const ScalePlaneDev* __restrict__ planes) {
    const ScalePlaneDev& p = planes[blockIdx.y];
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (p.tiled || !p.resize || i >= p.src_w * p.copy_h) return;
    const uint32_t r = i / p.src_w, x = i - r * p.src_w;
    const int32_t left = p.vleft[r];
    const uint32_t count = p.vcount[r];
    const float* w = p.vw + (size_t)r * p.vmax;
    const uint8_t* src = p.src + (size_t)left * p.src_stride + x;
    float t = 0.0f;
    for (uint32_t k = 0; k < count; ++k) t += (float)src[(size_t)k * p.src_stride] * w[k];
    p.tmp[i] = t;
}

// Horizontal pass and placement (center_crop_or_pad, compositor.rs:277-305): thread (x, y) of the
// canvas plane writes the background outside the window; inside, the horizontal f32 sum of the
// intermediate row clamped to [0, 255] and rounded half away from zero, or the source sample when
// the plane keeps its size.
__global__ __launch_bounds__(256) void scale_output_kernel(const ScalePlaneDev* __restrict__ planes) {
    const ScalePlaneDev& p = planes[blockIdx.y];
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (p.tiled || i >= p.dst_w * p.dst_h) return;
    const uint32_t y = i / p.dst_w, x = i - y * p.dst_w;
    uint32_t v = p.background;
    if (x >= p.dst_x && x < p.dst_x + p.copy_w && y >= p.dst_y && y < p.dst_y + p.copy_h) {
        const uint32_t c = x - p.dst_x, r = y - p.dst_y;
        if (p.resize) {
            const int32_t left = p.hleft[c];
            const uint32_t count = p.hcount[c];
            const float* w = p.hw + (size_t)c * p.hmax;
            const float* row = p.tmp + (size_t)r * p.src_w + left;
            float t = 0.0f;
            for (uint32_t k = 0; k < count; ++k) t += row[k] * w[k];
            t = t < 0.0f ? 0.0f : (t > 255.0f ? 255.0f : t);
            v = (uint32_t)roundf(t);
        } else {
            v = p.src[(size_t)(p.sy + r) * p.src_stride + p.sx + c];
        }
    }
    p.dst[(size_t)y * p.dst_stride + x] = (uint8_t)v;
}

// One-pass scaler: workgroup per kTileW x kTileH canvas tile of plane blockIdx.y. The vertical
// pass of the tile's window rows over the source columns its outputs reach goes to LDS (source
// read straight from global memory, the overlapping taps of neighbouring rows served by the
// caches), then the horizontal pass, clamp, rounding and placement -- the same f32 operations in
// the same order as the two-pass kernels, without the intermediate's round trip through HBM.
__global__ __launch_bounds__(256) void scale_tile_kernel(const ScalePlaneDev* __restrict__ planes) {
    const ScalePlaneDev& p = planes[blockIdx.y];
    if (!p.tiled) return;
    const uint32_t tiles_x = (p.dst_w + kScaleTileW - 1) / kScaleTileW;
    if (blockIdx.x >= tiles_x * ((p.dst_h + kScaleTileH - 1) / kScaleTileH)) return;
    __shared__ float tmp[kScaleTileH * kScaleTileSrc];
    __shared__ float hw_s[kScaleTileTaps][kScaleTileW];   // the tile columns' weights, tap-major
    __shared__ int32_t hl_s[kScaleTileW];                 // first tap, relative to the tile's source column 0
    __shared__ uint32_t hc_s[kScaleTileW];
    const int tx0 = (int)((blockIdx.x % tiles_x) * kScaleTileW), ty0 = (int)((blockIdx.x / tiles_x) * kScaleTileH);
    const int tx1 = min(tx0 + kScaleTileW, (int)p.dst_w), ty1 = min(ty0 + kScaleTileH, (int)p.dst_h);
    // The tile's part of the window, in window coordinates (empty when wx0 >= wx1 or wy0 >= wy1).
    const int wx0 = max(tx0, (int)p.dst_x) - (int)p.dst_x, wx1 = min(tx1, (int)(p.dst_x + p.copy_w)) - (int)p.dst_x;
    const int wy0 = max(ty0, (int)p.dst_y) - (int)p.dst_y, wy1 = min(ty1, (int)(p.dst_y + p.copy_h)) - (int)p.dst_y;
    const bool inside = wx0 < wx1 && wy0 < wy1;
    int sc0 = 0;
    if (inside && p.resize) {   // uniform per workgroup
        sc0 = p.hleft[wx0];
        const int sw = p.hleft[wx1 - 1] + (int)p.hcount[wx1 - 1] - sc0;   // <= kScaleTileSrc (host check)
        const int rows = wy1 - wy0, cols = wx1 - wx0, hmax = (int)p.hmax;   // hmax <= kScaleTileTaps (host check)
        for (int i = (int)threadIdx.x; i < cols * hmax; i += 256) {
            const int c = i / hmax, k = i - c * hmax;
            hw_s[k][c] = p.hw[(size_t)(wx0 + c) * hmax + k];
        }
        for (int i = (int)threadIdx.x; i < cols; i += 256) {
            hl_s[i] = p.hleft[wx0 + i] - sc0;
            hc_s[i] = p.hcount[wx0 + i];
        }
        for (int i = (int)threadIdx.x; i < rows * sw; i += 256) {
            const int r = i / sw, sc = i - r * sw;
            const int y = wy0 + r;
            const uint32_t count = p.vcount[y];
            const float* w = p.vw + (size_t)y * p.vmax;
            const uint8_t* src = p.src + (size_t)p.vleft[y] * p.src_stride + sc0 + sc;
            const size_t st = p.src_stride;
            float t = 0.0f;
            uint32_t k = 0;
            // four taps' loads in flight at once; the sum still takes them one by one in order
            for (; k + 4 <= count; k += 4) {
                const uint32_t a0 = src[k * st], a1 = src[(k + 1) * st], a2 = src[(k + 2) * st], a3 = src[(k + 3) * st];
                t += (float)a0 * w[k];
                t += (float)a1 * w[k + 1];
                t += (float)a2 * w[k + 2];
                t += (float)a3 * w[k + 3];
            }
            for (; k < count; ++k) t += (float)src[k * st] * w[k];
            tmp[r * kScaleTileSrc + sc] = t;
        }
        __syncthreads();
    }
    for (int i = (int)threadIdx.x; i < kScaleTileW * kScaleTileH; i += 256) {
        const int x = tx0 + (i % kScaleTileW), y = ty0 + (i / kScaleTileW);
        if (x >= tx1 || y >= ty1) continue;
        const int c = x - (int)p.dst_x, r = y - (int)p.dst_y;
        uint32_t v = p.background;
        if (inside && c >= wx0 && c < wx1 && r >= wy0 && r < wy1) {
            if (p.resize) {
                const int cc = c - wx0;
                const uint32_t count = hc_s[cc];
                const float* row = tmp + (r - wy0) * kScaleTileSrc + hl_s[cc];
                float t = 0.0f;
                uint32_t k = 0;
                for (; k + 4 <= count; k += 4) {
                    const float r0 = row[k], r1 = row[k + 1], r2 = row[k + 2], r3 = row[k + 3];
                    t += r0 * hw_s[k][cc];
                    t += r1 * hw_s[k + 1][cc];
                    t += r2 * hw_s[k + 2][cc];
                    t += r3 * hw_s[k + 3][cc];
                }
                for (; k < count; ++k) t += row[k] * hw_s[k][cc];
                t = t < 0.0f ? 0.0f : (t > 255.0f ? 255.0f : t);
                v = (uint32_t)roundf(t);
            } else {
                v = p.src[(size_t)(p.sy + r) * p.src_stride + p.sx + c];
            }
        }
        p.dst[(size_t)y * p.dst_stride + x] = (uint8_t)v;
    }
}

__global__ void fill_zero_kernel(uint32_t* p, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

hipError_t launch_yuv_to_rgb(const RgbJob* d_jobs, const uint32_t* d_tile_job, uint32_t tile_base, uint32_t tiles, hipStream_t s) {
    if (tiles == 0) return hipSuccess;
    const uint32_t per = xcd_per(tiles);
    hipLaunchKernelGGL(yuv_to_rgb_kernel, dim3(per * kXcds), dim3(256), 0, s, d_jobs, d_tile_job, tile_base, tiles, per);
    return hipGetLastError();
}

hipError_t launch_linear_composite(const LinearCompositeDev* d_jobs, const uint32_t* d_tile_job, uint32_t tile_base, uint32_t tiles,
                                  const uint16_t* d_to_linear, hipStream_t s) {
    if (tiles == 0) return hipSuccess;
    const uint32_t per = xcd_per(tiles);
    hipLaunchKernelGGL(linear_composite_kernel, dim3(per * kXcds), dim3(256), 0, s, d_jobs, d_tile_job, tile_base, tiles, d_to_linear,
                       per);
    return hipGetLastError();
}

hipError_t launch_linear_convert(const LinearConvertDev* d_jobs, const uint32_t* d_tile_job, uint32_t tile_base, uint32_t tiles,
                                const uint8_t* d_tables, hipStream_t s) {
    if (tiles == 0) return hipSuccess;
    const uint32_t per = xcd_per(tiles);
    hipLaunchKernelGGL(linear_convert_kernel, dim3(per * kXcds), dim3(256), 0, s, d_jobs, d_tile_job, tile_base, tiles, d_tables, per);
    return hipGetLastError();
}

hipError_t launch_composite(const CompositeJobDev* d_jobs, const uint32_t* d_tile_job, uint32_t tile_base, uint32_t tiles, hipStream_t s) {
    if (tiles == 0) return hipSuccess;
    const uint32_t per = xcd_per(tiles);
    hipLaunchKernelGGL(composite_kernel, dim3(per * kXcds), dim3(256), 0, s, d_jobs, d_tile_job, tile_base, tiles, per);
    return hipGetLastError();
}

hipError_t launch_yuv_composite(const RgbJob* d_rjobs, const CompositeJobDev* d_cjobs, const uint32_t* d_tile_job, uint32_t tile_base,
                                uint32_t tiles, hipStream_t s) {
    if (tiles == 0) return hipSuccess;
    const uint32_t per = xcd_per(tiles);
    hipLaunchKernelGGL(yuv_composite_kernel, dim3(per * kXcds), dim3(256), 0, s, d_rjobs, d_cjobs, d_tile_job, tile_base, tiles, per);
    return hipGetLastError();
}

hipError_t launch_yuv_blend(const mmr_recon_yuv_blend_job* d_jobs, uint32_t job_count, uint32_t max_px, hipStream_t s) {
    if (max_px == 0 || job_count == 0) return hipSuccess;
    hipLaunchKernelGGL(yuv_blend_kernel, dim3((max_px + 255) / 256, job_count), dim3(256), 0, s, d_jobs);
    return hipGetLastError();
}

hipError_t launch_overlay_prepare(const mmr_recon_overlay_prepare_job* d_jobs, uint32_t job_count, uint32_t max_cells, hipStream_t s) {
    if (max_cells == 0 || job_count == 0) return hipSuccess;
    hipLaunchKernelGGL(overlay_prepare_kernel, dim3((max_cells + 255) / 256, job_count), dim3(256), 0, s, d_jobs);
    return hipGetLastError();
}

hipError_t launch_scale(const ScalePlaneDev* d_planes, uint32_t plane_count, uint32_t max_tiles, uint32_t max_vertical,
                        uint32_t max_output, hipStream_t s) {
    if (plane_count == 0) return hipSuccess;
    if (max_tiles) hipLaunchKernelGGL(scale_tile_kernel, dim3(max_tiles, plane_count), dim3(256), 0, s, d_planes);
    if (max_vertical) hipLaunchKernelGGL(scale_vertical_kernel, dim3((max_vertical + 255) / 256, plane_count), dim3(256), 0, s, d_planes);
    if (max_output) hipLaunchKernelGGL(scale_output_kernel, dim3((max_output + 255) / 256, plane_count), dim3(256), 0, s, d_planes);
    return hipGetLastError();
}

hipError_t launch_fill_zero(uint32_t* d_words, uint32_t count, hipStream_t s) {
    if (count == 0) return hipSuccess;
    uint32_t blocks = (count + 255) / 256;
    if (blocks > 1024) blocks = 1024;
    hipLaunchKernelGGL(fill_zero_kernel, dim3(blocks), dim3(256), 0, s, d_words, count);
    return hipGetLastError();
}

}  // namespace mmr
