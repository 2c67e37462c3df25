// Internal job descriptors (device-resident, built by the host code of recon_host.cpp) and the
// kernel launchers. Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mmr_recon.h"

namespace mmr {

// ---------------------------------------------------------------- MJPEG
struct MjpegJob {
    const int16_t* levels;
    const uint16_t* quant;        // [n][64] zigzag
    uint8_t* planes[4];
    uint32_t plane_w[4];
    uint32_t plane_h[4];
    uint32_t mcu_cols;
    uint32_t blocks_per_mcu;      // 1 when not interleaved
    uint32_t tile_begin;          // first entry of this frame in the launch's tile -> job table
    uint32_t block_count;
    uint8_t blk_comp[16];         // per block-in-MCU: component, column, row
    uint8_t blk_bx[16];
    uint8_t blk_by[16];
    uint8_t comp_h[4];            // blocks per MCU row / column of each component (1,1 when not interleaved)
    uint8_t comp_v[4];
    uint8_t qsel[4];
};

// ---------------------------------------------------------------- MPEG-2
// One plane of an MPEG-2 picture with its reference planes; a lane loads the record of its
// block's plane with two 16-byte loads.
struct alignas(16) Mpeg2Plane {
    uint8_t* dst;
    const uint8_t* fwd;           // the plane itself when the picture has no forward reference
    const uint8_t* bwd;
    uint32_t stride;
    uint32_t height;
};
static_assert(sizeof(Mpeg2Plane) == 32, "plane record is two 16-byte loads");

struct Mpeg2Job {
    Mpeg2Plane plane[3];          // Y, Cb, Cr
    const uint8_t* matrices;      // [4][64] natural
    uint32_t* order;              // [6 * macroblocks][4] coded-first block entries (submission scratch)
    const mmr_recon_mpeg2_macroblock* mbs;
    const int16_t* levels;
    const mmr_recon_mpeg2_field_motion* field_motion;   // [macroblocks] or NULL
    int32_t fwd_src;
    int32_t bwd_src;
    uint32_t mb_w;
    uint32_t mb_h;
    uint32_t alt_scan;
    uint32_t dc_shift;            // 3 - intra_dc_precision
    uint32_t tile_begin;          // first entry of this picture in the submission's tile -> job table
    uint32_t tile_count;
    uint32_t refs_aligned;        // every reference plane pointer is 4-byte aligned
    uint32_t row_begin;           // first macroblock row of this picture in the order pass's row space
};

// ---------------------------------------------------------------- DV
struct DvJob {
    const int16_t* levels;
    const uint8_t* info;
    uint8_t* planes[3];
    uint32_t system;
    uint32_t tile_begin;          // first entry of this frame in the submission's tile -> job table
    uint32_t block_count;
    uint32_t pad;
};

// ---------------------------------------------------------------- H.264
struct H264Job {
    const mmr_recon_h264_macroblock* mbs;
    const mmr_recon_h264_block* blocks;
    const int16_t* levels;
    const mmr_recon_h264_weight* weights;
    const mmr_recon_h264_bi_weight* bi_weights;
    const uint8_t* scaling;       // never NULL on device (flat table substituted)
    uint8_t* planes[3];
    const uint8_t* ref_planes[2][MMR_H264_MAX_REFS][3];
    int32_t ref_poc[2][MMR_H264_MAX_REFS];
    uint32_t ref_count[2];
    uint32_t width_mbs;
    uint32_t height_mbs;
    uint32_t weighted;
    uint32_t deblock;
    int32_t chroma_qp_offset[2];
    uint32_t tile_begin;          // first entry of this picture in the tile -> job table (4 macroblocks per tile)
    uint32_t refs_aligned;        // every reference plane is 4-byte aligned
    uint32_t band_begin;          // first deblocking band ticket of this picture
    uint32_t pad;
    uint32_t* bs;                 // [macroblocks][4] packed boundary strengths (submission scratch)
    const mmr_recon_h264_intra* intra;   // [macroblocks] or NULL (intra macroblocks unsupported)
    uint32_t intra_band_begin;    // first intra band ticket of this picture (bands only with intra records)
    uint32_t transform_bypass;    // qpprime_y_zero_transform_bypass_flag
};

// ---------------------------------------------------------------- colour / blend
struct RgbJob {
    const uint8_t* planes[3];
    uint8_t* rgb;
    uint8_t* rgba;
    uint32_t width, height, cw, ch, luma_stride, chroma_stride, range, matrix;
    uint32_t tile_begin;          // first workgroup tile of this job in the launch's tile -> job table
    uint32_t waves_per_row;       // ceil(width / 256): a wave covers 256 pixels of one row
    double inv_width, inv_height; // 1.0 / width, 1.0 / height (chroma coordinate estimates)
    // Chroma terms X_c = kcb[c] * Cb + kcr[c] * Cr + k0[c] (c = R, G, B), each the integer form's
    // (A_c (Cb - 128) + B_c (Cr - 128) + 128) / 256 as an exact binary32 value; set_rgb_terms().
    float kcb[3], kcr[3], k0[3];
    uint32_t limited;             // range == MMR_RANGE_LIMITED
    uint32_t half_chroma;         // 2 cw == width: pixels 2k and 2k + 1 share chroma sample k
    uint32_t flat;                // segments run through the pixels in raster order (width % 4 == 0)
    uint32_t segments;            // 256-pixel segments of the job (one wave pass, 4 pixels per lane)
};
// Segments of the YUV->RGB kernel: 256 pixels, 4 per lane of a wave. When the width is a multiple
// of 4 (and the picture below 2^31 pixels) they run through the pixels in raster order across rows
// (every segment but the last full); otherwise each row has ceil(width / 256) segments. A wave
// converts kRgbSegmentsPerWave consecutive segments and a workgroup tile holds 4 waves, never
// straddling two jobs.
constexpr uint32_t kRgbSegmentsPerWave = 4;
inline void set_rgb_layout(RgbJob& j) {
    j.waves_per_row = (j.width + 255) / 256;
    const uint64_t px = (uint64_t)j.width * j.height;
    j.flat = j.width % 4 == 0 && px < (1ull << 31) ? 1u : 0u;
    j.segments = j.flat ? (uint32_t)((px + 255) / 256) : j.waves_per_row * j.height;
}
inline uint32_t rgb_job_tiles(const RgbJob& j) { return (j.segments + 4 * kRgbSegmentsPerWave - 1) / (4 * kRgbSegmentsPerWave); }

// Fills the chroma-term constants, limited and half_chroma of j from range, matrix, width and cw.
// Integer forms (b = Cb - 128, r = Cr - 128, yy = max(Y - 16, 0)), ycbcr_to_rgb
// (crates/cli/src/terminal_preview.rs:7159-7238) for BT.601:
//   limited  R = (298 yy + 409 r + 128) >> 8, G = (298 yy - 100 b - 208 r + 128) >> 8,
//            B = (298 yy + 516 b + 128) >> 8            (BT.709: 459; 55, 136; 541)
//   full     R = Y + ((359 r + 128) >> 8), G = Y - ((88 b + 183 r + 128) >> 8),
//            B = Y + ((454 b + 128) >> 8)                (BT.709: 403; 48, 120; 475)
// With X = (A b + B r + 128) / 256 = (A / 256) Cb + (B / 256) Cr + 1/2 - (A + B) / 2, limited
// channels are floor(298 / 256 yy + X_c) (X_G carrying the minus signs) and full-range ones
// Y + floor(X_R), Y - floor(X_G), Y + floor(X_B). Every constant is a multiple of 1/256 below
// 2^10 and every partial sum one of magnitude below 2^10, so binary32 holds them all exactly.
inline void set_rgb_terms(RgbJob& j) {
    const bool lim = j.range == MMR_RANGE_LIMITED, bt709 = j.matrix == MMR_MATRIX_BT709;
    int A[3], B[3];
    if (lim) {
        A[0] = 0;                 B[0] = bt709 ? 459 : 409;
        A[1] = bt709 ? -55 : -100; B[1] = bt709 ? -136 : -208;
        A[2] = bt709 ? 541 : 516; B[2] = 0;
    } else {
        A[0] = 0;                 B[0] = bt709 ? 403 : 359;
        A[1] = bt709 ? 48 : 88;   B[1] = bt709 ? 120 : 183;
        A[2] = bt709 ? 475 : 454; B[2] = 0;
    }
    for (int c = 0; c < 3; ++c) {
        j.kcb[c] = (float)A[c] / 256.0f;
        j.kcr[c] = (float)B[c] / 256.0f;
        j.k0[c] = 0.5f - (float)(A[c] + B[c]) / 2.0f;
    }
    j.limited = lim ? 1u : 0u;
    j.half_chroma = (uint64_t)j.cw * 2 == j.width ? 1u : 0u;
}

// MMFX linear-light jobs (u16 premultiplied RGBA): one lane per pixel, one wave per 64 pixels of
// a row, four waves per workgroup tile; job j has ceil(waves_per_row * height / 4) tiles.
struct LinearCompositeDev {
    uint16_t* target;
    uint32_t width, height, stride, layer_count;
    uint32_t tile_begin, waves_per_row;
    mmr_recon_linear_layer layers[MMR_RECON_MAX_LAYERS];
};
struct LinearConvertDev {
    uint16_t* linear;
    uint8_t* rgba;
    uint32_t direction, width, height, linear_stride, rgba_stride;
    uint32_t tile_begin, waves_per_row, pad;
};
inline uint32_t linear_job_tiles(uint32_t waves_per_row, uint32_t height) { return (waves_per_row * height + 3) / 4; }
// Device copy of srgb_to_linear_u16 (256 x u16) followed by linear_u16_to_srgb (65536 x u8).
constexpr size_t kLinearTableBytes = 256 * 2 + 65536;

struct CompositeJobDev {
    uint8_t* target;
    uint32_t width, height, stride, layer_count;
    uint32_t tile_begin;          // first workgroup tile of this job in the launch's tile -> job table
    uint32_t waves_per_row;       // ceil(width / 256): a wave covers 256 pixels of one row
    mmr_recon_layer layers[MMR_RECON_MAX_LAYERS];
};
inline uint32_t composite_job_tiles(uint32_t waves_per_row, uint32_t height) { return (waves_per_row * height + 3) / 4; }

// Kernel launchers (stream-ordered). Each returns hipSuccess or the launch error.
// The block kernels (MJPEG, MPEG-2, DV) run one workgroup per tile of group_blocks blocks (16, 32
// or 64): workgroup i takes tile tile_base + i (XCD-remapped), whose job index is
// d_tile_job[tile]. Tiles never straddle jobs.
hipError_t launch_mjpeg(const MjpegJob* d_jobs, const uint32_t* d_tile_job, uint32_t tile_base, uint32_t tiles, uint32_t group_blocks,
                        hipStream_t s);
hipError_t launch_mpeg2_order(const Mpeg2Job* d_jobs, uint32_t job_count, uint32_t total_rows, uint32_t* d_status,
                              hipStream_t s);
hipError_t launch_mpeg2(const Mpeg2Job* d_jobs, const uint32_t* d_tile_job, uint32_t tile_base, uint32_t tiles,
                        uint32_t group_blocks, uint32_t* d_status, hipStream_t s);
hipError_t launch_dv(const DvJob* d_jobs, const uint32_t* d_tile_job, uint32_t tile_base, uint32_t tiles, uint32_t group_blocks,
                     hipStream_t s);
hipError_t dv_upload_tables(hipStream_t s);
hipError_t launch_h264_mc(const H264Job* d_jobs, const uint32_t* d_tile_job, uint32_t tile_base, uint32_t tiles, uint32_t* d_status,
                          hipStream_t s);
hipError_t launch_h264_intra(const H264Job* d_jobs, uint32_t job_count, uint32_t total_bands, uint32_t* d_counters, uint32_t groups,
                             uint32_t publish, uint32_t* d_status, hipStream_t s);
uint32_t h264_intra_band_rows();   // macroblock rows per intra band (one workgroup)
// Deblocking in bands of `band_rows` (4, 8 or 16) macroblock rows: d_counters[0] is the band
// ticket, d_counters[1 + t] the progress of band t's last row (zeroed before the launch), published
// every `publish` macroblocks.
bool h264_deblock_rows_supported(uint32_t rows);
// Boundary strengths of every macroblock of the deblocked pictures into job.bs (MC tiles).
hipError_t launch_h264_bs(const H264Job* d_jobs, const uint32_t* d_tile_job, uint32_t tile_base, uint32_t tiles, hipStream_t s);
hipError_t launch_h264_deblock(const H264Job* d_jobs, uint32_t job_count, uint32_t total_bands, uint32_t* d_counters,
                               uint32_t groups, uint32_t band_rows, uint32_t publish, uint32_t* d_status, hipStream_t s);
hipError_t launch_linear_composite(const LinearCompositeDev* d_jobs, const uint32_t* d_tile_job, uint32_t tile_base, uint32_t tiles,
                                  const uint16_t* d_to_linear, hipStream_t s);
hipError_t launch_linear_convert(const LinearConvertDev* d_jobs, const uint32_t* d_tile_job, uint32_t tile_base, uint32_t tiles,
                                const uint8_t* d_tables, hipStream_t s);
hipError_t launch_yuv_to_rgb(const RgbJob* d_jobs, const uint32_t* d_tile_job, uint32_t tile_base, uint32_t tiles, hipStream_t s);
hipError_t launch_composite(const CompositeJobDev* d_jobs, const uint32_t* d_tile_job, uint32_t tile_base, uint32_t tiles, hipStream_t s);
// Conversion rjobs[i] and composite cjobs[i] in one pass over the composite's tiles (job i's
// target is conversion i's RGBA output, rows packed): the result of launch_yuv_to_rgb followed by
// launch_composite.
hipError_t launch_yuv_composite(const RgbJob* d_rjobs, const CompositeJobDev* d_cjobs, const uint32_t* d_tile_job, uint32_t tile_base,
                                uint32_t tiles, hipStream_t s);
hipError_t launch_yuv_blend(const mmr_recon_yuv_blend_job* d_jobs, uint32_t job_count, uint32_t max_px, hipStream_t s);

// One plane of a scale job on the device: the canvas window (copy_w x copy_h at dst_x, dst_y)
// shows the scaled plane's window at sx, sy; with resize, scaled rows / columns of the window have
// their taps (first source index, count, weights at stride vmax / hmax) and the vertical pass's
// f32 rows (src_w wide, one per window row) go to tmp; without resize sx, sy index the source.
struct ScalePlaneDev {
    const uint8_t* src;
    uint8_t* dst;
    float* tmp;
    const int32_t* vleft;
    const uint32_t* vcount;
    const float* vw;
    const int32_t* hleft;
    const uint32_t* hcount;
    const float* hw;
    uint32_t src_w, src_h, src_stride;
    uint32_t dst_w, dst_h, dst_stride;
    uint32_t copy_w, copy_h, dst_x, dst_y;
    uint32_t sx, sy;
    uint32_t vmax, hmax;
    uint32_t resize, background;
    uint32_t tiled;   // done by scale_tile_kernel (windows within kScaleTileSrc columns, hmax <= kScaleTileTaps)
    uint32_t pad;
};
// Canvas tiles of the one-pass scaler: kScaleTileW x kScaleTileH samples, the vertical pass of
// at most kScaleTileSrc source columns per tile and horizontal taps of at most kScaleTileTaps per
// column held in LDS.
constexpr int kScaleTileW = 64, kScaleTileH = 16, kScaleTileSrc = 320, kScaleTileTaps = 32;
hipError_t launch_overlay_prepare(const mmr_recon_overlay_prepare_job* d_jobs, uint32_t job_count, uint32_t max_cells, hipStream_t s);
hipError_t launch_scale(const ScalePlaneDev* d_planes, uint32_t plane_count, uint32_t max_tiles, uint32_t max_vertical,
                        uint32_t max_output, hipStream_t s);
hipError_t launch_fill_zero(uint32_t* d_words, uint32_t count, hipStream_t s);

// Bits of the per-submission device status word.
constexpr uint32_t kDevStatusUnsupported = 1u;
constexpr uint32_t kDevStatusInvalidData = 2u;
constexpr uint32_t kDevStatusTimeout = 4u;   // a bounded device wait gave up (internal error)

// Binary search over ascending begin offsets: index of the last job whose begin <= key.
template <typename Job, typename Get>
__device__ __forceinline__ uint32_t find_job(const Job* jobs, uint32_t count, uint32_t key, Get get) {
    uint32_t lo = 0, hi = count;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (get(jobs[mid]) <= key) lo = mid; else hi = mid;
    }
    return lo;
}

}  // namespace mmr
