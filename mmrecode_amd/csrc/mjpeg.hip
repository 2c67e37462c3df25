// MJPEG block reconstruction: dequantisation, zig-zag scatter, f64 IDCT, placement.
//
// Reference: decode_block's dequantisation (crates/codecs/mjpeg/src/decoder.rs:502-546),
// inverse_dct (crates/codecs/mjpeg/src/transform.rs:86-109), write_block (decoder.rs:550-565),
// MCU walk (decoder.rs:346-401).
//
// Layout: 8 lanes per 8x8 block, 32 blocks (one tile, never straddling frames) per 256-thread
// workgroup; a host-built tile -> frame table gives each workgroup its frame with one scalar
// load. Lane j loads scan positions 8j..8j+7 (16-byte levels, 16-byte quantisers, 8-byte natural
// positions), dequantises them and scatters to natural order in LDS. Lane v then takes natural
// row v for the IDCT row pass, lane x runs column x of the column pass (the fused form, with the
// reference order where it could round differently), and after an LDS
// transpose lane r stores row r of the block with one 8-byte store (write_block's right / bottom
// clipping handled per row). The lanes of a block share a wave, so the exchanges need no
// workgroup barrier.
#include "device_common.h"
#include "recon_internal.h"

namespace mmr {

// BPG 8x8 blocks (8 lanes each) per workgroup.
template <int BPG>
__global__ __launch_bounds__(BPG * 8) void mjpeg_blocks_kernel(const MjpegJob* __restrict__ jobs,
                                                                const uint32_t* __restrict__ tile_job,
                                                                uint32_t tile_base, uint32_t tiles, uint32_t per_xcd) {
    // Per block group: coefficients (int32, natural order), then the IDCT's row-pass output, then
    // the output samples, one after the other in the same LDS area.
    __shared__ __align__(16) double s_tmp[BPG * kIdctTmpStride];
    const uint32_t local = xcd_tile(blockIdx.x, per_xcd);
    if (local >= tiles) return;
    const uint32_t tile = tile_base + local;
    const MjpegJob& job = jobs[tile_job[tile]];
    const int g = threadIdx.x >> 3, lane = threadIdx.x & 7;
    const uint32_t b = (tile - job.tile_begin) * BPG + g;
    const bool active = b < job.block_count;
    const uint32_t mcu = active ? b / job.blocks_per_mcu : 0, k = active ? b % job.blocks_per_mcu : 0;
    const int comp = job.blk_comp[k];
    double* tmp = s_tmp + g * kIdctTmpStride;
    int32_t* coef = reinterpret_cast<int32_t*>(tmp);

    if (active) {
        const int4 raw = load_global_int4(job.levels + (size_t)b * 64 + lane * 8);
        const int4 qraw = load_global_int4(job.quant + 64 * job.qsel[comp] + lane * 8);
        const uint2 nraw = *reinterpret_cast<const uint2*>(&kZigzag[lane * 8]);
        const int16_t* l8 = reinterpret_cast<const int16_t*>(&raw);
        const uint16_t* q8 = reinterpret_cast<const uint16_t*>(&qraw);
        const uint8_t* n8 = reinterpret_cast<const uint8_t*>(&nraw);
#pragma unroll
        for (int i = 0; i < 8; ++i)   // coefficients[0] = dc_prediction * Q[0]; AC: level * Q[zz] at natural[zz]
            coef[n8[i]] = (int32_t)l8[i] * (int32_t)q8[i];
    }
    wave_lds_sync();
    const int4 lo = *reinterpret_cast<const int4*>(coef + lane * 8);
    const int4 hi = *reinterpret_cast<const int4*>(coef + lane * 8 + 4);
    const int32_t c[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
    const uint32_t rows = wave_nonzero_rows(active && (lo.x | lo.y | lo.z | lo.w | hi.x | hi.y | hi.z | hi.w) != 0);
    wave_lds_sync();   // every row is in registers before the row pass overwrites the area
    // Fast fused IDCT; the reference order again for the whole wave if a row is past the fast
    // form's coefficient bound or a sample (after + 128) lies within the tie window of a
    // half-integer (device_common.h).
    if (active) idct8x8_row_pass_fast(c, tmp + lane * 8);
    wave_lds_sync();
    double col[8];
    bool near = false;
    if (active) {
        idct8x8_column_pass_fast(tmp, lane, col);
#pragma unroll
        for (int y = 0; y < 8; ++y) col[y] = dadd(col[y], 128.0);
        near = !idct_row_within_fast_bound(c) || idct_near_half(col);
    }
    if (__ballot(near) != 0) {   // wave-uniform
        wave_lds_sync();   // the fast column reads are done before the row pass rewrites the area
        if (active) idct8x8_row_pass(c, tmp + lane * 8);
        wave_lds_sync();
        if (active) {
            idct8x8_column_pass(tmp, lane, rows, col);
#pragma unroll
            for (int y = 0; y < 8; ++y) col[y] = dadd(col[y], 128.0);
        }
    }
    wave_lds_sync();
    uint8_t* pix = reinterpret_cast<uint8_t*>(tmp);
    if (active) {
#pragma unroll
        for (int y = 0; y < 8; ++y) pix[y * 8 + lane] = (uint8_t)clampi(round_half_away_i32(col[y]), 0, 255);
    }
    wave_lds_sync();
    if (!active) return;
    // write_block: clip at the right / bottom plane edges; lane r writes row r.
    const uint32_t mcu_x = mcu % job.mcu_cols, mcu_y = mcu / job.mcu_cols;
    const uint32_t x0 = (mcu_x * job.comp_h[comp] + job.blk_bx[k]) * 8;
    const uint32_t y = (mcu_y * job.comp_v[comp] + job.blk_by[k]) * 8 + lane;
    const uint32_t pw = job.plane_w[comp], ph = job.plane_h[comp];
    if (y >= ph || x0 >= pw) return;
    const uint64_t row = *reinterpret_cast<const uint64_t*>(pix + lane * 8);
    uint8_t* dst = job.planes[comp] + (size_t)y * pw + x0;
    if (x0 + 8 <= pw) {
        store_row8(dst, row);
    } else {
        for (uint32_t i = 0; i < pw - x0; ++i) as_global(dst)[i] = (uint8_t)(row >> (8 * i));
    }
}

hipError_t launch_mjpeg(const MjpegJob* d_jobs, const uint32_t* d_tile_job, uint32_t tile_base, uint32_t tiles, uint32_t group_blocks,
                        hipStream_t s) {
    if (tiles == 0) return hipSuccess;
    const uint32_t per = xcd_per(tiles);
    switch (group_blocks) {
    case 16: hipLaunchKernelGGL(mjpeg_blocks_kernel<16>, dim3(per * kXcds), dim3(16 * 8), 0, s, d_jobs, d_tile_job, tile_base, tiles, per); break;
    case 32: hipLaunchKernelGGL(mjpeg_blocks_kernel<32>, dim3(per * kXcds), dim3(32 * 8), 0, s, d_jobs, d_tile_job, tile_base, tiles, per); break;
    case 64: hipLaunchKernelGGL(mjpeg_blocks_kernel<64>, dim3(per * kXcds), dim3(64 * 8), 0, s, d_jobs, d_tile_job, tile_base, tiles, per); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace mmr
