// MJPEG block reconstruction: dequantisation, zig-zag scatter, f64 IDCT, placement.
//
// Reference: decode_block's dequantisation (crates/codecs/mjpeg/src/decoder.rs:502-546),
// inverse_dct (crates/codecs/mjpeg/src/transform.rs:86-109), write_block (decoder.rs:550-565),
// MCU walk (decoder.rs:346-401).
//
// Layout: 8 lanes per 8x8 block, 32 blocks per 256-thread workgroup. Lane j loads scan
// positions 8j..8j+7 (one 16-byte load; a group reads its block's 128 bytes contiguously),
// dequantises them and scatters to natural order in LDS. Lane x then runs the column-x IDCT and
// stores column x of the block.
#include "device_common.h"
#include "recon_internal.h"

namespace mmr {

__global__ __launch_bounds__(kGroupThreads) void mjpeg_blocks_kernel(const MjpegJob* __restrict__ jobs,
                                                                      uint32_t job_count, uint32_t total_blocks) {
    __shared__ __align__(16) int32_t coef[kBlocksPerGroup][64];
    const int g = threadIdx.x >> 3, lane = threadIdx.x & 7;
    const uint32_t gb = blockIdx.x * kBlocksPerGroup + g;
    const bool active = gb < total_blocks;
    uint32_t ji = 0;
    if (active) ji = find_job(jobs, job_count, gb, [](const MjpegJob& j) { return j.block_begin; });
    const MjpegJob& job = jobs[ji];
    const uint32_t b = active ? gb - job.block_begin : 0;
    const uint32_t mcu = b / job.blocks_per_mcu, k = b % job.blocks_per_mcu;
    const int comp = job.blk_comp[k];

    if (active) {
        const int16_t* lv = job.levels + (size_t)b * 64 + lane * 8;
        const int4 raw = *reinterpret_cast<const int4*>(lv);
        const int16_t* l8 = reinterpret_cast<const int16_t*>(&raw);
        const uint16_t* q = job.quant + 64 * job.qsel[comp];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int zz = lane * 8 + i;
            // coefficients[0] = dc_prediction * Q[0]; AC: level * Q[zz] at ZIGZAG_TO_NATURAL[zz].
            coef[g][kZigzag[zz]] = (int32_t)l8[i] * (int32_t)q[zz];
        }
    }
    __syncthreads();
    uint32_t row_mask = 0;
    if (active) {
        const int4 lo = *reinterpret_cast<const int4*>(&coef[g][lane * 8]);
        const int4 hi = *reinterpret_cast<const int4*>(&coef[g][lane * 8 + 4]);
        const uint32_t nz = (lo.x | lo.y | lo.z | lo.w | hi.x | hi.y | hi.z | hi.w) != 0;
        row_mask = group8_or(nz << lane);
    }
    if (!active) return;
    double col[8];
    idct8x8_column(coef[g], lane, row_mask, col);

    // write_block: clip at the right / bottom plane edges.
    const uint32_t mcu_x = mcu % job.mcu_cols, mcu_y = mcu / job.mcu_cols;
    const uint32_t x0 = (mcu_x * job.comp_h[comp] + job.blk_bx[k]) * 8;
    const uint32_t y0 = (mcu_y * job.comp_v[comp] + job.blk_by[k]) * 8;
    const uint32_t pw = job.plane_w[comp], ph = job.plane_h[comp];
    uint8_t* plane = job.planes[comp];
    const uint32_t x = x0 + lane;
    if (x < pw) {
#pragma unroll
        for (int y = 0; y < 8; ++y) {
            const uint32_t yy = y0 + y;
            if (yy < ph) plane[(size_t)yy * pw + x] = clamp_u8_d(round_half_away(dadd(col[y], 128.0)));
        }
    }
}

hipError_t launch_mjpeg(const MjpegJob* d_jobs, uint32_t job_count, uint32_t total_blocks, hipStream_t s) {
    if (total_blocks == 0) return hipSuccess;
    const uint32_t groups = (total_blocks + kBlocksPerGroup - 1) / kBlocksPerGroup;
    hipLaunchKernelGGL(mjpeg_blocks_kernel, dim3(groups), dim3(kGroupThreads), 0, s, d_jobs, job_count, total_blocks);
    return hipGetLastError();
}

}  // namespace mmr
