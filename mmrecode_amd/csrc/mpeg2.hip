// MPEG-2 frame-picture reconstruction: intra / non-intra dequantisation with mismatch control,
// f64 IDCT, half-sample motion compensation and put_block, for a batch of pictures in one
// persistent launch.
//
// Reference: decode_intra_block / decode_non_intra_block (crates/codecs/mpeg2/src/decoder.rs:734-846),
// inverse_dct (crates/codecs/mpeg2/src/transform.rs:90-110), predict_macroblock / predict_plane /
// halfpel_sample / clamped_sample / chroma_motion (decoder.rs:944-1046, :1150-1228),
// put_block (decoder.rs:1231-1298).
//
// Work decomposition: one task per (picture, macroblock, block 0..5); every task writes exactly
// the samples of its block (prediction, residual, or both), so tasks of a picture are
// independent. Tasks are grouped in tiles of 32 (8 lanes per block, 256 threads) and persistent
// workgroups take tiles from a device counter. Pictures that depend on each other (the anchor
// chain I -> P -> P, B after both anchors) are launched in dependency levels by the host: every
// picture of one launch only reads pictures completed by earlier launches, so the kernel needs no
// cross-workgroup synchronisation and no cache maintenance beyond the launch boundary (on MI355X
// the eight XCDs have separate L2 caches; agent-scope release / acquire inside a launch costs an
// L2 write-back / invalidate per tile).
#include "device_common.h"
#include "recon_internal.h"

namespace mmr {

MMR_DEV int div_euclid2(int v) { return v >= 0 ? (v >> 1) : -((-v + 1) >> 1); }

// clamped_sample (decoder.rs:1208-1222) + halfpel_sample (decoder.rs:1181-1206).
MMR_DEV uint32_t halfpel_sample(const uint8_t* __restrict__ ref, int stride, int height, int x, int y, int mvx, int mvy) {
    const int bx = x + div_euclid2(mvx), by = y + div_euclid2(mvy);
    const bool fx = (mvx & 1) != 0, fy = (mvy & 1) != 0;
    const int x0 = clampi(bx, 0, stride - 1), x1 = clampi(bx + 1, 0, stride - 1);
    const int y0 = clampi(by, 0, height - 1), y1 = clampi(by + 1, 0, height - 1);
    const uint32_t a = ref[(size_t)y0 * stride + x0];
    if (!fx && !fy) return a;
    if (fx && !fy) return (a + ref[(size_t)y0 * stride + x1] + 1) >> 1;
    if (!fx && fy) return (a + ref[(size_t)y1 * stride + x0] + 1) >> 1;
    const uint32_t b = ref[(size_t)y0 * stride + x1], c = ref[(size_t)y1 * stride + x0], d = ref[(size_t)y1 * stride + x1];
    return (a + b + c + d + 2) >> 2;
}

__global__ __launch_bounds__(kGroupThreads) void mpeg2_kernel(const Mpeg2Job* __restrict__ jobs, uint32_t job_count,
                                                              uint32_t total_tiles, uint32_t* __restrict__ counters,
                                                              uint32_t* __restrict__ status) {
    __shared__ __align__(16) int32_t coef[kBlocksPerGroup][64];
    __shared__ uint32_t s_tile;
    const int g = threadIdx.x >> 3, lane = threadIdx.x & 7;
    for (;;) {
        if (threadIdx.x == 0) s_tile = fetch_add_relaxed(&counters[0], 1);
        __syncthreads();
        const uint32_t tile = s_tile;
        if (tile >= total_tiles) break;
        const uint32_t ji = find_job(jobs, job_count, tile, [](const Mpeg2Job& j) { return j.tile_begin; });
        const Mpeg2Job& job = jobs[ji];

        const uint32_t mbs = job.mb_w * job.mb_h;
        const uint32_t task = (tile - job.tile_begin) * kBlocksPerGroup + g;
        const bool active = task < mbs * 6;
        const uint32_t mb = active ? task / 6 : 0;
        const int b = active ? (int)(task % 6) : 0;
        mmr_recon_mpeg2_macroblock m = job.mbs[mb];
        const bool intra = (m.flags & MMR_MPEG2_MB_INTRA) != 0;
        const bool coded = active && (intra || ((m.cbp >> (5 - b)) & 1));
        // Field prediction is not reconstructed on the device: flag the submission (the caller falls
        // back to CPU reconstruction) and predict as frame motion so the kernel still completes.
        if (active && b == 0 && (m.flags & MMR_MPEG2_MB_FIELD_MOTION) && lane == 0) atomicOr(status, kDevStatusUnsupported);

        if (coded) {
            const uint32_t index = m.block_offset + (intra ? (uint32_t)b : (uint32_t)__popc((m.cbp & 0x3fu) >> (6 - b)));
            const int4 raw = *reinterpret_cast<const int4*>(job.levels + (size_t)index * 64 + lane * 8);
            const int16_t* l8 = reinterpret_cast<const int16_t*>(&raw);
            const uint8_t* scan = job.alt_scan ? kAlternateScan : kZigzag;
            const int component = b < 4 ? 0 : 1;
            const uint8_t* W = job.matrices + 64 * (intra ? (component ? 2 : 0) : (component ? 3 : 1));
            const int qscale = m.qscale;
            int32_t val[8];
            uint32_t x = 0;
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int p = lane * 8 + i;
                const int32_t level = l8[i];
                const int pos = scan[p];
                int32_t v;
                if (intra && p == 0) {
                    v = level << job.dc_shift;  // DC: predictor << (3 - intra_dc_precision)
                } else if (level == 0) {
                    v = 0;
                } else {
                    const int32_t a = level < 0 ? -level : level;
                    const int32_t mag = intra ? (a * qscale * (int32_t)W[pos]) / 16 : ((2 * a + 1) * qscale * (int32_t)W[pos]) / 32;
                    v = clampi(level < 0 ? -mag : mag, -2048, 2047);
                }
                val[i] = v;
                x ^= (uint32_t)v;
            }
            // Mismatch control: coefficient 63 ^= (1 ^ XOR of all values) & 1; scan[63] == 63 for
            // both scans, so it is the last value of lane 7.
            const uint32_t all = group8_xor(x);
            if (lane == 7) val[7] ^= (int32_t)((1u ^ all) & 1u);
#pragma unroll
            for (int i = 0; i < 8; ++i) coef[g][scan[lane * 8 + i]] = val[i];
        }
        __syncthreads();

        if (active) {
            uint32_t row_mask = 0;
            if (coded) {
                const int4 lo = *reinterpret_cast<const int4*>(&coef[g][lane * 8]);
                const int4 hi = *reinterpret_cast<const int4*>(&coef[g][lane * 8 + 4]);
                const uint32_t nz = (lo.x | lo.y | lo.z | lo.w | hi.x | hi.y | hi.z | hi.w) != 0;
                row_mask = group8_or(nz << lane);
            }
            double col[8];
            if (coded) idct8x8_column(coef[g], lane, row_mask, col);

            const uint32_t mb_x = mb % job.mb_w, mb_y = mb / job.mb_w;
            const bool interlaced = (m.flags & MMR_MPEG2_MB_INTERLACED_DCT) != 0;
            const int stride_y = (int)job.mb_w * 16, height_y = (int)job.mb_h * 16;
            int plane, stride, height, x0, y0, step;
            if (b < 4) {
                plane = 0; stride = stride_y; height = height_y;
                x0 = (int)mb_x * 16 + (b & 1) * 8;
                y0 = (int)mb_y * 16 + (interlaced ? (b >> 1) : (b >> 1) * 8);
                step = interlaced ? 2 : 1;
            } else {
                plane = b - 3; stride = stride_y / 2; height = height_y / 2;
                x0 = (int)mb_x * 8; y0 = (int)mb_y * 8; step = 1;
            }
            uint8_t* out = job.planes[plane];
            const int xx = x0 + lane;
            if (intra) {
#pragma unroll
                for (int r = 0; r < 8; ++r) {
                    double s = round_half_away(col[r]);
                    s = s < -256.0 ? -256.0 : (s > 255.0 ? 255.0 : s);
                    out[(size_t)(y0 + r * step) * stride + xx] = (uint8_t)clampi((int)s, 0, 255);
                }
            } else {
                const bool has_f = (m.flags & MMR_MPEG2_MB_FORWARD) != 0, has_b = (m.flags & MMR_MPEG2_MB_BACKWARD) != 0;
                int fmx = m.mv[0][0], fmy = m.mv[0][1], bmx = m.mv[1][0], bmy = m.mv[1][1];
                if (b >= 4) {  // chroma_motion: v / 2 truncating toward zero
                    fmx /= 2; fmy /= 2; bmx /= 2; bmy /= 2;
                }
                const uint8_t* fr = job.fwd[plane];
                const uint8_t* br = job.bwd[plane];
#pragma unroll
                for (int r = 0; r < 8; ++r) {
                    const int yy = y0 + r * step;
                    uint32_t pred;
                    if (has_f && has_b) {
                        const uint32_t a = halfpel_sample(fr, stride, height, xx, yy, fmx, fmy);
                        const uint32_t c = halfpel_sample(br, stride, height, xx, yy, bmx, bmy);
                        pred = (a + c + 1) >> 1;
                    } else if (has_f) {
                        pred = halfpel_sample(fr, stride, height, xx, yy, fmx, fmy);
                    } else {
                        pred = halfpel_sample(br, stride, height, xx, yy, bmx, bmy);
                    }
                    int v = (int)pred;
                    if (coded) {
                        double s = round_half_away(col[r]);
                        s = s < -256.0 ? -256.0 : (s > 255.0 ? 255.0 : s);
                        v += (int)s;
                    }
                    out[(size_t)yy * stride + xx] = (uint8_t)clampi(v, 0, 255);
                }
            }
        }
        __syncthreads();   // s_tile is rewritten by the next iteration
    }
}

hipError_t launch_mpeg2(const Mpeg2Job* d_jobs, uint32_t job_count, uint32_t total_tiles, uint32_t* d_counters,
                        uint32_t groups, uint32_t* d_status, hipStream_t s) {
    if (total_tiles == 0) return hipSuccess;
    if (groups > total_tiles) groups = total_tiles;
    hipLaunchKernelGGL(mpeg2_kernel, dim3(groups), dim3(kGroupThreads), 0, s, d_jobs, job_count, total_tiles, d_counters, d_status);
    return hipGetLastError();
}

}  // namespace mmr
