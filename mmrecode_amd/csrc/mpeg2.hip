// MPEG-2 frame-picture reconstruction: intra / non-intra dequantisation with mismatch control,
// f64 IDCT, half-sample motion compensation and put_block, for a batch of pictures per launch.
//
// Reference: decode_intra_block / decode_non_intra_block (crates/codecs/mpeg2/src/decoder.rs:734-846),
// inverse_dct (crates/codecs/mpeg2/src/transform.rs:90-110), predict_macroblock / predict_plane /
// halfpel_sample / clamped_sample / chroma_motion (decoder.rs:944-1046, :1150-1228),
// put_block (decoder.rs:1231-1298).
//
// Work decomposition: one task per (picture, macroblock, block 0..5); every task writes exactly
// the samples of its block (prediction, residual, or both), so tasks of a picture are
// independent. Tasks are grouped in tiles of 32 (8 lanes per block, 256 threads); tiles never
// straddle pictures and a host-built tile -> picture table gives each workgroup its picture with
// one scalar load. Pictures that depend on each other (the anchor chain I -> P -> P, B after both
// anchors) are launched in dependency levels by the host: every picture of one launch only reads
// pictures completed by earlier launches, so the kernel needs no cross-workgroup
// synchronisation (on MI355X the eight XCDs have separate L2 caches; ordering inside a launch
// would cost an L2 write-back / invalidate per tile).
//
// Per block, lane j issues its loads first (levels of scan positions 8j..8j+7, the quantiser
// matrix row of natural row j, and the reference samples of block row j: 9 per row and reference,
// as three aligned dwords when the row segment lies inside the picture). The levels go to natural
// order in LDS, lane v dequantises natural row v and runs the IDCT row pass, lane x the column
// pass, the residual goes back to LDS as rows, and lane j forms prediction row j with packed-byte
// averages, adds residual row j and stores the 8 samples with one 8-byte store. The 8 lanes of a
// block share a wave, so the LDS exchanges need no workgroup barrier.
#include "device_common.h"
#include "recon_internal.h"

namespace mmr {

MMR_DEV int div_euclid2(int v) { return v >= 0 ? (v >> 1) : -((-v + 1) >> 1); }

constexpr uint64_t kLo7 = 0xfefefefefefefefeull;
constexpr uint64_t kEven = 0x00ff00ff00ff00ffull;

// (a + b + 1) >> 1 in each byte: a + b = 2(a|b) - (a^b).
MMR_DEV uint64_t avg2_u8x8(uint64_t a, uint64_t b) { return (a | b) - (((a ^ b) & kLo7) >> 1); }

// (a + b + c + d + 2) >> 2 in each byte, with the even and odd bytes widened to 16-bit fields.
MMR_DEV uint64_t avg4_u8x8(uint64_t a, uint64_t b, uint64_t c, uint64_t d) {
    const uint64_t se = (a & kEven) + (b & kEven) + (c & kEven) + (d & kEven) + 0x0002000200020002ull;
    const uint64_t so = ((a >> 8) & kEven) + ((b >> 8) & kEven) + ((c >> 8) & kEven) + ((d >> 8) & kEven) + 0x0002000200020002ull;
    return ((se >> 2) & kEven) | (((so >> 2) & kEven) << 8);
}

// Samples x..x+8 of a reference row (the 9 a half-sample row needs), as three dwords w0..w2 whose
// bytes o..o+8 are the samples. clamped_sample (decoder.rs:1208-1222) clamps x into the row.
// When x >= 0 and x + 8 < stride (and plane and stride are 4-byte aligned) the prefetch reads the
// three aligned dwords that hold the samples. Otherwise (a row segment reaching past the picture
// edge, rare) it reads a dummy in-bounds dword triple and the samples are gathered one by one with
// clamped x when the row is consumed, so the common path has no branch between the loads and
// their use.
struct RefRow {
    uint32_t w0, w1, w2;
};

MMR_DEV RefRow load_ref_dwords(const MMR_GLOBAL uint8_t* p) {
    const MMR_GLOBAL uint32_t* q = reinterpret_cast<const MMR_GLOBAL uint32_t*>(p);
    RefRow r;
    r.w0 = q[0];
    r.w1 = q[1];
    r.w2 = q[2];
    return r;
}

MMR_DEV RefRow gather_ref_row(const MMR_GLOBAL uint8_t* row, int stride, int x) {
    uint32_t s[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) s[i] = row[clampi(x + i, 0, stride - 1)];
    RefRow r;
    r.w0 = s[0] | (s[1] << 8) | (s[2] << 16) | (s[3] << 24);
    r.w1 = s[4] | (s[5] << 8) | (s[6] << 16) | (s[7] << 24);
    r.w2 = s[8];
    return r;
}

// (samples 0..7, samples 1..8) of a row whose sample 0 is byte o of w0.
MMR_DEV void ref_row_bytes(const RefRow& r, uint32_t o, uint64_t& a, uint64_t& b) {
    const uint64_t w01 = ((uint64_t)r.w1 << 32) | r.w0;
    const uint64_t w12 = ((uint64_t)r.w2 << 32) | r.w1;
    const uint64_t lo = (uint32_t)(w01 >> (8 * o)) | ((uint64_t)(uint32_t)(w12 >> (8 * o)) << 32);
    const uint64_t hi = (r.w2 >> (8 * o)) & 0xffu;
    a = lo;
    b = (lo >> 8) | (hi << 56);
}

struct RefFetch {
    RefRow r0, r1;
    const MMR_GLOBAL uint8_t* row0;
    const MMR_GLOBAL uint8_t* row1;
    int x, stride;
    bool fast, fx, fy;
};

// Issues the loads of block row `yy` (8 samples from x0) predicted with half-sample vector (mvx, mvy).
MMR_DEV RefFetch fetch_ref(const uint8_t* ref, int stride, int height, int x0, int yy, int mvx, int mvy) {
    RefFetch f;
    const int bx = x0 + div_euclid2(mvx), by = yy + div_euclid2(mvy);
    f.fx = (mvx & 1) != 0;
    f.fy = (mvy & 1) != 0;
    f.x = bx;
    f.stride = stride;
    f.fast = bx >= 0 && bx + 8 < stride && ((reinterpret_cast<uintptr_t>(ref) | (uintptr_t)stride) & 3) == 0;
    const MMR_GLOBAL uint8_t* base = as_global(ref);
    f.row0 = base + (size_t)clampi(by, 0, height - 1) * stride;
    f.row1 = f.fy ? base + (size_t)clampi(by + 1, 0, height - 1) * stride : f.row0;
    const int off = f.fast ? (bx & ~3) : 0;
    f.r0 = load_ref_dwords((f.fast ? f.row0 : base) + off);
    f.r1 = load_ref_dwords((f.fast ? f.row1 : base) + off);
    return f;
}

// halfpel_sample (decoder.rs:1181-1206) for the 8 samples of the fetched row.
MMR_DEV uint64_t halfpel_row(const RefFetch& f) {
    RefRow r0 = f.r0, r1 = f.r1;
    uint32_t o = (uint32_t)(f.x & 3);
    if (!f.fast) {
        r0 = gather_ref_row(f.row0, f.stride, f.x);
        r1 = gather_ref_row(f.row1, f.stride, f.x);
        o = 0;
    }
    uint64_t a, b;
    ref_row_bytes(r0, o, a, b);
    if (!f.fy) return f.fx ? avg2_u8x8(a, b) : a;
    uint64_t c, d;
    ref_row_bytes(r1, o, c, d);
    return f.fx ? avg4_u8x8(a, b, c, d) : avg2_u8x8(a, c);
}

__global__ __launch_bounds__(kGroupThreads) void mpeg2_kernel(const Mpeg2Job* __restrict__ jobs,
                                                              const uint32_t* __restrict__ tile_job, uint32_t tile_base,
                                                              uint32_t* __restrict__ status) {
    __shared__ __align__(16) double s_tmp[kBlocksPerGroup * kIdctTmpStride];
    __shared__ __align__(16) int16_t s_blk[kBlocksPerGroup][64];   // levels in natural order, then residual rows
    const uint32_t tile = tile_base + blockIdx.x;
    const Mpeg2Job& job = jobs[tile_job[tile]];
    const int g = threadIdx.x >> 3, lane = threadIdx.x & 7;
    const uint32_t mbs = job.mb_w * job.mb_h;
    // Task (macroblock, block) of this group: the tile's first task is split on the scalar unit,
    // the per-group offset (0..36) with small-constant arithmetic.
    const uint32_t task0 = (tile - job.tile_begin) * kBlocksPerGroup;
    const uint32_t mb0 = task0 / 6, t = task0 % 6 + (uint32_t)g;
    const uint32_t dmb = (t * 43u) >> 8;   // t / 6 for t < 128
    const uint32_t task = task0 + (uint32_t)g;
    const bool active = task < mbs * 6;
    const uint32_t mb = active ? mb0 + dmb : 0;
    const int b = active ? (int)(t - dmb * 6) : 0;
    uint32_t mb_x = mb0 % job.mb_w + (active ? dmb : 0), mb_y = mb0 / job.mb_w;
    while (mb_x >= job.mb_w) {   // at most 6 steps (a tile spans at most 6 macroblocks)
        mb_x -= job.mb_w;
        ++mb_y;
    }
    mmr_recon_mpeg2_macroblock m;
    {
        static_assert(sizeof(mmr_recon_mpeg2_macroblock) == 16, "macroblock record is 4 dwords");
        const MMR_GLOBAL uint32_t* q = reinterpret_cast<const MMR_GLOBAL uint32_t*>(as_global(job.mbs + mb));
        const uint32_t w[4] = {q[0], q[1], q[2], q[3]};
        __builtin_memcpy(&m, w, sizeof m);
    }
    const bool intra = (m.flags & MMR_MPEG2_MB_INTRA) != 0;
    const bool coded = active && (intra || ((m.cbp >> (5 - b)) & 1));
    int16_t* blk = s_blk[g];
    double* tmp = s_tmp + g * kIdctTmpStride;
    // Field prediction is not reconstructed on the device: flag the submission (the caller falls
    // back to CPU reconstruction) and predict as frame motion so the kernel still completes.
    if (active && b == 0 && (m.flags & MMR_MPEG2_MB_FIELD_MOTION) && lane == 0) atomicOr(status, kDevStatusUnsupported);

    // Loads, in the order they are consumed: levels (scan positions 8*lane..), the scan positions,
    // the quantiser matrix row of natural row `lane`, then the reference rows (consumed after the
    // IDCT). They are unconditional (uncoded blocks read the 256-byte matrix block instead of
    // levels), so no branch joins force an early wait.
    const int component = b < 4 ? 0 : 1;
    const uint32_t index = m.block_offset + (intra ? (uint32_t)b : (uint32_t)__popc((m.cbp & 0x3fu) >> (6 - b)));
    const int16_t* lv = coded ? job.levels + (size_t)index * 64 : reinterpret_cast<const int16_t*>(job.matrices);
    const int4 raw = load_global_int4(lv + lane * 8);
    const uint2 sraw = *reinterpret_cast<const uint2*>((job.alt_scan ? kAlternateScan : kZigzag) + lane * 8);
    const uint8_t* W = job.matrices + 64 * (intra ? (component ? 2 : 0) : (component ? 3 : 1));
    const uint2 wrow = load_global_uint2(W + lane * 8);

    // Geometry of block b and of the row this lane predicts / stores.
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
    const int yy = y0 + lane * step;
    const bool has_f = !intra && (m.flags & MMR_MPEG2_MB_FORWARD) != 0;
    const bool has_b = !intra && (m.flags & MMR_MPEG2_MB_BACKWARD) != 0;
    int fmx = m.mv[0][0], fmy = m.mv[0][1], bmx = m.mv[1][0], bmy = m.mv[1][1];
    if (b >= 4) {  // chroma_motion: v / 2 truncating toward zero
        fmx /= 2; fmy /= 2; bmx /= 2; bmy /= 2;
    }
    // Unused directions fetch row 0 of their plane (a dummy that hits one cache line).
    const uint8_t* fwd = pick3(job.fwd, plane);
    const uint8_t* bwd = pick3(job.bwd, plane);
    const RefFetch ff = fetch_ref(fwd, stride, height, x0, has_f ? yy : 0, has_f ? fmx : -2 * x0, has_f ? fmy : 0);
    const RefFetch fb = fetch_ref(bwd, stride, height, x0, has_b ? yy : 0, has_b ? bmx : -2 * x0, has_b ? bmy : 0);

    // Levels to natural order.
    if (coded) {
        const uint8_t* s8 = reinterpret_cast<const uint8_t*>(&sraw);
        const int16_t* l8 = reinterpret_cast<const int16_t*>(&raw);
#pragma unroll
        for (int i = 0; i < 8; ++i) blk[s8[i]] = l8[i];
    }
    wave_lds_sync();
    // Dequantisation of natural row `lane` and the IDCT row pass.
    bool row_nonzero = false;
    if (coded) {
        const int4 rr = *reinterpret_cast<const int4*>(blk + lane * 8);
        const int16_t* l8 = reinterpret_cast<const int16_t*>(&rr);
        const uint8_t* w8 = reinterpret_cast<const uint8_t*>(&wrow);
        const int qscale = m.qscale;
        int32_t val[8];
        uint32_t x = 0;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int32_t level = l8[i];
            int32_t v;
            if (intra && lane == 0 && i == 0) {
                v = level << job.dc_shift;  // DC: predictor << (3 - intra_dc_precision)
            } else if (level == 0) {
                v = 0;
            } else {
                const int32_t a = level < 0 ? -level : level;
                const int32_t w = w8[i];
                const int32_t mag = intra ? (a * qscale * w) / 16 : ((2 * a + 1) * qscale * w) / 32;
                v = clampi(level < 0 ? -mag : mag, -2048, 2047);
            }
            val[i] = v;
            x ^= (uint32_t)v;
        }
        // Mismatch control: coefficient 63 ^= (1 ^ XOR of all values) & 1 (natural row 7, column 7).
        const uint32_t all = group8_xor(x);
        if (lane == 7) val[7] ^= (int32_t)((1u ^ all) & 1u);
        row_nonzero = (val[0] | val[1] | val[2] | val[3] | val[4] | val[5] | val[6] | val[7]) != 0;
        idct8x8_row_pass(val, tmp + lane * 8);
    }
    const uint32_t rows = wave_nonzero_rows(row_nonzero);
    wave_lds_sync();
    if (coded) {
        double col[8];
        idct8x8_column_pass(tmp, lane, rows, col);
        // put_block: residual rounded and clamped to [-256, 255] (transform.rs:90-110).
#pragma unroll
        for (int r = 0; r < 8; ++r) blk[r * 8 + lane] = (int16_t)clampi(round_half_away_i32(col[r]), -256, 255);
    }
    wave_lds_sync();
    if (!active) return;

    // Row `lane` of the block: prediction (0 for intra) + residual, clamped to 0..255.
    uint64_t pred = 0;
    if (has_f && has_b) {
        pred = avg2_u8x8(halfpel_row(ff), halfpel_row(fb));
    } else if (has_f) {
        pred = halfpel_row(ff);
    } else if (has_b) {
        pred = halfpel_row(fb);
    }
    uint64_t out = pred;
    if (coded) {   // prediction + residual in 16-bit pairs, clamped to 0..255
        const int4 rr = *reinterpret_cast<const int4*>(blk + lane * 8);
        const uint32_t r32[4] = {(uint32_t)rr.x, (uint32_t)rr.y, (uint32_t)rr.z, (uint32_t)rr.w};
        const uint32_t p32[2] = {(uint32_t)pred, (uint32_t)(pred >> 32)};
        uint32_t o32[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            uint32_t packed[2];
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const uint32_t pw = (p32[h] >> (16 * k)) & 0xffffu;      // two prediction bytes
                const mmr_v2i16 pv = {(int16_t)(pw & 0xffu), (int16_t)(pw >> 8)};
                mmr_v2i16 rv;
                __builtin_memcpy(&rv, &r32[h * 2 + k], 4);
                mmr_v2i16 v = pv + rv;                                    // |v| <= 511: no overflow
                v = __builtin_elementwise_max(v, (mmr_v2i16){0, 0});
                v = __builtin_elementwise_min(v, (mmr_v2i16){255, 255});
                uint32_t vw;
                __builtin_memcpy(&vw, &v, 4);
                packed[k] = (vw & 0xffu) | ((vw >> 8) & 0xff00u);
            }
            o32[h] = packed[0] | (packed[1] << 16);
        }
        out = (uint64_t)o32[0] | ((uint64_t)o32[1] << 32);
    }
    uint8_t* dst = pick3(job.planes, plane);
    store_row8(dst + (size_t)yy * stride + x0, out);
}

hipError_t launch_mpeg2(const Mpeg2Job* d_jobs, const uint32_t* d_tile_job, uint32_t tile_base, uint32_t tiles,
                        uint32_t* d_status, hipStream_t s) {
    if (tiles == 0) return hipSuccess;
    hipLaunchKernelGGL(mpeg2_kernel, dim3(tiles), dim3(kGroupThreads), 0, s, d_jobs, d_tile_job, tile_base, d_status);
    return hipGetLastError();
}

}  // namespace mmr
