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
// Per block, lane j issues its loads first (levels of scan positions 8j..8j+7 and the reference
// samples of block row j: 9 per row and reference, as three aligned dwords when the row segment
// lies inside the picture). Lane j dequantises its scan positions against a per-wave copy of the
// picture's quantiser matrices in scan order and scatters them to natural order in LDS; lane v
// then runs the IDCT row pass on natural row v, lane x the column pass, the residual goes back to
// LDS as rows, and lane j forms prediction row j with packed-byte averages, adds residual row j and
// stores the 8 samples with one 8-byte store. The 8 lanes of a block share a wave, so the LDS
// exchanges need no workgroup barrier.
#include "device_common.h"
#include "recon_internal.h"

namespace mmr {

// Byte-wise rounded averages with v_lerp_u8: lerp(a, b, r) = (a + b + r) >> 1 in every byte, r the
// low bit of each byte of the third operand.
constexpr uint32_t kOnes = 0x01010101u;
MMR_DEV uint32_t avg2_u8x4(uint32_t a, uint32_t b) { return __builtin_amdgcn_lerp(a, b, kOnes); }

// (a + b + c + d + 2) >> 2 in every byte: with p = (a + b + 1) >> 1 and q = (c + d + 1) >> 1, the
// result is (p + q + r) >> 1 where r = 1 exactly when a + b and c + d are both even (the two
// round-ups then did not happen and the final one must).
MMR_DEV uint32_t avg4_u8x4(uint32_t a, uint32_t b, uint32_t c, uint32_t d) {
    const uint32_t r = ~((a ^ b) | (c ^ d)) & kOnes;
    return __builtin_amdgcn_lerp(avg2_u8x4(a, b), avg2_u8x4(c, d), r);
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

// Samples 0..7 (a0, a1) and 1..8 (b0, b1) of a row whose sample 0 is byte o of w0.
MMR_DEV void ref_row_bytes(const RefRow& r, uint32_t o, uint32_t& a0, uint32_t& a1, uint32_t& b0, uint32_t& b1) {
    a0 = __builtin_amdgcn_alignbyte(r.w1, r.w0, o);
    a1 = __builtin_amdgcn_alignbyte(r.w2, r.w1, o);
    const uint32_t s8 = __builtin_amdgcn_alignbyte(r.w2, r.w2, o);   // byte 0: sample 8
    b0 = __builtin_amdgcn_alignbyte(a1, a0, 1);
    b1 = __builtin_amdgcn_alignbyte(s8, a1, 1);
}

// The plane record of a lane's block (two 16-byte loads from the job).
MMR_DEV Mpeg2Plane load_plane(const Mpeg2Plane* p) {
    const int4 a = load_global_int4(p);
    const int4 c = load_global_int4(reinterpret_cast<const char*>(p) + 16);
    Mpeg2Plane r;
    r.dst = reinterpret_cast<uint8_t*>(((uint64_t)(uint32_t)a.y << 32) | (uint32_t)a.x);
    r.fwd = reinterpret_cast<const uint8_t*>(((uint64_t)(uint32_t)a.w << 32) | (uint32_t)a.z);
    r.bwd = reinterpret_cast<const uint8_t*>(((uint64_t)(uint32_t)c.y << 32) | (uint32_t)c.x);
    r.stride = (uint32_t)c.z;
    r.height = (uint32_t)c.w;
    return r;
}

// chroma_motion (decoder.rs:1220-1222) on a packed (x, y) vector: both halves v / 2, truncated
// toward zero.
typedef uint16_t mmr_v2u16 __attribute__((ext_vector_type(2)));
MMR_DEV uint32_t halve_mv(uint32_t packed) {
    mmr_v2i16 v;
    __builtin_memcpy(&v, &packed, 4);
    mmr_v2u16 u;
    __builtin_memcpy(&u, &packed, 4);
    const mmr_v2u16 neg = u >> (mmr_v2u16){15, 15};
    mmr_v2i16 n;
    __builtin_memcpy(&n, &neg, 4);
    v = (v + n) >> (mmr_v2i16){1, 1};
    uint32_t out;
    __builtin_memcpy(&out, &v, 4);
    return out;
}

struct RefFetch {
    RefRow r0, r1;
    const MMR_GLOBAL uint8_t* row0;
    const MMR_GLOBAL uint8_t* row1;
    int x, stride;
    bool fast, fx, fy;
};

// Issues the loads of block row `yy` (8 samples from x0) predicted with half-sample vector
// (mvx, mvy); halfpel_sample's integer part is floor(v / 2) = v >> 1. Rows are `pitch` bytes apart
// and `stride` samples wide (a field of a frame plane: pitch = 2 * stride). With use == false (the
// direction is not predicted) both dword triples come from the start of the plane. `aligned`: the
// plane pointers of the picture are 4-byte aligned (strides are multiples of 8).
MMR_DEV RefFetch fetch_ref(const uint8_t* ref, int stride, int pitch, int height, int x0, int yy, int mvx, int mvy,
                           bool use, bool aligned) {
    RefFetch f;
    const int bx = x0 + (mvx >> 1), by = yy + (mvy >> 1);
    f.fx = (mvx & 1) != 0;
    f.fy = (mvy & 1) != 0;
    f.x = bx;
    f.stride = stride;
    f.fast = aligned && bx >= 0 && bx + 8 < stride;
    const uint32_t o0 = __umul24((uint32_t)clampi(by, 0, height - 1), (uint32_t)pitch);
    const uint32_t o1 = __umul24((uint32_t)clampi(by + (f.fy ? 1 : 0), 0, height - 1), (uint32_t)pitch);
    const MMR_GLOBAL uint8_t* base = as_global(ref);
    f.row0 = base + o0;
    f.row1 = base + o1;
    const uint32_t xo = (uint32_t)bx & ~3u;
    const bool load_rows = use && f.fast;
    f.r0 = load_ref_dwords(base + (load_rows ? o0 + xo : 0u));
    f.r1 = load_ref_dwords(base + (load_rows ? o1 + xo : 0u));
    return f;
}

// halfpel_sample (decoder.rs:1181-1206) for the 8 samples of the fetched row, as two dwords,
// without branches on the half-sample flags: with p = fx ? avg2(a, b) : a and q the same on the
// next row, the result is fy ? (fx ? avg4 : avg2(a, c)) : p, and avg4 = lerp(p, q, r) with r the
// avg4_u8x4 correction; lerp(x, x, 0) = x covers the cases without averaging.
MMR_DEV uint2 halfpel_row(const RefFetch& f) {
    RefRow r0 = f.r0, r1 = f.r1;
    uint32_t o = (uint32_t)(f.x & 3);
    if (!f.fast) {
        r0 = gather_ref_row(f.row0, f.stride, f.x);
        r1 = gather_ref_row(f.row1, f.stride, f.x);
        o = 0;
    }
    uint32_t a[2], b[2], c[2], d[2];
    ref_row_bytes(r0, o, a[0], a[1], b[0], b[1]);
    ref_row_bytes(r1, o, c[0], c[1], d[0], d[1]);
    const uint32_t rx = f.fx ? kOnes : 0u;
    uint32_t out[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const uint32_t p = __builtin_amdgcn_lerp(a[h], f.fx ? b[h] : a[h], rx);
        const uint32_t q = __builtin_amdgcn_lerp(c[h], f.fx ? d[h] : c[h], rx);
        const uint32_t r4 = ~((a[h] ^ b[h]) | (c[h] ^ d[h])) & kOnes;
        const uint32_t ry = f.fy ? (f.fx ? r4 : kOnes) : 0u;
        out[h] = __builtin_amdgcn_lerp(p, f.fy ? q : p, ry);
    }
    return make_uint2(out[0], out[1]);
}

// Block order of a picture, row-local: the 6 * mb_w blocks of macroblock row r occupy positions
// 6 * mb_w * r .. 6 * mb_w * (r + 1) - 1, coded blocks first (in macroblock order), then the blocks
// without residual, so that tiles of coded blocks run the dequantisation and IDCT on every lane,
// tiles without residual skip them, and both stay near each other in the picture (and in the tile
// sequence). One wave per macroblock row. An entry is 16 bytes and holds everything the block's
// reconstruction needs from its macroblock record, so the main kernel's first load is the entry
// itself: x = block | mb_x << 3 | mb_y << 13 | coded << 23 | macroblock flags << 24,
// y = level record index | qscale << 24, z / w = forward / backward vector (packed int16 x, y).
// Field-predicted macroblocks are checked here: the picture has field records and their field
// selects are 0 or 1.
constexpr int kOrderThreads = 256;

MMR_DEV uint32_t wave_exclusive_sum(uint32_t v) {
    const uint32_t lane = __lane_id();
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, d, 64);
        if (lane >= (uint32_t)d) x += y;
    }
    return x - v;
}
MMR_DEV uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

__global__ __launch_bounds__(kOrderThreads) void mpeg2_order_kernel(const Mpeg2Job* __restrict__ jobs, uint32_t job_count,
                                                                    uint32_t total_rows, uint32_t* __restrict__ status) {
    const uint32_t grow = blockIdx.x * (kOrderThreads / 64) + (threadIdx.x >> 6);
    if (grow >= total_rows) return;
    const Mpeg2Job& job = jobs[find_job(jobs, job_count, grow, [](const Mpeg2Job& j) { return j.row_begin; })];
    const uint32_t r = grow - job.row_begin, mb_w = job.mb_w, lane = __lane_id();
    const MMR_GLOBAL uint32_t* rec = reinterpret_cast<const MMR_GLOBAL uint32_t*>(as_global(job.mbs));
    constexpr uint32_t kRecWords = sizeof(mmr_recon_mpeg2_macroblock) / 4;
    // first dword of a record: flags | cbp << 8 | qscale << 16; bit (5 - b) set when block b is coded
    auto coded_mask = [](uint32_t w) -> uint32_t { return (w & MMR_MPEG2_MB_INTRA) ? 0x3fu : ((w >> 8) & 0x3fu); };
    uint32_t row_coded = 0;
    bool bad_field = false;
    for (uint32_t c0 = 0; c0 < mb_w; c0 += 64) {
        const uint32_t x = c0 + lane;
        const uint32_t w = x < mb_w ? rec[((size_t)r * mb_w + x) * kRecWords] : 0u;
        row_coded += wave_sum((uint32_t)__popc(coded_mask(w)));
        if ((w & MMR_MPEG2_MB_FIELD_MOTION) && !(w & MMR_MPEG2_MB_INTRA)) {
            if (!job.field_motion) {
                bad_field = true;
            } else {
                const MMR_GLOBAL uint8_t* sel = reinterpret_cast<const MMR_GLOBAL uint8_t*>(
                    as_global(job.field_motion + (size_t)r * mb_w + x)) + 8;   // field_select[2][2]
                bad_field |= (sel[0] | sel[1] | sel[2] | sel[3]) > 1u;
            }
        }
    }
    if (bad_field) atomicOr(status, kDevStatusInvalidData);
    MMR_GLOBAL mmr_v4i32* order = reinterpret_cast<MMR_GLOBAL mmr_v4i32*>(as_global(job.order)) + (size_t)6 * mb_w * r;
    uint32_t coded_before = 0;   // coded blocks of the row before this chunk
    for (uint32_t c0 = 0; c0 < mb_w; c0 += 64) {
        const uint32_t x = c0 + lane;
        mmr_v4i32 mbr = {0, 0, 0, 0};
        if (x < mb_w) mbr = *reinterpret_cast<const MMR_GLOBAL mmr_v4i32*>(rec + ((size_t)r * mb_w + x) * kRecWords);
        const uint32_t w = (uint32_t)mbr.x;
        const uint32_t mask = coded_mask(w), n = (uint32_t)__popc(mask);
        uint32_t cpos = coded_before + wave_exclusive_sum(n);   // coded blocks before this macroblock
        if (x < mb_w) {
            const bool intra = (w & MMR_MPEG2_MB_INTRA) != 0;
            const uint32_t base = (uint32_t)mbr.w;   // block_offset: first level record of the macroblock
            const uint32_t head = (r << 13) | (x << 3) | ((w & 0xffu) << 24);
            const uint32_t qs = ((w >> 16) & 0xffu) << 24;
#pragma unroll
            for (uint32_t b = 0; b < 6; ++b) {
                const bool coded = (mask >> (5 - b)) & 1u;
                // level record of block b: intra blocks are all coded; non-intra ones count the
                // coded blocks before b in the coded-block pattern
                const uint32_t index = base + (intra ? b : (uint32_t)__popc(((w >> 8) & 0x3fu) >> (6 - b)));
                if (coded && index >= (1u << 24)) atomicOr(status, kDevStatusInvalidData);   // 24-bit entry field
                const mmr_v4i32 e = {(int32_t)(head | b | (coded ? 1u << 23 : 0u)), (int32_t)((coded ? index : 0u) | qs), mbr.y, mbr.z};
                if (coded) {
                    order[cpos++] = e;
                } else {
                    // blocks of the row before this one minus coded blocks before it
                    order[row_coded + (x * 6 + b) - cpos] = e;
                }
            }
        }
        coded_before += wave_sum(n);
    }
}

// Scan position -> byte offset of the natural-order int32 coefficient (4 * natural index), for the
// zig-zag and the alternate scan.
__constant__ static const uint8_t kScanCoefOffset[2][64] = {
    {0, 4, 32, 64, 36, 8, 12, 40, 68, 96, 128, 100, 72, 44, 16, 20, 48, 76, 104, 132, 160, 192, 164, 136, 108, 80, 52, 24,
     28, 56, 84, 112, 140, 168, 196, 224, 228, 200, 172, 144, 116, 88, 60, 92, 120, 148, 176, 204, 232, 236, 208, 180,
     152, 124, 156, 184, 212, 240, 244, 216, 188, 220, 248, 252},
    {0, 32, 64, 96, 4, 36, 8, 40, 68, 100, 128, 160, 192, 224, 228, 196, 164, 132, 104, 72, 12, 44, 16, 48, 76, 108, 136,
     168, 200, 232, 140, 172, 204, 236, 80, 112, 20, 52, 24, 56, 84, 116, 144, 176, 208, 240, 148, 180, 212, 244, 88, 120,
     28, 60, 92, 124, 152, 184, 216, 248, 156, 188, 220, 252}};

// BPG 8x8 blocks (8 lanes each) per workgroup.
// At least 7 waves per SIMD (72 VGPRs; the LDS of 32-block workgroups allows 8): measured 3 % faster
// than the 6 waves the unconstrained allocation gives, although 4 VGPRs then spill.
template <int BPG>
__global__ __launch_bounds__(BPG * 8, 7) void mpeg2_kernel(const Mpeg2Job* __restrict__ jobs,
                                                        const uint32_t* __restrict__ tile_job, uint32_t tile_base,
                                                        uint32_t tiles, uint32_t per_xcd, uint32_t* __restrict__ status) {
    // Per block group: natural-order coefficients, then the IDCT row-pass output, then the residual
    // rows, one after the other in the same LDS area.
    __shared__ __align__(16) double s_tmp[BPG * kIdctTmpStride];
    __shared__ __align__(16) uint8_t s_wscan[BPG / 8][256];
    const uint32_t local = xcd_tile(blockIdx.x, per_xcd);
    if (local >= tiles) return;
    const uint32_t tile = tile_base + local;
    const Mpeg2Job& job = jobs[tile_job[tile]];
    const int g = threadIdx.x >> 3, lane = threadIdx.x & 7;
    const uint32_t mbs = job.mb_w * job.mb_h;
    // Block of this group, from the picture's coded-first block order.
    const uint32_t pos = (tile - job.tile_begin) * BPG + (uint32_t)g;
    const bool active = pos < mbs * 6;
    const int4 ent = load_global_int4(job.order + (size_t)(active ? pos : 0) * 4);
    const uint32_t entry = (uint32_t)ent.x;
    const uint32_t mb_x = (entry >> 3) & 0x3ffu, mb_y = (entry >> 13) & 0x3ffu;
    const int b = (int)(entry & 7u);
    const uint32_t mb_flags = entry >> 24;
    const uint32_t mv_words[2] = {(uint32_t)ent.z, (uint32_t)ent.w};   // forward, backward (x, y) as packed int16
    // Plane of block b (Y for 0..3, Cb for 4, Cr for 5) with its reference planes.
    const int plane = b < 4 ? 0 : b - 3;
    const Mpeg2Plane pr = load_plane(job.plane + plane);
    const bool intra = (mb_flags & MMR_MPEG2_MB_INTRA) != 0;
    const bool coded = active && ((entry >> 23) & 1u);
    double* tmp = s_tmp + g * kIdctTmpStride;

    // Quantiser matrices in scan order, one copy per wave: wscan[m][p] = W_m[scan[p]] (all blocks
    // of a workgroup belong to one picture).
    uint8_t* wscan = s_wscan[threadIdx.x >> 6];
    {
        const int p = threadIdx.x & 63;
        const uint32_t sp = (job.alt_scan ? kAlternateScan : kZigzag)[p];
        const MMR_GLOBAL uint8_t* Wg = as_global(job.matrices);
#pragma unroll
        for (int mi = 0; mi < 4; ++mi) wscan[mi * 64 + p] = Wg[mi * 64 + sp];
    }
    // Loads: levels (scan positions 8*lane..) and the scan positions here, unconditional (uncoded
    // blocks read the 256-byte matrix block instead of levels) so that no branch join forces an
    // early wait; the reference rows are fetched after the IDCT (issued before it, they kept ~18
    // more registers live across it and cost a wave per SIMD of occupancy, which measured slower).
    const int component = b < 4 ? 0 : 1;
    const uint32_t index = (uint32_t)ent.y & 0xffffffu;
    const int16_t* lv = coded ? job.levels + (size_t)index * 64 : reinterpret_cast<const int16_t*>(job.matrices);
    const int4 raw = load_global_int4(lv + lane * 8);
    const uint2 sraw = *reinterpret_cast<const uint2*>(kScanCoefOffset[job.alt_scan ? 1 : 0] + lane * 8);

    // Geometry of block b and of the row this lane predicts / stores.
    const bool interlaced = (mb_flags & MMR_MPEG2_MB_INTERLACED_DCT) != 0;
    const int stride = (int)pr.stride, height = (int)pr.height;
    int x0, y0, step;
    if (b < 4) {
        x0 = (int)mb_x * 16 + (b & 1) * 8;
        y0 = (int)mb_y * 16 + (interlaced ? (b >> 1) : (b >> 1) * 8);
        step = interlaced ? 2 : 1;
    } else {
        x0 = (int)mb_x * 8;
        y0 = (int)mb_y * 8;
        step = 1;
    }
    const int yy = y0 + lane * step;
    const bool has_f = !intra && (mb_flags & MMR_MPEG2_MB_FORWARD) != 0;
    const bool has_b = !intra && (mb_flags & MMR_MPEG2_MB_BACKWARD) != 0;
    uint32_t fmv = mv_words[0], bmv = mv_words[1];
    const uint8_t* fref = pr.fwd;
    const uint8_t* bref = pr.bwd;
    int pitch = stride, ry = yy, fh = height, bh = height;
    if (!intra && (mb_flags & MMR_MPEG2_MB_FIELD_MOTION) && job.field_motion) {
        // predict_field_plane / halfpel_field_sample (decoder.rs:1053-1148): the lane's row has
        // parity yy & 1 and takes that parity's vector and field select in each direction (parity
        // 0: the record's vectors, parity 1: the field record's). It reads field `sel` of the
        // reference -- rows sel, sel + 2, ..., a plane of pitch 2 * stride and (height - 1 - sel) / 2
        // + 1 rows -- at field row yy / 2 + floor(mvy / 2). The order pass reports a missing field
        // array (then predicted as frame motion here) and selects other than 0 or 1.
        const MMR_GLOBAL uint16_t* fm = reinterpret_cast<const MMR_GLOBAL uint16_t*>(
            as_global(job.field_motion + (size_t)mb_y * job.mb_w + mb_x));
        const uint32_t parity = (uint32_t)yy & 1u;
        if (parity) {
            fmv = (uint32_t)fm[0] | ((uint32_t)fm[1] << 16);
            bmv = (uint32_t)fm[2] | ((uint32_t)fm[3] << 16);
        }
        const uint32_t fsel = ((uint32_t)fm[4] >> (parity * 8)) & 1u;   // field_select[0][parity]
        const uint32_t bsel = ((uint32_t)fm[5] >> (parity * 8)) & 1u;   // field_select[1][parity]
        fref += fsel * (uint32_t)stride;
        bref += bsel * (uint32_t)stride;
        fh = (height - 1 - (int)fsel) / 2 + 1;
        bh = (height - 1 - (int)bsel) / 2 + 1;
        pitch = 2 * stride;
        ry = yy >> 1;
    }
    if (b >= 4) {
        fmv = halve_mv(fmv);
        bmv = halve_mv(bmv);
    }
    const int fmx = (int16_t)(fmv & 0xffffu), fmy = (int16_t)(fmv >> 16);
    const int bmx = (int16_t)(bmv & 0xffffu), bmy = (int16_t)(bmv >> 16);

    // Dequantisation in scan order (decode_intra_block / decode_non_intra_block), branch-free with
    // 24-bit multiplies. Intra: |level| * W * qscale / 16; non-intra: (2 |level| + 1) * W * qscale
    // / 32 for level != 0. Both are (k * W * qscale) >> 5 with k = 2 |level| + (non-intra and
    // level != 0) = (|level| << 1) | min(|level|, non_intra), which is 0 for a zero level.
    // |level| <= 2^15 and W * qscale < 2^15, so k * W * qscale < 2^32 is exact in 32 bits and equals
    // the reference's i32 product wherever that one does not overflow.
    wave_lds_sync();   // wscan
    int32_t* coef = reinterpret_cast<int32_t*>(tmp);   // natural-order coefficients, then tmp, then residual
    if (coded) {
        const uint2 wrow = *reinterpret_cast<const uint2*>(wscan + 64 * (intra ? (component ? 2 : 0) : (component ? 3 : 1)) + lane * 8);
        const int16_t* l8 = reinterpret_cast<const int16_t*>(&raw);
        const uint8_t* s8 = reinterpret_cast<const uint8_t*>(&sraw);
        const uint32_t qs = (uint32_t)ent.y >> 24, non_intra = intra ? 0u : 1u;
        int32_t val[8];
        uint32_t x = 0;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int32_t level = l8[i];
            const uint32_t a = (uint32_t)(level < 0 ? -level : level);
            const uint32_t w = ((i < 4 ? wrow.x : wrow.y) >> (8 * (i & 3))) & 0xffu;
            const uint32_t k = (a << 1) | min(a, non_intra);
            const int32_t mag = (int32_t)(__umul24(k, __umul24(w, qs)) >> 5);
            int32_t v = clampi(level < 0 ? -mag : mag, -2048, 2047);
            if (i == 0 && intra && lane == 0) v = level << job.dc_shift;   // DC: predictor << (3 - intra_dc_precision)
            val[i] = v;
            x ^= (uint32_t)v;
        }
        // Mismatch control: coefficient 63 ^= (1 ^ XOR of all values) & 1; scan position 63 is
        // natural 63 in both scans.
        const uint32_t all = group8_xor(x);
        if (lane == 7) val[7] ^= (int32_t)((1u ^ all) & 1u);
        char* coef_bytes = reinterpret_cast<char*>(coef);
#pragma unroll
        for (int i = 0; i < 8; ++i) *reinterpret_cast<int32_t*>(coef_bytes + s8[i]) = val[i];
    }
    wave_lds_sync();
    int32_t c[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (coded) {
        const int4 lo = *reinterpret_cast<const int4*>(coef + lane * 8);
        const int4 hi = *reinterpret_cast<const int4*>(coef + lane * 8 + 4);
        c[0] = lo.x; c[1] = lo.y; c[2] = lo.z; c[3] = lo.w; c[4] = hi.x; c[5] = hi.y; c[6] = hi.z; c[7] = hi.w;
    }
    const uint32_t rows = wave_nonzero_rows((c[0] | c[1] | c[2] | c[3] | c[4] | c[5] | c[6] | c[7]) != 0);
    wave_lds_sync();   // every row is in registers before the row pass overwrites the area
    // inverse_dct (transform.rs:90-110): the fused fast form, then the reference's order for the
    // whole wave if any of its samples lies within the tie window (device_common.h: the AC
    // coefficients are saturated to [-2048, 2047] above and the DC is an int16 << dc_shift <= 3,
    // the block sizes the window's bound covers).
    if (coded) idct8x8_row_pass_fast(c, tmp + lane * 8);
    wave_lds_sync();
    double col[8];
    bool near = false;
    if (coded) {
        idct8x8_column_pass_fast(tmp, lane, col);
        near = idct_near_half(col);
    }
    if (__ballot(near) != 0) {   // wave-uniform
        wave_lds_sync();   // the fast column reads are done before the row pass rewrites the area
        if (coded) idct8x8_row_pass(c, tmp + lane * 8);
        wave_lds_sync();
        if (coded) idct8x8_column_pass(tmp, lane, rows, col);
    }
    wave_lds_sync();   // the column reads are done before the residual overwrites the area
    int16_t* blk = reinterpret_cast<int16_t*>(tmp);
    if (coded) {
        // put_block: residual rounded and clamped to [-256, 255] (transform.rs:90-110).
#pragma unroll
        for (int r = 0; r < 8; ++r) blk[r * 8 + lane] = (int16_t)clampi(round_half_away_i32(col[r]), -256, 255);
    }
    wave_lds_sync();
    if (!active) return;

    // Row `lane` of the block: prediction (0 for intra) + residual, clamped to 0..255. Each
    // direction is formed only where a lane of the wave predicts from it; the average of both is
    // lerp(f, b, 1), a single direction is lerp(x, x, 0).
    const RefFetch ff = fetch_ref(fref, stride, pitch, fh, x0, ry, fmx, fmy, has_f, job.refs_aligned != 0);
    const RefFetch fb = fetch_ref(bref, stride, pitch, bh, x0, ry, bmx, bmy, has_b, job.refs_aligned != 0);
    uint32_t pf[2] = {0u, 0u}, pb[2] = {0u, 0u};
    if (has_f) {
        const uint2 t = halfpel_row(ff);
        pf[0] = t.x;
        pf[1] = t.y;
    }
    if (has_b) {
        const uint2 t = halfpel_row(fb);
        pb[0] = t.x;
        pb[1] = t.y;
    }
    const uint32_t rb = (has_f && has_b) ? kOnes : 0u;
    uint32_t o32[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) o32[h] = __builtin_amdgcn_lerp(has_f ? pf[h] : pb[h], has_b ? pb[h] : pf[h], rb);
    if (coded) {   // prediction + residual as 16-bit pairs, clamped to 0..255
        const int4 rr = *reinterpret_cast<const int4*>(blk + lane * 8);
        const uint32_t r32[4] = {(uint32_t)rr.x, (uint32_t)rr.y, (uint32_t)rr.z, (uint32_t)rr.w};
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            uint32_t sums[2];
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                // prediction bytes 2k, 2k + 1 zero-extended to 16 bits
                const uint32_t pw = __builtin_amdgcn_perm(0u, o32[h], k ? 0x0c030c02u : 0x0c010c00u);
                mmr_v2i16 pv, rv;
                __builtin_memcpy(&pv, &pw, 4);
                __builtin_memcpy(&rv, &r32[h * 2 + k], 4);
                mmr_v2i16 v = pv + rv;                                    // |v| <= 511: no overflow
                v = __builtin_elementwise_max(v, (mmr_v2i16){0, 0});
                v = __builtin_elementwise_min(v, (mmr_v2i16){255, 255});
                __builtin_memcpy(&sums[k], &v, 4);
            }
            o32[h] = __builtin_amdgcn_perm(sums[1], sums[0], 0x06040200u);   // low bytes of the four halves
        }
    }
    store_row8(pr.dst + (uint32_t)(yy * stride + x0), (uint64_t)o32[0] | ((uint64_t)o32[1] << 32));
}

hipError_t launch_mpeg2_order(const Mpeg2Job* d_jobs, uint32_t job_count, uint32_t total_rows, uint32_t* d_status,
                              hipStream_t s) {
    if (job_count == 0 || total_rows == 0) return hipSuccess;
    const uint32_t waves = kOrderThreads / 64;
    hipLaunchKernelGGL(mpeg2_order_kernel, dim3((total_rows + waves - 1) / waves), dim3(kOrderThreads), 0, s, d_jobs, job_count,
                       total_rows, d_status);
    return hipGetLastError();
}

hipError_t launch_mpeg2(const Mpeg2Job* d_jobs, const uint32_t* d_tile_job, uint32_t tile_base, uint32_t tiles,
                        uint32_t group_blocks, uint32_t* d_status, hipStream_t s) {
    if (tiles == 0) return hipSuccess;
    const uint32_t per = xcd_per(tiles);
    switch (group_blocks) {
    case 16: hipLaunchKernelGGL(mpeg2_kernel<16>, dim3(per * kXcds), dim3(16 * 8), 0, s, d_jobs, d_tile_job, tile_base, tiles, per, d_status); break;
    case 32: hipLaunchKernelGGL(mpeg2_kernel<32>, dim3(per * kXcds), dim3(32 * 8), 0, s, d_jobs, d_tile_job, tile_base, tiles, per, d_status); break;
    case 64: hipLaunchKernelGGL(mpeg2_kernel<64>, dim3(per * kXcds), dim3(64 * 8), 0, s, d_jobs, d_tile_job, tile_base, tiles, per, d_status); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace mmr
