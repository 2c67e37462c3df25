// H.264 loop filter over frame pictures as an edge-order wavefront inside one workgroup.
//
// Reference: filter_picture / derive_boundary_strengths / boundary_strength / motion_pairs_match /
// filter_luma_macroblock / filter_chroma_macroblock / filter_vertical/horizontal_edge /
// filter_samples / strong_filter / normal_filter / ALPHA / BETA / TC0
// (crates/codecs/h264/src/deblock.rs:3-587).
//
// Ordering. The reference filters macroblocks in raster order. Macroblock (x, y) reads and writes
// samples of (x-1, y) and (x, y-1), and (x+1, y-1) writes samples that (x, y) reads, so (x, y) may
// run once (x-1, y) and (x+1, y-1) are done: a wavefront over t = x + 2y.
//
// Mapping. Pictures are cut into bands of 16 macroblock rows; a 1024-thread workgroup filters one
// band at a time (band tickets from a device counter, in picture and band order) and wave w owns
// row w of the band, walking it left to right with a private LDS window (20x20 luma, 2 x 12x12
// chroma: 4 samples of left / top context). The left context stays in the window (shifted by one
// macroblock per step). Inside a band the top context comes from the wave of the row above
// through an LDS ring: after its step x that wave stores the bottom four sample rows of macroblock
// x (and completes those of x-1, whose right columns step x modified) and publishes progress x+1;
// the row below reads slot x once progress >= x+2. A writer waits before reusing a ring slot until
// the reader has consumed it (progress of the row below >= x - ring + 2); waits between adjacent
// rows cannot form a cycle (ring >= 4). Across bands the hand-over goes through the picture in
// global memory: the band's last row writes its bottom rows as well and publishes its progress
// with an agent-scope release every kDbPublish macroblocks; the next band's first row waits for
// progress >= x+2 with relaxed loads, then an agent-scope acquire, and reads the four rows above
// from the picture. Tickets are taken in order, so the band a workgroup waits for is held by a
// running workgroup (or done), and no band waits for a later one.
//
// Write-back. Each row writes its final samples once: luma rows y*16-3 .. y*16+12 (chroma
// y*8-3 .. y*8+4) of columns x*16-4 .. x*16+11 after step x (the rows below are handed to and
// written by the next row, which modifies them; the last row / column write the rest). The
// samples a step writes are not touched by any other row afterwards.
#include "device_common.h"
#include "recon_internal.h"

namespace mmr {

__constant__ static const int kAlpha[52] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 4, 4, 5, 6, 7, 8, 9, 10, 12, 13,
                                            15, 17, 20, 22, 25, 28, 32, 36, 40, 45, 50, 56, 63, 71, 80, 90, 101, 113, 127, 144, 162, 182, 203, 226, 255, 255};
__constant__ static const int kBeta[52] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4,
                                           6, 6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12, 13, 13, 14, 14, 15, 15, 16, 16, 17, 17, 18, 18};
__constant__ static const int kTc0[3][52] = {
    {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 4, 4, 4, 5, 6, 6, 7, 8, 9, 10, 11, 13},
    {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 4, 4, 5, 5, 6, 7, 8, 8, 10, 11, 12, 13, 15, 17},
    {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 4, 4, 4, 5, 6, 6, 7, 8, 9, 10, 11, 13, 14, 16, 18, 20, 23, 25}};
__constant__ static const int kDbChromaQp[22] = {29, 30, 31, 32, 32, 33, 34, 34, 35, 35, 36, 36, 37, 37, 37, 38, 38, 38, 39, 39, 39, 39};

// Ring slots per wave: a row runs at most kDbRing-1 macroblocks ahead of the row below.
constexpr int kDbRing = 16;
// Row stride of the chroma windows: the luma one's, so one code path filters luma and chroma
// lines (rows 12..19 and columns 12..19 of a chroma window are padding).
constexpr int kLc = 20;

// Per-workgroup LDS copies of the tables: tab[i] = ALPHA[i] | BETA[i] << 8 | TC0[0][i] << 16 |
// TC0[1][i] << 21 | TC0[2][i] << 26, cqp[i] = the chroma QP of luma QP index i. Per-lane lookups in
// __constant__ memory would be vector memory loads on every filtered edge.
struct DbTables {
    uint32_t tab[52];
    int32_t cqp[52];
};
MMR_DEV void db_load_tables(DbTables& t) {
    for (int i = threadIdx.x; i < 52; i += blockDim.x) {
        t.tab[i] = (uint32_t)kAlpha[i] | ((uint32_t)kBeta[i] << 8) | ((uint32_t)kTc0[0][i] << 16) | ((uint32_t)kTc0[1][i] << 21) |
                   ((uint32_t)kTc0[2][i] << 26);
        t.cqp[i] = i < 30 ? i : kDbChromaQp[i - 30];
    }
}
MMR_DEV int db_chroma_qp(const DbTables& t, int luma_qp, int offset) { return t.cqp[clampi(luma_qp + offset, 0, 51)]; }

// Filter parameters of one edge class: alpha and TC0 from index A = clamp(qp + offset_a), beta
// from index B = clamp(qp + offset_b).
struct EdgeParams {
    uint32_t a, b;   // tab[A], tab[B]
};
MMR_DEV EdgeParams edge_params(const DbTables& t, int qp, int offset_a, int offset_b) {
    EdgeParams e;
    e.a = t.tab[clampi(qp + offset_a, 0, 51)];
    e.b = t.tab[clampi(qp + offset_b, 0, 51)];
    return e;
}
MMR_DEV int iabs(int v) { return v < 0 ? -v : v; }

// LDS ordering between the lanes of one wave (the wave issues LDS operations in order; this keeps
// the compiler from moving accesses across the point and waits for the outstanding ones).
MMR_DEV void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// filter_samples + strong_filter + normal_filter (deblock.rs:469-569) on one line of samples held in
// registers as packed bytes: sample i is byte i of the line, p_i is sample k - 1 - i and q_i sample
// k + i (k is a compile-time index, so every extract / insert is one bitfield instruction).
// Samples are only rewritten when the filter applies, which leaves the same values as the
// reference's unconditional write-back of unchanged samples.
template <int NW>
MMR_DEV int get_byte(const uint32_t (&w)[NW], int i) { return (w[i >> 2] >> (8 * (i & 3))) & 0xff; }
template <int NW>
MMR_DEV void set_byte(uint32_t (&w)[NW], int i, int v) {
    const uint32_t sh = 8 * (i & 3);
    w[i >> 2] = (w[i >> 2] & ~(0xffu << sh)) | ((uint32_t)v << sh);
}

template <int k>
MMR_DEV void filter_line(uint32_t (&w)[5], EdgeParams ep, int bs, bool chroma) {
    const int alpha = (int)(ep.a & 0xffu), beta = (int)((ep.b >> 8) & 0x1fu);
    const int p0 = get_byte(w, k - 1), p1 = get_byte(w, k - 2), p2 = get_byte(w, k - 3), p3 = get_byte(w, k - 4);
    const int q0 = get_byte(w, k), q1 = get_byte(w, k + 1), q2 = get_byte(w, k + 2), q3 = get_byte(w, k + 3);
    if (bs == 0 || iabs(p0 - q0) >= alpha || iabs(p1 - p0) >= beta || iabs(q1 - q0) >= beta) return;
    // Both filters are evaluated and the results selected (no branches inside a wave).
    const bool luma = !chroma;
    const bool ap = iabs(p2 - p0) < beta, aq = iabs(q2 - q0) < beta;
    // bS < 4
    const int tc0 = (int)((ep.a >> (16 + 5 * (bs - 1))) & 0x1fu);   // bs 1..3 (unused for 4)
    const int tc = tc0 + (chroma ? 1 : ((int)ap + (int)aq));
    const int delta = clampi((((q0 - p0) << 2) + (p1 - q1) + 4) >> 3, -tc, tc);
    const int avg = (p0 + q0 + 1) >> 1;
    const int np0 = clampi(p0 + delta, 0, 255), nq0 = clampi(q0 - delta, 0, 255);
    const int np1 = (luma && ap) ? clampi(p1 + clampi((p2 + avg - (p1 << 1)) >> 1, -tc0, tc0), 0, 255) : p1;
    const int nq1 = (luma && aq) ? clampi(q1 + clampi((q2 + avg - (q1 << 1)) >> 1, -tc0, tc0), 0, 255) : q1;
    // bS == 4
    const bool near = iabs(p0 - q0) < (alpha >> 2) + 2;
    const bool sp = luma && ap && near, sq = luma && aq && near;
    const int sp0 = sp ? (p2 + 2 * p1 + 2 * p0 + 2 * q0 + q1 + 4) >> 3 : (2 * p1 + p0 + q1 + 2) >> 2;
    const int sp1 = sp ? (p2 + p1 + p0 + q0 + 2) >> 2 : p1;
    const int sp2 = sp ? (2 * p3 + 3 * p2 + p1 + p0 + q0 + 4) >> 3 : p2;
    const int sq0 = sq ? (p1 + 2 * p0 + 2 * q0 + 2 * q1 + q2 + 4) >> 3 : (2 * q1 + q0 + p1 + 2) >> 2;
    const int sq1 = sq ? (p0 + q0 + q1 + q2 + 2) >> 2 : q1;
    const int sq2 = sq ? (2 * q3 + 3 * q2 + q1 + q0 + p0 + 4) >> 3 : q2;
    const bool strong = bs == 4;
    set_byte(w, k - 3, strong ? sp2 : p2);
    set_byte(w, k - 2, strong ? sp1 : np1);
    set_byte(w, k - 1, strong ? sp0 : np0);
    set_byte(w, k, strong ? sq0 : nq0);
    set_byte(w, k + 1, strong ? sq1 : nq1);
    set_byte(w, k + 2, strong ? sq2 : q2);
}

struct RefMotion {
    int x, y, index, poc;
    bool has_ref;
};

MMR_DEV RefMotion motion_of(const H264Job& job, const mmr_recon_h264_block& b, int list) {
    RefMotion m;
    m.x = 0; m.y = 0; m.index = -1; m.poc = 0; m.has_ref = false;   // MotionInfo::default() when unused
    if (b.ref_idx[list] >= 0) {
        m.x = b.mv[list][0];
        m.y = b.mv[list][1];
        m.index = b.ref_idx[list];
        m.poc = job.ref_poc[list][min((int)b.ref_idx[list], (int)MMR_H264_MAX_REFS - 1)];
        m.has_ref = true;
    }
    return m;
}
MMR_DEV bool vec_match(const RefMotion& a, const RefMotion& b) { return iabs(a.x - b.x) < 4 && iabs(a.y - b.y) < 4; }
MMR_DEV bool ref_match(const RefMotion& a, const RefMotion& b) { return a.has_ref == b.has_ref && (!a.has_ref || a.poc == b.poc); }
MMR_DEV bool motion_match(const RefMotion& a, const RefMotion& b) { return ref_match(a, b) && vec_match(a, b); }
MMR_DEV bool indexed_match(const RefMotion& a, const RefMotion& b) { return a.index == b.index && vec_match(a, b); }

// boundary_strength + motion_pairs_match (deblock.rs:391-439) from prefetched records.
MMR_DEV int boundary_strength(const H264Job& job, uint8_t p_flags, uint8_t q_flags, const mmr_recon_h264_block& P,
                              const mmr_recon_h264_block& Q, bool mb_edge) {
    if ((p_flags | q_flags) & MMR_H264_MB_INTRA) return mb_edge ? 4 : 3;
    if (P.total_coeff != 0 || Q.total_coeff != 0) return 2;
    const RefMotion l0 = motion_of(job, P, 0), l1 = motion_of(job, P, 1), r0 = motion_of(job, Q, 0), r1 = motion_of(job, Q, 1);
    bool match;
    if (!l1.has_ref && !r1.has_ref) match = indexed_match(l0, r0);
    else if (!l0.has_ref && !r0.has_ref) match = indexed_match(l1, r1);
    else match = (motion_match(l0, r0) && motion_match(l1, r1)) || (motion_match(l0, r1) && motion_match(l1, r0));
    return match ? 0 : 1;
}

MMR_DEV int luma_block_index(int x, int y) { return (y / 2) * 8 + (x / 2) * 4 + (y % 2) * 2 + x % 2; }

struct DbWave {
    uint8_t ly[20 * 20];               // rows y*16-4 .., cols x*16-4 ..
    uint8_t lc[2][20 * kLc];           // rows y*8-4 .., cols x*8-4 .. (12 x 12 used)
    uint8_t ring_y[kDbRing][4 * 16];   // bottom 4 luma rows of a macroblock, for the row below
    uint8_t ring_c[kDbRing][2][4 * 8];
};

// Per-lane prefetch of one macroblock step: its record, own samples and boundary strengths.
struct DbPrefetch {
    mmr_recon_h264_macroblock m;   // record of the macroblock
    int qp_top;            // QP of the macroblock above (0 in the first row)
    uint32_t own_y;        // luma dword (row l>>2, cols 4*(l&3)) of the macroblock
    uint32_t own_c;        // chroma dword, lanes < 32 (comp l>>4, row (l>>1)&7, cols 4*(l&1))
    int4 bs;               // the macroblock's 32 strengths (h264_bs_kernel)
};

// Records read through the global address space (the job holds generic pointers).
MMR_DEV mmr_recon_h264_macroblock db_mb(const H264Job& job, uint32_t a) {
    const MMR_GLOBAL uint32_t* q = reinterpret_cast<const MMR_GLOBAL uint32_t*>(as_global(job.mbs + a));
    const uint32_t w[3] = {q[0], q[1], q[2]};
    mmr_recon_h264_macroblock m;
    __builtin_memcpy(&m, w, sizeof m);
    return m;
}
MMR_DEV mmr_recon_h264_block db_block(const H264Job& job, size_t i) {
    const MMR_GLOBAL uint32_t* q = reinterpret_cast<const MMR_GLOBAL uint32_t*>(as_global(job.blocks + i));
    const uint32_t w[3] = {q[0], q[1], q[2]};
    mmr_recon_h264_block b;
    __builtin_memcpy(&b, w, sizeof b);
    return b;
}

MMR_DEV void db_prefetch(const H264Job& job, int MW, int x, int y, int l, DbPrefetch& f) {
    const int W = MW * 16, CW = MW * 8;
    const uint32_t a = (uint32_t)(y * MW + x);
    f.m = db_mb(job, a);
    f.qp_top = y > 0 ? as_global(job.mbs)[a - MW].qp : 0;
    f.bs = load_global_int4(job.bs + (size_t)a * 4);
    f.own_y = *reinterpret_cast<const MMR_GLOBAL uint32_t*>(as_global(job.planes[0]) + (size_t)(y * 16 + (l >> 2)) * W + x * 16 + 4 * (l & 3));
    if (l < 32) {
        const int comp = l >> 4, row = (l >> 1) & 7;
        f.own_c = *reinterpret_cast<const MMR_GLOBAL uint32_t*>(as_global(job.planes[1 + comp]) + (size_t)(y * 8 + row) * CW + x * 8 + 4 * (l & 1));
    }
}

// Strength n of a macroblock's record: n = e * 4 + segment for vertical edge e, 16 + e * 4 +
// segment for horizontal edge e; four bits each, eight per dword.
MMR_DEV int bs_nibble(const int4& r, int n) {
    const uint32_t w = (uint32_t)(n < 8 ? r.x : n < 16 ? r.y : n < 24 ? r.z : r.w);
    return (int)((w >> (4 * (n & 7))) & 0xfu);
}

// derive_boundary_strengths (deblock.rs:144-200) for every macroblock of the deblocked pictures,
// ahead of the wavefront: one lane per (macroblock, edge segment), lanes 0-15 of a 32-lane group
// the vertical edges (e = lane >> 2, segment lane & 3), 16-31 the horizontal ones. Edges that are
// not filtered get strength 0: external edges on the picture border or across a slice edge with
// filtering off there, the internal edges 1 and 3 of 8x8-transform macroblocks, and every edge of
// a macroblock with filtering disabled. Writes 16 bytes per macroblock (job.bs).
__global__ __launch_bounds__(128) void h264_bs_kernel(const H264Job* __restrict__ jobs, const uint32_t* __restrict__ tile_job,
                                                      uint32_t tile_base, uint32_t tiles, uint32_t per_xcd) {
    const uint32_t local = xcd_tile(blockIdx.x, per_xcd);
    if (local >= tiles) return;
    const uint32_t tile = tile_base + local;
    const H264Job& job = jobs[tile_job[tile]];
    if (!job.deblock) return;
    const uint32_t a = (tile - job.tile_begin) * 4 + (threadIdx.x >> 5);   // uniform over 32 lanes
    const int MW = (int)job.width_mbs;
    if (a >= (uint32_t)MW * job.height_mbs) return;
    const int x = (int)(a % (uint32_t)MW), y = (int)(a / (uint32_t)MW), l = threadIdx.x & 31;
    const mmr_recon_h264_macroblock q = db_mb(job, a);
    const bool vertical = l < 16;
    const int e = (l & 15) >> 2, sgm = l & 3;
    const bool external = e == 0;
    bool on = !(q.flags & MMR_H264_MB_NO_DEBLOCK) && !(e & 1 && (q.flags & MMR_H264_MB_TRANSFORM_8X8));
    uint32_t pm = a;
    if (external) {
        on = on && (vertical ? x > 0 : y > 0);
        pm = on ? (vertical ? a - 1 : a - (uint32_t)MW) : a;
        if (on)   // external_edge_enabled
            on = !(q.flags & MMR_H264_MB_SLICE_EDGE_OFF) || db_mb(job, pm).slice_id == q.slice_id;
    }
    int bs = 0;
    if (on) {
        const int pb = vertical ? luma_block_index(external ? 3 : e - 1, sgm) : luma_block_index(sgm, external ? 3 : e - 1);
        const int qb = vertical ? luma_block_index(e, sgm) : luma_block_index(sgm, e);
        const mmr_recon_h264_block P = db_block(job, (size_t)pm * 16 + pb), Q = db_block(job, (size_t)a * 16 + qb);
        bs = boundary_strength(job, as_global(job.mbs)[pm].flags, q.flags, P, Q, external);
    }
    const uint32_t packed = group8_or((uint32_t)bs << (4 * (l & 7)));
    if ((l & 7) == 0) as_global(job.bs)[(size_t)a * 4 + (l >> 3)] = packed;
}

hipError_t launch_h264_bs(const H264Job* d_jobs, const uint32_t* d_tile_job, uint32_t tile_base, uint32_t tiles, hipStream_t s) {
    if (tiles == 0) return hipSuccess;
    const uint32_t per = xcd_per(tiles);
    hipLaunchKernelGGL(h264_bs_kernel, dim3(per * kXcds), dim3(128), 0, s, d_jobs, d_tile_job, tile_base, tiles, per);
    return hipGetLastError();
}

// Bounded wait on a workgroup progress counter: gives up (and reports) after about a second, so a
// protocol error surfaces as a status instead of a hung kernel. Rows of a band hand over only
// through LDS (the ring and these counters): a row reads no sample another row of its band wrote
// in global memory (its own macroblocks come from earlier kernels, the band above's rows through
// the agent-scope hand-over), so the acquire and the matching release cover LDS only and the
// release does not wait for the write-back of the picture.
constexpr uint32_t kDbSpinLimit = 1u << 25;
MMR_DEV bool db_wait(uint32_t* counter, uint32_t need, uint32_t* status) {
    uint32_t spins = 0;
    while (__hip_atomic_load(counter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) < need) {
        if (++spins > kDbSpinLimit) {
            atomicOr(status, kDevStatusTimeout);
            return false;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
    return true;
}

// Wait for the previous band's published progress (another workgroup): relaxed agent-scope loads,
// then one agent-scope acquire, so the rows that band wrote before its release are visible. `seen`
// keeps the last value observed (and acquired).
MMR_DEV bool db_wait_band(uint32_t* counter, uint32_t need, uint32_t& seen, uint32_t* status) {
    if (seen >= need) return true;
    uint32_t spins = 0;
    for (;;) {
        seen = __hip_atomic_load(counter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (seen >= need) break;
        if (++spins > kDbSpinLimit) {
            atomicOr(status, kDevStatusTimeout);
            return false;
        }
        __builtin_amdgcn_s_sleep(2);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    return true;
}

// One macroblock row. band_first: the top context comes from the picture (gprev: the previous
// band's progress); band_last: no row below in this band (gself: progress for the next band, or
// null for the picture's last row) -- the row writes its bottom sample rows itself.
template <int ROWS>
MMR_DEV void db_row(const H264Job& job, const DbTables& T, int y, int w, int l, DbWave* waves, uint32_t* prog, uint32_t* status,
                    bool band_first, bool band_last, uint32_t* gprev, uint32_t* gself, uint32_t publish) {
    DbWave& S = waves[w];
    const DbWave& Above = waves[(w + ROWS - 1) % ROWS];
    const int MW = (int)job.width_mbs, MH = (int)job.height_mbs;
    const int W = MW * 16, CW = MW * 8;
    const bool last_row = band_last;   // writes its own bottom rows
    uint32_t seen = 0;
    DbPrefetch cur, nxt;
    db_prefetch(job, MW, 0, y, l, cur);
    // Lanes 0-15 filter luma lines, 16-23 Cb and 24-31 Cr lines (rows in the vertical pass, columns
    // in the horizontal one); their QPs go through the chroma QP table.
    const bool chroma_lane = l >= 16;
    const int chroma_off = job.chroma_qp_offset[(l >> 3) & 1];
    int qp_left = 0;
    for (int x = 0; x < MW; ++x) {
        const mmr_recon_h264_macroblock m = cur.m;
        const bool filter = !(m.flags & MMR_H264_MB_NO_DEBLOCK);
        // filter parameters of the left edge, the internal edges and the top edge
        auto lane_qp = [&](int qp) { return chroma_lane ? db_chroma_qp(T, qp, chroma_off) : qp; };
        const int qpc = lane_qp(m.qp);
        const EdgeParams ev = edge_params(T, (lane_qp(qp_left) + qpc + 1) >> 1, m.offset_a, m.offset_b);
        const EdgeParams ei = edge_params(T, qpc, m.offset_a, m.offset_b);
        const EdgeParams eh = edge_params(T, (lane_qp(cur.qp_top) + qpc + 1) >> 1, m.offset_a, m.offset_b);
        // own samples into the window (rows 4.., cols 4..)
        *reinterpret_cast<uint32_t*>(&S.ly[(4 + (l >> 2)) * 20 + 4 + 4 * (l & 3)]) = cur.own_y;
        if (l < 32) {
            const int comp = l >> 4, row = (l >> 1) & 7;
            *reinterpret_cast<uint32_t*>(&S.lc[comp][(4 + row) * kLc + 4 + 4 * (l & 1)]) = cur.own_c;
        }
        if (x + 1 < MW) db_prefetch(job, MW, x + 1, y, l, nxt);
        // top context from the row above (slot x), after it finished macroblock x + 1
        if (y > 0 && band_first) {
            if (!db_wait_band(gprev, (uint32_t)min(x + 2, MW), seen, status)) return;
            if (l < 16) {
                *reinterpret_cast<uint32_t*>(&S.ly[(l >> 2) * 20 + 4 + 4 * (l & 3)]) = *reinterpret_cast<const MMR_GLOBAL uint32_t*>(
                    as_global(job.planes[0]) + (size_t)(y * 16 - 4 + (l >> 2)) * W + x * 16 + 4 * (l & 3));
            } else if (l < 32) {
                const int k = l - 16, comp = k >> 3, row = (k >> 1) & 3;
                *reinterpret_cast<uint32_t*>(&S.lc[comp][row * kLc + 4 + 4 * (k & 1)]) = *reinterpret_cast<const MMR_GLOBAL uint32_t*>(
                    as_global(job.planes[1 + comp]) + (size_t)(y * 8 - 4 + row) * CW + x * 8 + 4 * (k & 1));
            }
        } else if (y > 0) {
            if (!db_wait(&prog[w - 1], (uint32_t)min(x + 2, MW), status)) return;
            const int slot = x % kDbRing;
            if (l < 16) {
                *reinterpret_cast<uint32_t*>(&S.ly[(l >> 2) * 20 + 4 + 4 * (l & 3)]) =
                    *reinterpret_cast<const uint32_t*>(&Above.ring_y[slot][(l >> 2) * 16 + 4 * (l & 3)]);
            } else if (l < 32) {
                const int k = l - 16, comp = k >> 3, row = (k >> 1) & 3;
                *reinterpret_cast<uint32_t*>(&S.lc[comp][row * kLc + 4 + 4 * (k & 1)]) =
                    *reinterpret_cast<const uint32_t*>(&Above.ring_c[slot][comp][row * 8 + 4 * (k & 1)]);
            }
        }
        wave_sync();
        if (filter && l < 32) {
            // One line per lane, in registers: lanes 0-15 luma (rows in the vertical pass,
            // columns in the horizontal one: 4 samples of context + 16), 16-23 Cb and 24-31 Cr (4 +
            // 8). Edge slots k = 4, 8, 12, 16 are the luma edges 0..3 and the chroma edges 0 and 2
            // (slots 12, 16 have strength 0 on chroma lines). Strengths come from the macroblock's
            // record: vertical edge e, segment g at nibble e*4+g of dwords x / y, horizontal at
            // nibble e*4+g of dwords z / w.
            const int li = chroma_lane ? ((l - 16) & 7) : l;
            uint8_t* win = chroma_lane ? S.lc[(l - 16) >> 3] : S.ly;
            const uint32_t sh = 4u * (uint32_t)(chroma_lane ? (li >> 1) : (li >> 2));
            const int4 r = cur.bs;
            auto nib = [&](int v, uint32_t shift) { return (int)(((uint32_t)v >> shift) & 0xfu); };
            const int bv[4] = {nib(r.x, sh), chroma_lane ? nib(r.y, sh) : nib(r.x, sh + 16), chroma_lane ? 0 : nib(r.y, sh),
                               chroma_lane ? 0 : nib(r.y, sh + 16)};
            const int bh[4] = {nib(r.z, sh), chroma_lane ? nib(r.w, sh) : nib(r.z, sh + 16), chroma_lane ? 0 : nib(r.w, sh),
                               chroma_lane ? 0 : nib(r.w, sh + 16)};
            {   // vertical edges
                uint32_t* row = reinterpret_cast<uint32_t*>(win + (4 + li) * 20);
                uint32_t v[5];
#pragma unroll
                for (int i = 0; i < 5; ++i) v[i] = row[i];
                filter_line<4>(v, ev, bv[0], chroma_lane);
                filter_line<8>(v, ei, bv[1], chroma_lane);
                filter_line<12>(v, ei, bv[2], chroma_lane);
                filter_line<16>(v, ei, bv[3], chroma_lane);
#pragma unroll
                for (int i = 0; i < 5; ++i) row[i] = v[i];
            }
            wave_sync();
            {   // horizontal edges
                uint8_t* col = win + 4 + li;
                uint32_t v[5];
#pragma unroll
                for (int i = 0; i < 5; ++i)
                    v[i] = col[(4 * i) * 20] | (col[(4 * i + 1) * 20] << 8) | (col[(4 * i + 2) * 20] << 16) | ((uint32_t)col[(4 * i + 3) * 20] << 24);
                filter_line<4>(v, eh, bh[0], chroma_lane);
                filter_line<8>(v, ei, bh[1], chroma_lane);
                filter_line<12>(v, ei, bh[2], chroma_lane);
                filter_line<16>(v, ei, bh[3], chroma_lane);
#pragma unroll
                for (int i = 1; i < 19; ++i) col[i * 20] = (uint8_t)get_byte(v, i);
            }
        }
        wave_sync();
        // hand the bottom rows to the row below: slot x (all columns) and the right columns of
        // slot x - 1, which this step's left edge modified
        if (!band_last) {
            const int free_at = x - kDbRing + 2;
            if (free_at > 0 && !db_wait(&prog[w + 1], (uint32_t)free_at, status)) return;
            const int slot = x % kDbRing, prev = (x + kDbRing - 1) % kDbRing;
            if (l < 16) {
                *reinterpret_cast<uint32_t*>(&S.ring_y[slot][(l >> 2) * 16 + 4 * (l & 3)]) =
                    *reinterpret_cast<const uint32_t*>(&S.ly[(16 + (l >> 2)) * 20 + 4 + 4 * (l & 3)]);
            } else if (l < 32) {
                const int k = l - 16, comp = k >> 3, row = (k >> 1) & 3;
                *reinterpret_cast<uint32_t*>(&S.ring_c[slot][comp][row * 8 + 4 * (k & 1)]) =
                    *reinterpret_cast<const uint32_t*>(&S.lc[comp][(8 + row) * kLc + 4 + 4 * (k & 1)]);
            } else if (x > 0 && l < 36) {
                const int row = l - 32;
                *reinterpret_cast<uint32_t*>(&S.ring_y[prev][row * 16 + 12]) = *reinterpret_cast<const uint32_t*>(&S.ly[(16 + row) * 20]);
            } else if (x > 0 && l >= 40 && l < 48) {
                const int k = l - 40, comp = k >> 2, row = k & 3;
                *reinterpret_cast<uint32_t*>(&S.ring_c[prev][comp][row * 8 + 4]) = *reinterpret_cast<const uint32_t*>(&S.lc[comp][(8 + row) * kLc]);
            }
        }
        // write back the samples that are now final
        {
            // luma: rows y*16-3 .. y*16+12 (window rows 1..16), columns x*16-4 .. x*16+11 (window cols 0..15)
            const int r = 1 + (l >> 2), c4 = 4 * (l & 3);
            const int gy = y * 16 - 4 + r, gx = x * 16 - 4 + c4;
            if (gy >= 0 && gx >= 0)
                *reinterpret_cast<MMR_GLOBAL uint32_t*>(as_global(job.planes[0]) + (size_t)gy * W + gx) = *reinterpret_cast<const uint32_t*>(&S.ly[r * 20 + c4]);
            if (last_row && l < 12) {   // rows y*16+13 .. +15 (window rows 17..19)
                const int r2 = 17 + (l >> 2), gx2 = x * 16 - 4 + c4;
                if (gx2 >= 0)
                    *reinterpret_cast<MMR_GLOBAL uint32_t*>(as_global(job.planes[0]) + (size_t)(y * 16 - 4 + r2) * W + gx2) =
                        *reinterpret_cast<const uint32_t*>(&S.ly[r2 * 20 + c4]);
            }
            if (x == MW - 1 && l < 19) {   // last column: columns x*16+12 .. +15 (window col 16)
                const int r3 = 1 + l;
                const int gy3 = y * 16 - 4 + r3;
                if (gy3 >= 0 && (r3 <= 16 || last_row))
                    *reinterpret_cast<MMR_GLOBAL uint32_t*>(as_global(job.planes[0]) + (size_t)gy3 * W + x * 16 + 12) =
                        *reinterpret_cast<const uint32_t*>(&S.ly[r3 * 20 + 16]);
            }
            // chroma: rows y*8-3 .. y*8+4 (window rows 1..8), columns x*8-4 .. x*8+3 (window cols 0..7)
            if (l < 32) {
                const int comp = l >> 4, rr = 1 + ((l >> 1) & 7), cc = 4 * (l & 1);
                const int cy = y * 8 - 4 + rr, cx = x * 8 - 4 + cc;
                if (cy >= 0 && cx >= 0)
                    *reinterpret_cast<MMR_GLOBAL uint32_t*>(as_global(job.planes[1 + comp]) + (size_t)cy * CW + cx) =
                        *reinterpret_cast<const uint32_t*>(&S.lc[comp][rr * kLc + cc]);
            }
            // chroma extras: last row, rows y*8+5 .. +7 (window rows 9..11, 12 dwords); last column,
            // columns x*8+4 .. +7 (window col 8, 2 x 11 rows)
            const int extra_rows = last_row ? 12 : 0, extra_cols = x == MW - 1 ? 22 : 0;
            for (int i = l; i < extra_rows + extra_cols; i += 64) {
                if (i < extra_rows) {
                    const int comp = i / 6, rr = 9 + (i % 6) / 2, cc = 4 * (i & 1);
                    const int cx = x * 8 - 4 + cc;
                    if (cx >= 0)
                        *reinterpret_cast<MMR_GLOBAL uint32_t*>(as_global(job.planes[1 + comp]) + (size_t)(y * 8 - 4 + rr) * CW + cx) =
                            *reinterpret_cast<const uint32_t*>(&S.lc[comp][rr * kLc + cc]);
                } else {
                    const int k = i - extra_rows, comp = k / 11, rr = 1 + k % 11;
                    const int cy = y * 8 - 4 + rr;
                    if (cy >= 0 && (rr <= 8 || last_row))
                        *reinterpret_cast<MMR_GLOBAL uint32_t*>(as_global(job.planes[1 + comp]) + (size_t)cy * CW + x * 8 + 4) =
                            *reinterpret_cast<const uint32_t*>(&S.lc[comp][rr * kLc + 8]);
                }
            }
        }
        wave_sync();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
        if (l == 0) __hip_atomic_store(&prog[w], (uint32_t)(x + 1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (gself && ((x + 1) % publish == 0 || x == MW - 1))
            __hip_atomic_store(gself, (uint32_t)(x + 1), __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        // shift the window left by one macroblock (its right 4 columns become the left context)
        if (l < 20) {
            *reinterpret_cast<uint32_t*>(&S.ly[l * 20]) = *reinterpret_cast<const uint32_t*>(&S.ly[l * 20 + 16]);
        } else if (l < 44) {
            const int k = l - 20, comp = k / 12, row = k % 12;
            *reinterpret_cast<uint32_t*>(&S.lc[comp][row * kLc]) = *reinterpret_cast<const uint32_t*>(&S.lc[comp][row * kLc + 8]);
        }
        wave_sync();
        qp_left = m.qp;
        cur = nxt;
    }
}

// ROWS macroblock rows per band, one wave per row.
template <int ROWS>
__global__ __launch_bounds__(ROWS * 64) void h264_deblock_kernel(const H264Job* __restrict__ jobs, uint32_t job_count,
                                                                 uint32_t total_bands, uint32_t* __restrict__ counters,
                                                                 uint32_t publish, uint32_t* __restrict__ status) {
    __shared__ __align__(16) DbWave waves[ROWS];
    __shared__ uint32_t prog[ROWS];
    __shared__ uint32_t s_ticket;
    __shared__ DbTables s_tables;
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    db_load_tables(s_tables);   // visible after the loop's first barrier
    for (;;) {
        if (threadIdx.x == 0) s_ticket = fetch_add_relaxed(&counters[0], 1);
        if (threadIdx.x < ROWS) prog[threadIdx.x] = 0;
        __syncthreads();
        const uint32_t t = s_ticket;
        if (t >= total_bands) break;
        const H264Job& job = jobs[find_job(jobs, job_count, t, [](const H264Job& j) { return j.band_begin; })];
        const uint32_t band = t - job.band_begin;
        const int y = (int)(band * ROWS) + w;
        if (job.deblock && y < (int)job.height_mbs) {
            const bool first = w == 0, last = w == ROWS - 1 || y == (int)job.height_mbs - 1;
            uint32_t* gprev = first && band > 0 ? &counters[1 + t - 1] : nullptr;
            uint32_t* gself = last && y < (int)job.height_mbs - 1 ? &counters[1 + t] : nullptr;
            db_row<ROWS>(job, s_tables, y, w, l, waves, prog, status, first, last, gprev, gself, publish);
        }
        __syncthreads();
    }
}

bool h264_deblock_rows_supported(uint32_t rows) { return rows == 4 || rows == 8 || rows == 16; }

hipError_t launch_h264_deblock(const H264Job* d_jobs, uint32_t job_count, uint32_t total_bands, uint32_t* d_counters,
                               uint32_t groups, uint32_t band_rows, uint32_t publish, uint32_t* d_status, hipStream_t s) {
    if (job_count == 0 || total_bands == 0) return hipSuccess;
    if (groups > total_bands) groups = total_bands;
    if (publish == 0) publish = 1;
    switch (band_rows) {
    case 4:
        hipLaunchKernelGGL(h264_deblock_kernel<4>, dim3(groups), dim3(4 * 64), 0, s, d_jobs, job_count, total_bands, d_counters,
                           publish, d_status);
        break;
    case 8:
        hipLaunchKernelGGL(h264_deblock_kernel<8>, dim3(groups), dim3(8 * 64), 0, s, d_jobs, job_count, total_bands, d_counters,
                           publish, d_status);
        break;
    case 16:
        hipLaunchKernelGGL(h264_deblock_kernel<16>, dim3(groups), dim3(16 * 64), 0, s, d_jobs, job_count, total_bands, d_counters,
                           publish, d_status);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace mmr
