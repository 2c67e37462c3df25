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
// Mapping. Pictures are cut into bands of ROWS (4, 8 or 16) macroblock rows; a workgroup of ROWS
// waves filters one band at a time (band tickets from a device counter, in picture and band order)
// and wave w owns row w of the band, walking it left to right with a private LDS window (20x20
// luma, 2 x 12x12 chroma: 4 samples of left / top context). The left context stays in the window
// (shifted by one macroblock per step). Inside a band the top context comes from the wave of the
// row above through an LDS ring: the bottom four sample rows of macroblock x are final for that
// row once its step x + 1 has filtered the vertical edges (the left edge of x + 1 changes their
// right columns), so right there, before its horizontal edges, the row stores them as slot x and
// publishes progress x + 1; the row below waits for progress >= x + 1 before the horizontal edges
// of its step x. Rows therefore trail each other by one step plus the hand-over, not by a whole
// step plus the horizontal half. A writer waits before reusing a ring slot until the reader has
// consumed it (progress of the row below >= x - ring + 1); waits between adjacent rows cannot
// form a cycle (ring >= 2). Across bands the hand-over goes through the picture in
// global memory: the band's last row writes its bottom rows as well, with stores coherent across
// the XCDs, and publishes its progress every `publish` macroblocks once they have completed; the
// next band's first row waits for the progress that covers macroblock x + 2 and reads the four
// rows above that macroblock from the picture, with coherent loads, two steps before it filters
// it (db_put / publish_agent / db_wait_band: no L2 write-back or invalidate).
// Tickets are taken in order, so the band a workgroup waits for is held by a running workgroup (or
// done), and no band waits for a later one.
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

// Filter parameters of one edge class for this lane's line: alpha and TC0 from index A =
// clamp(qp + offset_a), beta from index B = clamp(qp + offset_b). The thresholds are held negated,
// as the bias an absolute difference is tested with (sad_bias): na = -alpha, nb = -beta, nnear =
// -((alpha >> 2) + 2). nb_ap / nb_aq are -beta on luma lines and -256 / 0 on chroma lines, where
// the p2 / q2 tests then read always / never true: tc0 + [ap] + [aq] is then chroma's tc0 + 1,
// and the p1 / q1 updates and the strong filter, luma only, are masked by luma_m.
struct EdgeParams {
    uint32_t a;                        // tab[A]: TC0 for bS 1, 2, 3 at bits 16, 21, 26
    int na, nb, nb_ap, nb_aq, nnear;
};
MMR_DEV EdgeParams edge_params(const DbTables& t, int qp, int offset_a, int offset_b, int luma_m) {
    EdgeParams e;
    e.a = t.tab[clampi(qp + offset_a, 0, 51)];
    const uint32_t b = t.tab[clampi(qp + offset_b, 0, 51)];
    const int alpha = (int)(e.a & 0xffu), beta = (int)((b >> 8) & 0x1fu);
    e.na = -alpha;
    e.nb = -beta;
    e.nb_ap = (-beta & luma_m) | (-256 & ~luma_m);
    e.nb_aq = -beta & luma_m;
    e.nnear = -((alpha >> 2) + 2);
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

// Lane masks for the edge filter: 0 or -1 (all bits), built from sign bits so that the filter is
// straight-line vector arithmetic (no per-lane branches or exec-mask changes).
MMR_DEV int mask_sel(int m, int a, int b) { return (a & m) | (b & ~m); }      // m ? a : b
MMR_DEV int clamp_sym(int v, int t) { return min(max(v, -t), t); }           // t >= 0
// |a - b| + bias as a signed value, one v_sad_u16 (its 32-bit accumulator takes the bias): for a, b
// in 0..255 and bias in -256..0 it is negative exactly when |a - b| < -bias.
MMR_DEV int sad_bias(int a, int b, int bias) { return (int)__builtin_amdgcn_sad_u16((uint32_t)a, (uint32_t)b, (uint32_t)bias); }
// bS != 0 && |p0 - q0| < alpha && |p1 - p0| < beta && |q1 - q0| < beta as 0 / -1: the three tests'
// maximum is negative when all hold, and -bS is negative when bS != 0.
MMR_DEV int edge_on(int p1, int p0, int q0, int q1, const EdgeParams& e, int bs) {
    return (max(max(sad_bias(p0, q0, e.na), sad_bias(p1, p0, e.nb)), sad_bias(q1, q0, e.nb)) & -bs) >> 31;
}
// (p0 + q0 + 1) >> 1 for p0, q0 in 0..255: one v_lerp_u8 (byte 0 with its rounding bit set).
MMR_DEV int avg_round(int p0, int q0) { return (int)__builtin_amdgcn_lerp((uint32_t)p0, (uint32_t)q0, 1u); }

// filter_samples + strong_filter + normal_filter (deblock.rs:469-569) on one line of samples, one
// sample per register: p_i is s[k - 1 - i] and q_i is s[k + i] (k is a compile-time index). The
// line's edges run one after another (each reads samples the previous one wrote), which makes
// this the wavefront's critical path, so it is straight-line code: both filters are evaluated and
// the results selected with lane masks; samples keep their values where the filter does not apply
// (the reference writes unchanged samples back, which leaves the same values). luma_m is -1 on
// luma lines, 0 on chroma lines.
template <int k>
MMR_DEV void filter_edge(int (&s)[20], const EdgeParams& e, int bs, int luma_m) {
    const int p0 = s[k - 1], p1 = s[k - 2], p2 = s[k - 3], p3 = s[k - 4];
    const int q0 = s[k], q1 = s[k + 1], q2 = s[k + 2], q3 = s[k + 3];
    const int on = edge_on(p1, p0, q0, q1, e, bs);
    // |p2 - p0| < beta, |q2 - q0| < beta (luma lines; chroma: -1 and 0)
    const int ap = sad_bias(p2, p0, e.nb_ap) >> 31, aq = sad_bias(q2, q0, e.nb_aq) >> 31;
    // bS < 4
    const int tc0 = (int)((e.a >> (16 + 5 * (bs - 1))) & 0x1fu);   // bS 1..3 (unused for 0 and 4)
    const int tc = tc0 - ap - aq;
    const int delta = clamp_sym((((q0 - p0) << 2) + (p1 - q1) + 4) >> 3, tc);
    const int avg = avg_round(p0, q0);
    const int np0 = clampi(p0 + delta, 0, 255), nq0 = clampi(q0 - delta, 0, 255);
    const int np1 = mask_sel(luma_m & ap, clampi(p1 + clamp_sym((p2 + avg - (p1 << 1)) >> 1, tc0), 0, 255), p1);
    const int nq1 = mask_sel(luma_m & aq, clampi(q1 + clamp_sym((q2 + avg - (q1 << 1)) >> 1, tc0), 0, 255), q1);
    // bS == 4: |p0 - q0| < (alpha >> 2) + 2
    const int near = sad_bias(p0, q0, e.nnear) >> 31;
    const int sp = luma_m & ap & near, sq = luma_m & aq & near;
    const int sp0 = mask_sel(sp, (p2 + 2 * p1 + 2 * p0 + 2 * q0 + q1 + 4) >> 3, (2 * p1 + p0 + q1 + 2) >> 2);
    const int sp1 = mask_sel(sp, (p2 + p1 + p0 + q0 + 2) >> 2, p1);
    const int sp2 = mask_sel(sp, (2 * p3 + 3 * p2 + p1 + p0 + q0 + 4) >> 3, p2);
    const int sq0 = mask_sel(sq, (p1 + 2 * p0 + 2 * q0 + 2 * q1 + q2 + 4) >> 3, (2 * q1 + q0 + p1 + 2) >> 2);
    const int sq1 = mask_sel(sq, (p0 + q0 + q1 + q2 + 2) >> 2, q1);
    const int sq2 = mask_sel(sq, (2 * q3 + 3 * q2 + q1 + q0 + p0 + 4) >> 3, q2);
    const int strong = (3 - bs) >> 31;   // bS == 4 (bS is 0..4)
    const int on_strong = on & strong;
    s[k - 3] = mask_sel(on_strong, sp2, p2);
    s[k - 2] = mask_sel(on, mask_sel(strong, sp1, np1), p1);
    s[k - 1] = mask_sel(on, mask_sel(strong, sp0, np0), p0);
    s[k] = mask_sel(on, mask_sel(strong, sq0, nq0), q0);
    s[k + 1] = mask_sel(on, mask_sel(strong, sq1, nq1), q1);
    s[k + 2] = mask_sel(on_strong, sq2, q2);
}

// filter_edge for an internal edge of a macroblock (luma edges 1-3, chroma edge 2), whose strength
// is at most 3 (boundary_strength gives 4 only on macroblock edges): the normal filter alone.
template <int k>
MMR_DEV void filter_edge_inner(int (&s)[20], const EdgeParams& e, int bs, int luma_m) {
    const int p0 = s[k - 1], p1 = s[k - 2], p2 = s[k - 3];
    const int q0 = s[k], q1 = s[k + 1], q2 = s[k + 2];
    const int on = edge_on(p1, p0, q0, q1, e, bs);
    const int ap = sad_bias(p2, p0, e.nb_ap) >> 31, aq = sad_bias(q2, q0, e.nb_aq) >> 31;
    const int tc0 = (int)((e.a >> (16 + 5 * (bs - 1))) & 0x1fu);
    const int tc = tc0 - ap - aq;
    const int delta = clamp_sym((((q0 - p0) << 2) + (p1 - q1) + 4) >> 3, tc);
    const int avg = avg_round(p0, q0);
    const int np1 = clampi(p1 + clamp_sym((p2 + avg - (p1 << 1)) >> 1, tc0), 0, 255);
    const int nq1 = clampi(q1 + clamp_sym((q2 + avg - (q1 << 1)) >> 1, tc0), 0, 255);
    s[k - 2] = mask_sel(on & luma_m & ap, np1, p1);
    s[k - 1] = mask_sel(on, clampi(p0 + delta, 0, 255), p0);
    s[k] = mask_sel(on, clampi(q0 - delta, 0, 255), q0);
    s[k + 1] = mask_sel(on & luma_m & aq, nq1, q1);
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

// The samples one step filters: the macroblock with 4 rows of top and 4 columns of left context.
struct DbWindow {
    uint8_t ly[20 * 20];               // rows y*16-4 .., cols x*16-4 ..
    uint8_t lc[2][20 * kLc];           // rows y*8-4 .., cols x*8-4 .. (12 x 12 used)
};
struct DbWave {
    DbWindow win[2];                   // even and odd steps
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

// The picture buffers a row reads and writes, taken from the job once into scalar registers. As
// far as the compiler knows a store to the picture may change the job record, so a field read
// through the record after a store is loaded again -- and that load's wait covers every store
// issued before it, which would put a memory round trip between consecutive stores.
struct DbPic {
    const MMR_GLOBAL mmr_recon_h264_macroblock* mbs;
    const MMR_GLOBAL uint32_t* bs;
    MMR_GLOBAL uint8_t* py;
    MMR_GLOBAL uint8_t* pcb;
    MMR_GLOBAL uint8_t* pcr;
    MMR_DEV MMR_GLOBAL uint8_t* chroma(int comp) const { return comp ? pcr : pcb; }
};
MMR_DEV DbPic db_pic(const H264Job& job) {
    DbPic p;
    p.mbs = uniform_ptr(as_global(job.mbs));
    p.bs = uniform_ptr(as_global(static_cast<const uint32_t*>(job.bs)));
    p.py = uniform_ptr(as_global(job.planes[0]));
    p.pcb = uniform_ptr(as_global(job.planes[1]));
    p.pcr = uniform_ptr(as_global(job.planes[2]));
    return p;
}

MMR_DEV void db_prefetch(const DbPic& P, int MW, int x, int y, int l, DbPrefetch& f) {
    const int W = MW * 16, CW = MW * 8;
    const uint32_t a = (uint32_t)(y * MW + x);
    {
        const MMR_GLOBAL uint32_t* q = reinterpret_cast<const MMR_GLOBAL uint32_t*>(P.mbs + a);
        const uint32_t w[3] = {q[0], q[1], q[2]};
        __builtin_memcpy(&f.m, w, sizeof f.m);
    }
    f.qp_top = y > 0 ? P.mbs[a - MW].qp : 0;
    {
        const mmr_v4i32 v = *reinterpret_cast<const MMR_GLOBAL mmr_v4i32*>(P.bs + (size_t)a * 4);
        f.bs = make_int4(v.x, v.y, v.z, v.w);
    }
    f.own_y = *reinterpret_cast<const MMR_GLOBAL uint32_t*>(P.py + (size_t)(y * 16 + (l >> 2)) * W + x * 16 + 4 * (l & 3));
    if (l < 32) {
        const int comp = l >> 4, row = (l >> 1) & 7;
        f.own_c = *reinterpret_cast<const MMR_GLOBAL uint32_t*>(P.chroma(comp) + (size_t)(y * 8 + row) * CW + x * 8 + 4 * (l & 1));
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
    while (poll_workgroup(counter) < need) {
        if (++spins > kDbSpinLimit) {
            atomicOr(status, kDevStatusTimeout);
            return false;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
    return true;
}

// Hand-over between bands: the row that hands samples to the next band stores all of them with
// agent_store and publishes with publish_agent; the next band's first row reads them with
// agent_load (coherent across the XCDs' L2 caches, no write-back or invalidate of the L2).
MMR_DEV void db_put(MMR_GLOBAL uint8_t* p, uint32_t v, bool agent) {
    MMR_GLOBAL uint32_t* q = reinterpret_cast<MMR_GLOBAL uint32_t*>(p);
    if (agent)
        agent_store(q, v);
    else
        *q = v;
}
MMR_DEV uint32_t db_get_coherent(const MMR_GLOBAL uint8_t* p) { return agent_load(reinterpret_cast<const MMR_GLOBAL uint32_t*>(p)); }

// Wait for the previous band's published progress (another workgroup) with relaxed agent-scope
// loads. The samples it covers are read afterwards with coherent loads (db_load_top), issued only
// once the progress load has returned (the compiler barrier keeps them after it). `seen` keeps the
// last value observed.
MMR_DEV bool db_wait_band(uint32_t* counter, uint32_t need, uint32_t& seen, uint32_t* status) {
    if (seen >= need) return true;
    uint32_t spins = 0;
    for (;;) {
        seen = poll_agent(counter);
        if (seen >= need) break;
        if (++spins > kDbSpinLimit) {
            atomicOr(status, kDevStatusTimeout);
            return false;
        }
        __builtin_amdgcn_s_sleep(2);
    }
    asm volatile("" ::: "memory");
    return true;
}

// The top context of macroblock x of row y (the four sample rows above it, luma and both chroma
// components), one dword per lane (lanes 0-15 luma, 16-31 chroma), read from the picture with
// coherent loads (the previous band's last row wrote them).
MMR_DEV uint32_t db_load_top(const DbPic& P, int MW, int x, int y, int l) {
    const int W = MW * 16, CW = MW * 8;
    if (l < 16) return db_get_coherent(P.py + (size_t)(y * 16 - 4 + (l >> 2)) * W + x * 16 + 4 * (l & 3));
    if (l < 32) {
        const int k = l - 16, comp = k >> 3, row = (k >> 1) & 3;
        return db_get_coherent(P.chroma(comp) + (size_t)(y * 8 - 4 + row) * CW + x * 8 + 4 * (k & 1));
    }
    return 0u;
}

// One macroblock row. band_first: the top context comes from the picture (gprev: the previous
// band's progress); band_last: no row below in this band (gself: progress for the next band, or
// null for the picture's last row) -- the row writes its bottom sample rows itself.
//
// A step is the wavefront's unit of latency, so it is kept short and free of memory round trips:
// the macroblock's own samples, record and strengths are loaded two steps ahead, and a band's
// first row loads its top context two steps ahead as well (it waits for progress >= x + 4 in the
// middle of step x, then reads the rows above macroblock x + 2). Even and odd steps filter in two
// windows; step x takes its left context from the other window and writes the other window's final
// samples (step x - 1's) back to the picture, so a step's stores have a whole step to complete
// before anything waits for memory again. The band hand-over (publish and wait) sits in the middle
// of a step, after the filtering and before the step's prefetch is issued. Published progress p means
// that steps 0 .. p-1 are complete and their samples written; the last row publishes p = x in the
// middle of step x (every `publish` macroblocks) and p = MW after its last store. The vertical
// edges do not read the rows above the macroblock, so the wait for the row above comes after them.
// A wait that gives up (it reports a time-out in the status word) marks the wave failed: later
// waits return at once and the row ends after the step.
template <int ROWS>
MMR_DEV void db_row(const H264Job& job, const DbTables& T, int y, int w, int l, DbWave* waves, uint32_t* prog, uint32_t* status,
                    bool band_first, bool band_last, uint32_t* gprev, uint32_t* gself, uint32_t publish) {
    DbWave& S = waves[w];
    DbWave& Above = waves[(w + ROWS - 1) % ROWS];
    const int MW = (int)job.width_mbs;
    const int W = MW * 16, CW = MW * 8;
    const DbPic P = db_pic(job);
    const int chroma_off0 = job.chroma_qp_offset[0], chroma_off1 = job.chroma_qp_offset[1];
    const bool last_row = band_last;   // writes its own bottom rows
    const bool global_top = y > 0 && band_first;
    const bool agent_out = gself != nullptr;   // the next band reads this row's samples
    uint32_t seen = 0;
    bool failed = false;
    // Two buffers, even and odd steps: step x reads its buffer, then refills it with step x + 2's
    // loads, so a loaded value is never copied (a copy would wait for the load) before its step.
    // (Named variables, not arrays: an array of them is placed in scratch memory.)
    DbPrefetch pf_even, pf_odd;
    uint32_t top_even = 0u, top_odd = 0u;   // top context of the buffer's macroblock (band_first rows)
    db_prefetch(P, MW, 0, y, l, pf_even);
    if (MW > 1) db_prefetch(P, MW, 1, y, l, pf_odd);
    if (global_top) {
        failed = !db_wait_band(gprev, (uint32_t)min(3, MW), seen, status);
        top_even = db_load_top(P, MW, 0, y, l);
        if (MW > 1) top_odd = db_load_top(P, MW, 1, y, l);
    }
    // Lane roles. Lanes 0-15 filter luma lines, 16-23 Cb and 24-31 Cr lines (rows in the vertical
    // pass, columns in the horizontal one); their QPs go through the chroma QP table. The same 32
    // lanes move one dword each of a macroblock's four-row top context (window rows 0-3, columns
    // 4-19; ring slot rows 0-3) and of its bottom rows (window rows 16-19 luma / 8-11 chroma).
    // Offsets are bytes from the start of a window (or of the wave's record for ring slots).
    const bool chroma_lane = l >= 16;
    const int luma_m = chroma_lane ? 0 : -1;
    const int chroma_off = (l >> 3) & 1 ? chroma_off1 : chroma_off0;
    const int li = chroma_lane ? ((l - 16) & 7) : l;   // line index within its plane
    constexpr int kLcOff = (int)offsetof(DbWindow, lc);
    const int line_off = chroma_lane ? kLcOff + ((l - 16) >> 3) * 20 * kLc : 0;   // the lane's plane in a window
    const int ck = l - 16, ccomp = (ck >> 3) & 1, crow = (ck >> 1) & 3;
    const int ctx_off = chroma_lane ? kLcOff + ccomp * 20 * kLc + crow * kLc + 4 + 4 * (ck & 1) : (l >> 2) * 20 + 4 + 4 * (l & 3);
    const int bot_off = ctx_off + (chroma_lane ? 8 * kLc : 16 * 20);
    const int ring_off = chroma_lane ? (int)offsetof(DbWave, ring_c) + ccomp * 32 + crow * 8 + 4 * (ck & 1)
                                     : (int)offsetof(DbWave, ring_y) + (l >> 2) * 16 + 4 * (l & 3);
    static_assert(sizeof(DbWave::ring_y[0]) == 64 && sizeof(DbWave::ring_c[0]) == 64, "ring slots are 64 bytes");
    // The right four columns of a macroblock's bottom rows (luma dword 3 of a row, chroma dword 1):
    // once the next macroblock's left edge has filtered them they are that step's window's left
    // context (columns 0-3), `right_shift` bytes before bot_off.
    const bool right_lane = chroma_lane ? (ck & 1) != 0 : (l & 3) == 3;
    const int right_shift = chroma_lane ? 8 : 16;
    // Left context: lanes 0-19 the luma rows, 20-43 the chroma rows (the other window's right 4
    // columns become this window's left 4).
    const int sk = l - 20, scomp = sk >= 12 ? 1 : 0, srow = sk - 12 * scomp;
    const int left_off = l < 20 ? l * 20 : kLcOff + scomp * 20 * kLc + srow * kLc;
    const int left_from = l < 20 ? 16 : 8;
    uint8_t* const wave_base = reinterpret_cast<uint8_t*>(&S);
    const uint8_t* const above_base = reinterpret_cast<const uint8_t*>(&Above);
    int qp_left = 0;

    // Writes the samples step xs made final (in window wb) to the picture: luma rows y*16-3 ..
    // y*16+12 of columns xs*16-4 .. xs*16+11 (chroma rows y*8-3 .. y*8+4 of columns xs*8-4 ..
    // xs*8+3); the rows below are handed to and written by the next row, which modifies them, except
    // on the last row, and the last column writes the right 4 columns as well.
    auto write_back = [&](const int xs, const int wb) {
        const DbWindow& Wn = S.win[wb];
        const int r = 1 + (l >> 2), c4 = 4 * (l & 3);
        const int gy = y * 16 - 4 + r, gx = xs * 16 - 4 + c4;
        if (gy >= 0 && gx >= 0)
            db_put(P.py + (size_t)gy * W + gx, *reinterpret_cast<const uint32_t*>(&Wn.ly[r * 20 + c4]), agent_out);
        if (l < 32) {   // chroma: window rows 1..8, cols 0..7
            const int comp = l >> 4, rr = 1 + ((l >> 1) & 7), cc = 4 * (l & 1);
            const int cy = y * 8 - 4 + rr, cx = xs * 8 - 4 + cc;
            if (cy >= 0 && cx >= 0)
                db_put(P.chroma(comp) + (size_t)cy * CW + cx, *reinterpret_cast<const uint32_t*>(&Wn.lc[comp][rr * kLc + cc]), agent_out);
        }
        if (last_row || xs == MW - 1) {
            if (last_row && l < 12) {   // rows y*16+13 .. +15 (window rows 17..19)
                const int r2 = 17 + (l >> 2), gx2 = xs * 16 - 4 + c4;
                if (gx2 >= 0)
                    db_put(P.py + (size_t)(y * 16 - 4 + r2) * W + gx2, *reinterpret_cast<const uint32_t*>(&Wn.ly[r2 * 20 + c4]), agent_out);
            }
            if (xs == MW - 1 && l < 19) {   // last column: columns xs*16+12 .. +15 (window col 16)
                const int r3 = 1 + l;
                const int gy3 = y * 16 - 4 + r3;
                if (gy3 >= 0 && (r3 <= 16 || last_row))
                    db_put(P.py + (size_t)gy3 * W + xs * 16 + 12, *reinterpret_cast<const uint32_t*>(&Wn.ly[r3 * 20 + 16]), agent_out);
            }
            // chroma extras: last row, rows y*8+5 .. +7 (window rows 9..11, 12 dwords); last column,
            // columns xs*8+4 .. +7 (window col 8, 2 x 11 rows): at most 34 dwords, one per lane
            const int extra_rows = last_row ? 12 : 0, extra_cols = xs == MW - 1 ? 22 : 0;
            if (l < extra_rows + extra_cols) {
                const int i = l;
                if (i < extra_rows) {
                    const int comp = i / 6, rr = 9 + (i % 6) / 2, cc = 4 * (i & 1);
                    const int cx = xs * 8 - 4 + cc;
                    if (cx >= 0)
                        db_put(P.chroma(comp) + (size_t)(y * 8 - 4 + rr) * CW + cx, *reinterpret_cast<const uint32_t*>(&Wn.lc[comp][rr * kLc + cc]), agent_out);
                } else {
                    const int k = i - extra_rows, comp = k / 11, rr = 1 + k % 11;
                    const int cy = y * 8 - 4 + rr;
                    if (cy >= 0 && (rr <= 8 || last_row))
                        db_put(P.chroma(comp) + (size_t)cy * CW + xs * 8 + 4, *reinterpret_cast<const uint32_t*>(&Wn.lc[comp][rr * kLc + 8]), agent_out);
                }
            }
        }
    };

    // Hands the bottom four rows of macroblock xs (window wmb; its right columns in window wright
    // at columns 0-3 unless `last`, when no later left edge changes them) to the row below as ring
    // slot xs, then publishes progress xs + 1: slots 0 .. xs are handed over. Every row publishes,
    // the band's last one too: a row's progress p also tells the row above that slots 0 .. p - 1
    // were read (slot j is read in step j, before progress j + 1 is published in step j + 1), so a
    // row writes slot xs over slot xs - kDbRing once the row below has progress >= xs - kDbRing + 1.
    auto hand_over = [&](const int xs, const uint8_t* wmb, const uint8_t* wright, const bool last) {
        if (!band_last) {
            const int free_at = xs + 1 - kDbRing;
            if (free_at > 0 && !failed && !db_wait(&prog[w + 1], (uint32_t)free_at, status)) failed = true;
            if (l < 32) {
                const uint8_t* src = !last && right_lane ? wright + bot_off - right_shift : wmb + bot_off;
                *reinterpret_cast<uint32_t*>(wave_base + ring_off + 64 * (xs % kDbRing)) = *reinterpret_cast<const uint32_t*>(src);
            }
        }
        wave_sync();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
        if (l == 0) __hip_atomic_store(&prog[w], (uint32_t)(xs + 1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    };

    // One step in window b = x & 1 (a constant after inlining: the loop below is unrolled by two).
    // The top context travels by value (returned: the rows above macroblock x + 2, or the input):
    // through a reference the compiler may load it through a pointer select with the ring slot.
    auto step = [&](const int x, const int b, DbPrefetch& buf, const uint32_t top_in) -> uint32_t {
        uint32_t top_next = top_in;
        uint8_t* const wbase = reinterpret_cast<uint8_t*>(&S.win[b]);
        const uint8_t* const wprev = reinterpret_cast<const uint8_t*>(&S.win[b ^ 1]);
        const DbPrefetch cur = buf;
        const mmr_recon_h264_macroblock m = cur.m;
        const bool filter = !(m.flags & MMR_H264_MB_NO_DEBLOCK);
        // filter parameters of the left edge, the internal edges and the top edge
        auto lane_qp = [&](int qp) { return chroma_lane ? db_chroma_qp(T, qp, chroma_off) : qp; };
        const int qpc = lane_qp(m.qp);
        const EdgeParams ev = edge_params(T, (lane_qp(qp_left) + qpc + 1) >> 1, m.offset_a, m.offset_b, luma_m);
        const EdgeParams ei = edge_params(T, qpc, m.offset_a, m.offset_b, luma_m);
        const EdgeParams eh = edge_params(T, (lane_qp(cur.qp_top) + qpc + 1) >> 1, m.offset_a, m.offset_b, luma_m);
        // Strengths from the macroblock's record: vertical edge e, segment g at nibble e*4+g of
        // dwords x / y, horizontal at nibble e*4+g of dwords z / w. Edge slots k = 4, 8, 12, 16 are
        // the luma edges 0..3 and the chroma edges 0 and 2 (slots 12, 16 have strength 0 on chroma
        // lines).
        const uint32_t sh = 4u * (uint32_t)(chroma_lane ? (li >> 1) : (li >> 2));
        const int4 r = cur.bs;
        auto nib = [&](int v, uint32_t shift) { return (int)(((uint32_t)v >> shift) & 0xfu); };
        const int bv0 = nib(r.x, sh), bv1 = chroma_lane ? nib(r.y, sh) : nib(r.x, sh + 16);
        const int bv2 = chroma_lane ? 0 : nib(r.y, sh), bv3 = chroma_lane ? 0 : nib(r.y, sh + 16);
        const int bh0 = nib(r.z, sh), bh1 = chroma_lane ? nib(r.w, sh) : nib(r.z, sh + 16);
        const int bh2 = chroma_lane ? 0 : nib(r.w, sh), bh3 = chroma_lane ? 0 : nib(r.w, sh + 16);
        // Edge slots whose strengths are all 0 (luma and chroma lines alike) change no sample; the
        // record is the same in every lane, so the test is a scalar branch. Slot 4: dword bits 0-15
        // of x (z); slot 8: bits 16-31 of x (z) and, for chroma, bits 0-15 of y (w); slot 12: bits
        // 0-15 of y (w); slot 16: bits 16-31 of y (w).
        const uint32_t vx = (uint32_t)__builtin_amdgcn_readfirstlane(r.x), vy = (uint32_t)__builtin_amdgcn_readfirstlane(r.y);
        const uint32_t hz = (uint32_t)__builtin_amdgcn_readfirstlane(r.z), hw = (uint32_t)__builtin_amdgcn_readfirstlane(r.w);
        const bool v4 = (vx & 0xffffu) != 0, v8 = (vx >> 16) != 0 || (vy & 0xffffu) != 0, v12 = (vy & 0xffffu) != 0, v16 = (vy >> 16) != 0;
        const bool h4 = (hz & 0xffffu) != 0, h8 = (hz >> 16) != 0 || (hw & 0xffffu) != 0, h12 = (hw & 0xffffu) != 0, h16 = (hw >> 16) != 0;
        // own samples (rows 4.., cols 4..) and left context (cols 0..3) into the window
        *reinterpret_cast<uint32_t*>(wbase + (4 + (l >> 2)) * 20 + 4 + 4 * (l & 3)) = cur.own_y;
        if (l < 32) *reinterpret_cast<uint32_t*>(wbase + kLcOff + (l >> 4) * 20 * kLc + (4 + ((l >> 1) & 7)) * kLc + 4 + 4 * (l & 1)) = cur.own_c;
        if (l < 44) *reinterpret_cast<uint32_t*>(wbase + left_off) = *reinterpret_cast<const uint32_t*>(wprev + left_off + left_from);
        // the previous step's final samples: its window is not written again before step x + 1
        if (x > 0) write_back(x - 1, b ^ 1);
        wave_sync();
        if (filter && l < 32) {   // vertical edges: the lane's row, 5 dwords of the window
            uint32_t* row = reinterpret_cast<uint32_t*>(wbase + line_off + (4 + li) * 20);
            int sm[20];
#pragma unroll
            for (int i = 0; i < 5; ++i) {
                const uint32_t v = row[i];
#pragma unroll
                for (int j = 0; j < 4; ++j) sm[4 * i + j] = (int)((v >> (8 * j)) & 0xffu);
            }
            if (v4) filter_edge<4>(sm, ev, bv0, luma_m);
            if (v8) filter_edge_inner<8>(sm, ei, bv1, luma_m);
            if (v12) filter_edge_inner<12>(sm, ei, bv2, luma_m);
            if (v16) filter_edge_inner<16>(sm, ei, bv3, luma_m);
#pragma unroll
            for (int i = 0; i < 5; ++i)
                row[i] = (uint32_t)sm[4 * i] | ((uint32_t)sm[4 * i + 1] << 8) | ((uint32_t)sm[4 * i + 2] << 16) | ((uint32_t)sm[4 * i + 3] << 24);
        }
        // macroblock x - 1's bottom rows are final now (this step's left edge filtered their right
        // columns): hand them to the row below
        wave_sync();
        if (x > 0) hand_over(x - 1, wprev, wbase, false);
        // top context: from the picture (loaded two steps ago) or from the row above's ring slot x,
        // once that row has handed it over (after the left edge of its macroblock x + 1)
        if (y > 0) {
            uint32_t top = top_in;
            if (!band_first) {
                if (!failed && !db_wait(&prog[w - 1], (uint32_t)(x + 1), status)) failed = true;
                top = l < 32 ? *reinterpret_cast<const uint32_t*>(above_base + ring_off + 64 * (x % kDbRing)) : 0u;
            }
            if (l < 32) *reinterpret_cast<uint32_t*>(wbase + ctx_off) = top;
        }
        wave_sync();
        if (filter && l < 32) {   // horizontal edges: the lane's column, 20 bytes of the window
            uint8_t* col = wbase + line_off + 4 + li;
            int sm[20];
#pragma unroll
            for (int i = 0; i < 20; ++i) sm[i] = col[i * 20];
            if (h4) filter_edge<4>(sm, eh, bh0, luma_m);
            if (h8) filter_edge_inner<8>(sm, ei, bh1, luma_m);
            if (h12) filter_edge_inner<12>(sm, ei, bh2, luma_m);
            if (h16) filter_edge_inner<16>(sm, ei, bh3, luma_m);
#pragma unroll
            for (int i = 1; i < 19; ++i) col[i * 20] = (uint8_t)sm[i];
        }
        wave_sync();
        // Mid-step: the agent-scope hand-over, then this step's loads for step x + 2.
        if (gself && x > 0 && x % publish == 0)   // steps 0 .. x-1 are written (step x-1's stores were issued above)
            publish_agent(gself, (uint32_t)x);
        if (x + 2 < MW) {
            if (global_top) {
                if (!failed && !db_wait_band(gprev, (uint32_t)min(x + 4, MW), seen, status)) failed = true;
                top_next = db_load_top(P, MW, x + 2, y, l);
            }
            db_prefetch(P, MW, x + 2, y, l, buf);
        }
        qp_left = m.qp;
        return top_next;
    };
    for (int x = 0; x < MW && !failed; x += 2) {
        top_even = step(x, 0, pf_even, top_even);
        if (x + 1 < MW && !failed) top_odd = step(x + 1, 1, pf_odd, top_odd);
    }
    if (!failed) {
        wave_sync();
        const uint8_t* const wlast = reinterpret_cast<const uint8_t*>(&S.win[(MW - 1) & 1]);
        hand_over(MW - 1, wlast, wlast, true);
        write_back(MW - 1, (MW - 1) & 1);
    }
    if (gself) publish_agent(gself, (uint32_t)MW);
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
        const uint32_t t = (uint32_t)__builtin_amdgcn_readfirstlane((int)s_ticket);
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
