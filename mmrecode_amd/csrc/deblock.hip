// H.264 loop filter over frame pictures as an edge-order wavefront.
//
// Reference: filter_picture / derive_boundary_strengths / boundary_strength / motion_pairs_match /
// filter_luma_macroblock / filter_chroma_macroblock / filter_vertical/horizontal_edge /
// filter_samples / strong_filter / normal_filter / ALPHA / BETA / TC0
// (crates/codecs/h264/src/deblock.rs:3-587).
//
// The reference filters macroblocks in raster order; macroblock (x, y) reads and writes samples
// of (x-1, y) and (x, y-1), and (x+1, y-1) writes samples that (x, y) reads, so (x, y) may run once
// (x-1, y) and (x+1, y-1) are done: a wavefront over t = x + 2y. Each wave owns one macroblock row
// and walks it left to right; before macroblock x it waits until the row above has published
// progress >= min(x + 2, width). Rows are handed out by a device ticket counter in order, so the
// row a wave waits on was taken earlier by a running wave (no deadlock for any residency).
//
// Per macroblock the wave stages a 20x20 luma and two 12x12 chroma windows (4 samples of left /
// top context) in LDS, computes the 32 boundary strengths from metadata, filters vertical edges
// (one lane per row: 16 luma + 8 Cb + 8 Cr), then horizontal edges (one lane per column), and
// writes the window back. Sample order inside a lane follows the reference exactly.
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

MMR_DEV int db_chroma_qp(int luma_qp, int offset) {
    const int index = clampi(luma_qp + offset, 0, 51);
    return index < 30 ? index : kDbChromaQp[index - 30];
}
MMR_DEV int iabs(int v) { return v < 0 ? -v : v; }

// filter_samples + strong_filter + normal_filter (deblock.rs:469-569) on LDS samples:
// p_i at s[-(i+1)*step], q_i at s[i*step].
MMR_DEV void filter_samples(uint8_t* s, int step, int qp, int bs, bool chroma, int offset_a, int offset_b) {
    const int ia = clampi(qp + offset_a, 0, 51), ib = clampi(qp + offset_b, 0, 51);
    const int alpha = kAlpha[ia], beta = kBeta[ib];
    int p[4], q[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        p[i] = s[-(i + 1) * step];
        q[i] = s[i * step];
    }
    if (iabs(p[0] - q[0]) >= alpha || iabs(p[1] - p[0]) >= beta || iabs(q[1] - q[0]) >= beta) return;
    int fp[3], fq[3];
    if (bs == 4) {
        const bool sp = !chroma && iabs(p[2] - p[0]) < beta && iabs(p[0] - q[0]) < (alpha >> 2) + 2;
        const bool sq = !chroma && iabs(q[2] - q[0]) < beta && iabs(p[0] - q[0]) < (alpha >> 2) + 2;
        if (sp) {
            fp[0] = (p[2] + 2 * p[1] + 2 * p[0] + 2 * q[0] + q[1] + 4) >> 3;
            fp[1] = (p[2] + p[1] + p[0] + q[0] + 2) >> 2;
            fp[2] = (2 * p[3] + 3 * p[2] + p[1] + p[0] + q[0] + 4) >> 3;
        } else {
            fp[0] = (2 * p[1] + p[0] + q[1] + 2) >> 2;
            fp[1] = p[1];
            fp[2] = p[2];
        }
        if (sq) {
            fq[0] = (p[1] + 2 * p[0] + 2 * q[0] + 2 * q[1] + q[2] + 4) >> 3;
            fq[1] = (p[0] + q[0] + q[1] + q[2] + 2) >> 2;
            fq[2] = (2 * q[3] + 3 * q[2] + q[1] + q[0] + p[0] + 4) >> 3;
        } else {
            fq[0] = (2 * q[1] + q[0] + p[1] + 2) >> 2;
            fq[1] = q[1];
            fq[2] = q[2];
        }
    } else {
        const int tc0 = kTc0[bs - 1][ia];
        const int ap = iabs(p[2] - p[0]), aq = iabs(q[2] - q[0]);
        const int tc = tc0 + (chroma ? 1 : ((ap < beta) + (aq < beta)));
        const int delta = clampi((((q[0] - p[0]) << 2) + (p[1] - q[1]) + 4) >> 3, -tc, tc);
        fp[0] = p[0] + delta;
        fp[1] = p[1];
        fp[2] = p[2];
        fq[0] = q[0] - delta;
        fq[1] = q[1];
        fq[2] = q[2];
        if (!chroma && ap < beta) fp[1] = p[1] + clampi((p[2] + ((p[0] + q[0] + 1) >> 1) - (p[1] << 1)) >> 1, -tc0, tc0);
        if (!chroma && aq < beta) fq[1] = q[1] + clampi((q[2] + ((p[0] + q[0] + 1) >> 1) - (q[1] << 1)) >> 1, -tc0, tc0);
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        s[-(i + 1) * step] = (uint8_t)clampi(fp[i], 0, 255);
        s[i * step] = (uint8_t)clampi(fq[i], 0, 255);
    }
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

// boundary_strength + motion_pairs_match (deblock.rs:391-439).
MMR_DEV int boundary_strength(const H264Job& job, uint32_t pm, int pb, uint32_t qm, int qb, bool mb_edge) {
    if ((job.mbs[pm].flags | job.mbs[qm].flags) & MMR_H264_MB_INTRA) return mb_edge ? 4 : 3;
    const mmr_recon_h264_block P = job.blocks[(size_t)pm * 16 + pb], Q = job.blocks[(size_t)qm * 16 + qb];
    if (P.total_coeff != 0 || Q.total_coeff != 0) return 2;
    const RefMotion l0 = motion_of(job, P, 0), l1 = motion_of(job, P, 1), r0 = motion_of(job, Q, 0), r1 = motion_of(job, Q, 1);
    bool match;
    if (!l1.has_ref && !r1.has_ref) match = indexed_match(l0, r0);
    else if (!l0.has_ref && !r0.has_ref) match = indexed_match(l1, r1);
    else match = (motion_match(l0, r0) && motion_match(l1, r1)) || (motion_match(l0, r1) && motion_match(l1, r0));
    return match ? 0 : 1;
}

MMR_DEV int luma_block_index(int x, int y) { return (y / 2) * 8 + (x / 2) * 4 + (y % 2) * 2 + x % 2; }

MMR_DEV bool edge_enabled(const H264Job& job, uint32_t pm, uint32_t qm) {
    const mmr_recon_h264_macroblock q = job.mbs[qm];
    return !(q.flags & MMR_H264_MB_SLICE_EDGE_OFF) || job.mbs[pm].slice_id == q.slice_id;
}

__global__ __launch_bounds__(64) void h264_deblock_kernel(const H264Job* __restrict__ jobs, uint32_t job_count,
                                                          uint32_t total_rows, uint32_t* __restrict__ counters) {
    __shared__ __align__(16) uint8_t ly[20 * 20];
    __shared__ __align__(16) uint8_t lc[2][12 * 12];
    __shared__ int bs_v[4][4], bs_h[4][4];
    __shared__ uint32_t s_ticket;
    const int l = threadIdx.x;
    uint32_t* ticket_counter = counters;
    uint32_t* progress = counters + 1;
    for (;;) {
        if (l == 0) s_ticket = fetch_add_relaxed(ticket_counter, 1);
        __syncthreads();
        const uint32_t ticket = s_ticket;
        __syncthreads();
        if (ticket >= total_rows) break;
        const uint32_t ji = find_job(jobs, job_count, ticket, [](const H264Job& j) { return j.row_begin; });
        const H264Job& job = jobs[ji];
        const int y = (int)(ticket - job.row_begin);
        const int MW = (int)job.width_mbs;
        const int W = MW * 16, CW = MW * 8;
        for (int x = 0; x < MW; ++x) {
            if (y > 0) {
                const uint32_t need = (uint32_t)min(x + 2, MW);
                while (load_acquire(&progress[ticket - 1]) < need) __builtin_amdgcn_s_sleep(1);
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            }
            const uint32_t a = (uint32_t)(y * MW + x);
            const mmr_recon_h264_macroblock m = job.mbs[a];
            if (!(m.flags & MMR_H264_MB_NO_DEBLOCK)) {
                const bool left_on = x > 0 && edge_enabled(job, a - 1, a);
                const bool top_on = y > 0 && edge_enabled(job, a - MW, a);
                // derive_boundary_strengths (deblock.rs:144-200)
                if (l < 32) {
                    const int e = (l & 15) >> 2, sgm = l & 3;
                    int bs = 0;
                    if (l < 16) {
                        if (e != 0 || left_on)
                            bs = boundary_strength(job, e == 0 ? a - 1 : a, luma_block_index(e == 0 ? 3 : e - 1, sgm), a,
                                                   luma_block_index(e == 0 ? 0 : e, sgm), e == 0);
                        bs_v[e][sgm] = bs;
                    } else {
                        if (e != 0 || top_on)
                            bs = boundary_strength(job, e == 0 ? a - MW : a, luma_block_index(sgm, e == 0 ? 3 : e - 1), a,
                                                   luma_block_index(sgm, e == 0 ? 0 : e), e == 0);
                        bs_h[e][sgm] = bs;
                    }
                }
                // Stage windows: luma rows y*16-4.., cols x*16-4..; chroma rows y*8-4.., cols x*8-4..
                const int lx0 = x * 16 - 4, ly0 = y * 16 - 4, cx0 = x * 8 - 4, cy0 = y * 8 - 4;
                for (int i = l; i < 20 * 5; i += 64) {
                    const int r = i / 5, c4 = (i % 5) * 4;
                    const int gy = ly0 + r, gx = lx0 + c4;
                    if (gy >= 0 && gx >= 0)
                        *reinterpret_cast<uint32_t*>(&ly[r * 20 + c4]) = *reinterpret_cast<const uint32_t*>(job.planes[0] + (size_t)gy * W + gx);
                }
                for (int i = l; i < 2 * 12 * 3; i += 64) {
                    const int comp = i / 36, r = (i % 36) / 3, c4 = (i % 3) * 4;
                    const int gy = cy0 + r, gx = cx0 + c4;
                    if (gy >= 0 && gx >= 0)
                        *reinterpret_cast<uint32_t*>(&lc[comp][r * 12 + c4]) =
                            *reinterpret_cast<const uint32_t*>(job.planes[1 + comp] + (size_t)gy * CW + gx);
                }
                __syncthreads();
                const int oa = m.offset_a, ob = m.offset_b;
                const bool t8 = (m.flags & MMR_H264_MB_TRANSFORM_8X8) != 0;
                const int qp_left = x > 0 ? job.mbs[a - 1].qp : 0, qp_top = y > 0 ? job.mbs[a - MW].qp : 0;
                // Vertical edges: lanes 0-15 luma rows, 16-23 Cb rows, 24-31 Cr rows.
                if (l < 16) {
                    for (int e = 0; e < 4; ++e) {
                        if (e == 0 && !left_on) continue;
                        if (t8 && (e == 1 || e == 3)) continue;
                        const int bs = bs_v[e][l >> 2];
                        if (bs == 0) continue;
                        const int qp = ((e == 0 ? qp_left : m.qp) + m.qp + 1) >> 1;
                        filter_samples(&ly[(4 + l) * 20 + 4 + e * 4], 1, qp, bs, false, oa, ob);
                    }
                } else if (l < 32) {
                    const int comp = (l - 16) >> 3, row = (l - 16) & 7;
                    const int off = job.chroma_qp_offset[comp];
                    for (int e = 0; e < 2; ++e) {
                        if (e == 0 && !left_on) continue;
                        const int bs = bs_v[e * 2][row >> 1];
                        if (bs == 0) continue;
                        const int qp = (db_chroma_qp(e == 0 ? qp_left : m.qp, off) + db_chroma_qp(m.qp, off) + 1) >> 1;
                        filter_samples(&lc[comp][(4 + row) * 12 + 4 + e * 4], 1, qp, bs, true, oa, ob);
                    }
                }
                __syncthreads();
                // Horizontal edges: lanes 0-15 luma columns, 16-23 Cb columns, 24-31 Cr columns.
                if (l < 16) {
                    for (int e = 0; e < 4; ++e) {
                        if (e == 0 && !top_on) continue;
                        if (t8 && (e == 1 || e == 3)) continue;
                        const int bs = bs_h[e][l >> 2];
                        if (bs == 0) continue;
                        const int qp = ((e == 0 ? qp_top : m.qp) + m.qp + 1) >> 1;
                        filter_samples(&ly[(4 + e * 4) * 20 + 4 + l], 20, qp, bs, false, oa, ob);
                    }
                } else if (l < 32) {
                    const int comp = (l - 16) >> 3, col = (l - 16) & 7;
                    const int off = job.chroma_qp_offset[comp];
                    for (int e = 0; e < 2; ++e) {
                        if (e == 0 && !top_on) continue;
                        const int bs = bs_h[e * 2][col >> 1];
                        if (bs == 0) continue;
                        const int qp = (db_chroma_qp(e == 0 ? qp_top : m.qp, off) + db_chroma_qp(m.qp, off) + 1) >> 1;
                        filter_samples(&lc[comp][(4 + e * 4) * 12 + 4 + col], 12, qp, bs, true, oa, ob);
                    }
                }
                __syncthreads();
                // Write back rows y*16-3.. (dword-aligned from column x*16-4; no other wave writes
                // those samples while this macroblock is in flight).
                for (int i = l; i < 20 * 5; i += 64) {
                    const int r = i / 5, c4 = (i % 5) * 4;
                    const int gy = ly0 + r, gx = lx0 + c4;
                    if (r >= 1 && gy >= 0 && gx >= 0)
                        *reinterpret_cast<uint32_t*>(job.planes[0] + (size_t)gy * W + gx) = *reinterpret_cast<const uint32_t*>(&ly[r * 20 + c4]);
                }
                for (int i = l; i < 2 * 12 * 3; i += 64) {
                    const int comp = i / 36, r = (i % 36) / 3, c4 = (i % 3) * 4;
                    const int gy = cy0 + r, gx = cx0 + c4;
                    if (r >= 1 && gy >= 0 && gx >= 0)
                        *reinterpret_cast<uint32_t*>(job.planes[1 + comp] + (size_t)gy * CW + gx) =
                            *reinterpret_cast<const uint32_t*>(&lc[comp][r * 12 + c4]);
                }
            }
            // Publish macroblock x of this row.
            __threadfence();
            __syncthreads();
            if (l == 0) add_release(&progress[ticket], 1);
        }
    }
}

hipError_t launch_h264_deblock(const H264Job* d_jobs, uint32_t job_count, uint32_t total_rows, uint32_t* d_counters,
                               uint32_t counter_words, uint32_t groups, hipStream_t s) {
    (void)counter_words;
    if (total_rows == 0) return hipSuccess;
    if (groups > total_rows) groups = total_rows;
    hipLaunchKernelGGL(h264_deblock_kernel, dim3(groups), dim3(64), 0, s, d_jobs, job_count, total_rows, d_counters);
    return hipGetLastError();
}

}  // namespace mmr
