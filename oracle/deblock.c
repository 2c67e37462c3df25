/* Oracle: H.264 loop filter restated from crates/codecs/h264/src/deblock.rs. Test infrastructure only. */
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* ALPHA, BETA, TC0: deblock.rs:3-27 (ITU-T H.264 tables 8-16, 8-17). */
static const int ALPHA[52] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 4, 4, 5, 6, 7, 8, 9, 10, 12, 13,
                              15, 17, 20, 22, 25, 28, 32, 36, 40, 45, 50, 56, 63, 71, 80, 90, 101, 113, 127, 144, 162, 182, 203, 226, 255, 255};
static const int BETA[52] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4,
                             6, 6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12, 13, 13, 14, 14, 15, 15, 16, 16, 17, 17, 18, 18};
static const int TC0[3][52] = {
    {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 4, 4, 4, 5, 6, 6, 7, 8, 9, 10, 11, 13},
    {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 4, 4, 5, 5, 6, 7, 8, 8, 10, 11, 12, 13, 15, 17},
    {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 4, 4, 4, 5, 6, 6, 7, 8, 9, 10, 11, 13, 14, 16, 18, 20, 23, 25}};

static int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
static uint8_t clip_sample(int v) { return (uint8_t)clampi(v, 0, 255); }

/* filter_samples + strong_filter + normal_filter, deblock.rs:469-569. All six p/q samples are
 * written back, unchanged ones included, as in the reference. */
void oracle_h264_filter_samples(uint8_t *plane, const int pi[4], const int qi[4], int qp, int bs, int chroma,
                                int offset_a, int offset_b) {
    int index_a = clampi(qp + offset_a, 0, 51), index_b = clampi(qp + offset_b, 0, 51);
    int alpha = ALPHA[index_a], beta = BETA[index_b];
    int p[4], q[4], fp[3], fq[3];
    for (int i = 0; i < 4; ++i) {
        p[i] = plane[pi[i]];
        q[i] = plane[qi[i]];
    }
    if (abs(p[0] - q[0]) >= alpha || abs(p[1] - p[0]) >= beta || abs(q[1] - q[0]) >= beta) return;
    if (bs == 4) {
        int sp = !chroma && abs(p[2] - p[0]) < beta && abs(p[0] - q[0]) < (alpha >> 2) + 2;
        int sq = !chroma && abs(q[2] - q[0]) < beta && abs(p[0] - q[0]) < (alpha >> 2) + 2;
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
        int tc0 = TC0[bs - 1][index_a];
        int ap = abs(p[2] - p[0]), aq = abs(q[2] - q[0]);
        int tc = tc0 + (chroma ? 1 : ((ap < beta) + (aq < beta)));
        int delta = clampi((((q[0] - p[0]) << 2) + (p[1] - q[1]) + 4) >> 3, -tc, tc);
        fp[0] = p[0] + delta;
        fp[1] = p[1];
        fp[2] = p[2];
        fq[0] = q[0] - delta;
        fq[1] = q[1];
        fq[2] = q[2];
        if (!chroma && ap < beta) fp[1] = p[1] + clampi((p[2] + ((p[0] + q[0] + 1) >> 1) - (p[1] << 1)) >> 1, -tc0, tc0);
        if (!chroma && aq < beta) fq[1] = q[1] + clampi((q[2] + ((p[0] + q[0] + 1) >> 1) - (q[1] << 1)) >> 1, -tc0, tc0);
    }
    for (int i = 0; i < 3; ++i) {
        plane[pi[i]] = clip_sample(fp[i]);
        plane[qi[i]] = clip_sample(fq[i]);
    }
}

typedef struct { int x, y, has_ref, index, poc; } ref_motion;

static ref_motion motion_of(const mmr_recon_h264_picture *p, int mb, int blk, int list) {
    const mmr_recon_h264_block *b = &p->blocks[(size_t)mb * 16 + blk];
    ref_motion m;
    memset(&m, 0, sizeof m);
    m.index = -1;
    if (b->ref_idx[list] >= 0) { /* MotionInfo::default() for an unused list (decoder.rs:11590) */
        m.x = b->mv[list][0];
        m.y = b->mv[list][1];
        m.has_ref = 1;
        m.index = b->ref_idx[list];
        m.poc = p->refs[list][b->ref_idx[list]].poc;
    }
    return m;
}

/* vector_matches / motion_matches / indexed_motion_matches / motion_pairs_match, deblock.rs:415-439. */
static int vector_matches(ref_motion a, ref_motion b) { return abs(a.x - b.x) < 4 && abs(a.y - b.y) < 4; }
static int same_ref(ref_motion a, ref_motion b) { return a.has_ref == b.has_ref && (!a.has_ref || a.poc == b.poc); }
static int motion_matches(ref_motion a, ref_motion b) { return same_ref(a, b) && vector_matches(a, b); }
static int indexed_matches(ref_motion a, ref_motion b) { return a.index == b.index && vector_matches(a, b); }
static int motion_pairs_match(const mmr_recon_h264_picture *p, int pm, int pb, int qm, int qb) {
    ref_motion l0 = motion_of(p, pm, pb, 0), l1 = motion_of(p, pm, pb, 1);
    ref_motion r0 = motion_of(p, qm, qb, 0), r1 = motion_of(p, qm, qb, 1);
    if (!l1.has_ref && !r1.has_ref) return indexed_matches(l0, r0);
    if (!l0.has_ref && !r0.has_ref) return indexed_matches(l1, r1);
    int same = motion_matches(l0, r0) && motion_matches(l1, r1);
    int swapped = motion_matches(l0, r1) && motion_matches(l1, r0);
    return same || swapped;
}

/* boundary_strength, deblock.rs:391-413. */
static int boundary_strength(const mmr_recon_h264_picture *p, int pm, int pb, int qm, int qb, int mb_edge) {
    if ((p->macroblocks[pm].flags | p->macroblocks[qm].flags) & MMR_H264_MB_INTRA) return mb_edge ? 4 : 3;
    if (p->blocks[(size_t)pm * 16 + pb].total_coeff != 0 || p->blocks[(size_t)qm * 16 + qb].total_coeff != 0) return 2;
    return !motion_pairs_match(p, pm, pb, qm, qb);
}

static int luma_block_index(int x, int y) { return (y / 2) * 8 + (x / 2) * 4 + (y % 2) * 2 + x % 2; }

/* external_edge_enabled, deblock.rs:380-389. */
static int external_edge_enabled(const mmr_recon_h264_picture *p, int pm, int qm) {
    const mmr_recon_h264_macroblock *q = &p->macroblocks[qm];
    return !(q->flags & MMR_H264_MB_SLICE_EDGE_OFF) || p->macroblocks[pm].slice_id == q->slice_id;
}

static int average_qp(int a, int b) { return (a + b + 1) >> 1; }

/* filter_picture, deblock.rs:84-142, with derive_boundary_strengths (:144-200),
 * filter_luma_macroblock (:202-276), filter_chroma_macroblock (:278-378) and the edge walkers
 * (:441-467). */
int oracle_h264_deblock(const mmr_recon_h264_picture *p) {
    int mw = (int)p->width_mbs, mh = (int)p->height_mbs;
    int W = mw * 16, CW = mw * 8;
    for (int a = 0; a < mw * mh; ++a) {
        const mmr_recon_h264_macroblock *m = &p->macroblocks[a];
        if (m->flags & MMR_H264_MB_NO_DEBLOCK) continue;
        int mx = a % mw, my = a / mw;
        int oa = m->offset_a, ob = m->offset_b;
        int sv[4][4], sh[4][4];
        memset(sv, 0, sizeof sv);
        memset(sh, 0, sizeof sh);
        for (int e = 0; e < 4; ++e) {
            if (e != 0 || (mx != 0 && external_edge_enabled(p, a - 1, a))) {
                int pa = e == 0 ? a - 1 : a;
                for (int s = 0; s < 4; ++s)
                    sv[e][s] = boundary_strength(p, pa, luma_block_index(e == 0 ? 3 : e - 1, s), a, luma_block_index(e == 0 ? 0 : e, s), e == 0);
            }
            if (e != 0 || (my != 0 && external_edge_enabled(p, a - mw, a))) {
                int pa = e == 0 ? a - mw : a;
                for (int s = 0; s < 4; ++s)
                    sh[e][s] = boundary_strength(p, pa, luma_block_index(s, e == 0 ? 3 : e - 1), a, luma_block_index(s, e == 0 ? 0 : e), e == 0);
            }
        }
        int t8 = (m->flags & MMR_H264_MB_TRANSFORM_8X8) != 0;
        /* luma vertical edges */
        for (int e = 0; e < 4; ++e) {
            if (e == 0 && (mx == 0 || !external_edge_enabled(p, a - 1, a))) continue;
            if (t8 && (e == 1 || e == 3)) continue;
            int pa = e == 0 ? a - 1 : a;
            int qp = average_qp(p->macroblocks[pa].qp, m->qp);
            for (int s = 0; s < 4; ++s) {
                if (sv[e][s] == 0) continue;
                for (int o = 0; o < 4; ++o) {
                    int q0 = (my * 16 + s * 4 + o) * W + mx * 16 + e * 4;
                    int pi[4] = {q0 - 1, q0 - 2, q0 - 3, q0 - 4}, qi[4] = {q0, q0 + 1, q0 + 2, q0 + 3};
                    oracle_h264_filter_samples(p->planes[0], pi, qi, qp, sv[e][s], 0, oa, ob);
                }
            }
        }
        /* luma horizontal edges */
        for (int e = 0; e < 4; ++e) {
            if (e == 0 && (my == 0 || !external_edge_enabled(p, a - mw, a))) continue;
            if (t8 && (e == 1 || e == 3)) continue;
            int pa = e == 0 ? a - mw : a;
            int qp = average_qp(p->macroblocks[pa].qp, m->qp);
            for (int s = 0; s < 4; ++s) {
                if (sh[e][s] == 0) continue;
                for (int o = 0; o < 4; ++o) {
                    int q0 = (my * 16 + e * 4) * W + mx * 16 + s * 4 + o;
                    int pi[4] = {q0 - W, q0 - 2 * W, q0 - 3 * W, q0 - 4 * W}, qi[4] = {q0, q0 + W, q0 + 2 * W, q0 + 3 * W};
                    oracle_h264_filter_samples(p->planes[0], pi, qi, qp, sh[e][s], 0, oa, ob);
                }
            }
        }
        /* chroma: Cb then Cr, vertical edges 0/2 then horizontal edges 0/2 of the luma strengths */
        for (int comp = 0; comp < 2; ++comp) {
            uint8_t *plane = p->planes[1 + comp];
            int off = p->chroma_qp_offset[comp];
            for (int e = 0; e < 2; ++e) {
                if (e == 0 && (mx == 0 || !external_edge_enabled(p, a - 1, a))) continue;
                int pa = e == 0 ? a - 1 : a;
                int qp = average_qp(oracle_h264_chroma_qp(p->macroblocks[pa].qp, off), oracle_h264_chroma_qp(m->qp, off));
                for (int s = 0; s < 4; ++s) {
                    int bs = sv[e * 2][s];
                    if (bs == 0) continue;
                    for (int o = 0; o < 2; ++o) {
                        int q0 = (my * 8 + s * 2 + o) * CW + mx * 8 + e * 4;
                        int pi[4] = {q0 - 1, q0 - 2, q0 - 3, q0 - 4}, qi[4] = {q0, q0 + 1, q0 + 2, q0 + 3};
                        oracle_h264_filter_samples(plane, pi, qi, qp, bs, 1, oa, ob);
                    }
                }
            }
            for (int e = 0; e < 2; ++e) {
                if (e == 0 && (my == 0 || !external_edge_enabled(p, a - mw, a))) continue;
                int pa = e == 0 ? a - mw : a;
                int qp = average_qp(oracle_h264_chroma_qp(p->macroblocks[pa].qp, off), oracle_h264_chroma_qp(m->qp, off));
                for (int s = 0; s < 4; ++s) {
                    int bs = sh[e * 2][s];
                    if (bs == 0) continue;
                    for (int o = 0; o < 2; ++o) {
                        int q0 = (my * 8 + e * 4) * CW + mx * 8 + s * 2 + o;
                        int pi[4] = {q0 - CW, q0 - 2 * CW, q0 - 3 * CW, q0 - 4 * CW}, qi[4] = {q0, q0 + CW, q0 + 2 * CW, q0 + 3 * CW};
                        oracle_h264_filter_samples(plane, pi, qi, qp, bs, 1, oa, ob);
                    }
                }
            }
        }
    }
    return 0;
}
