/* Oracle: H.264 CAVLC front end (test infrastructure only). Parses Baseline / Main / High frame
 * streams coded with CAVLC into the library's H.264 records (include/mmr_recon.h), so that bitstreams
 * reach the device reconstruction path. After crates/codecs/h264/src/{parser.rs, decoder.rs,
 * cavlc.rs}: NAL / RBSP handling, sequence and picture parameter sets, slice headers with reference
 * list modification, prediction weight tables and reference marking, picture order count types 0-2,
 * macroblock layer (decoder.rs decode_p_macroblocks / decode_b_macroblocks / decode_intra_nxn /
 * decode_inter_residuals :8097-8173, :10279-10420), CAVLC residual blocks (cavlc.rs:119-221), motion
 * vector prediction (ITU-T H.264 8.4.1.3) and P_Skip (8.4.1.1).
 * Not supported (reported as -2): CABAC, interlace / MBAFF, slice groups, scaling matrices,
 * B_Skip / B_Direct (direct prediction), SP / SI slices, memory_management_control_operation 5,
 * chroma formats other than 4:2:0, bit depths above 8. */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* ------------------------------------------------------------------ bits */
typedef struct {
    const uint8_t *d;
    size_t bits;     /* position of the rbsp_stop_one_bit (bits before it are payload) */
    size_t pos;
    int err;
} Bits;

static uint32_t rd_bit(Bits *b) {
    if (b->pos >= b->bits) {
        b->err = 1;
        return 0;
    }
    uint32_t v = (b->d[b->pos >> 3] >> (7 - (b->pos & 7))) & 1u;
    ++b->pos;
    return v;
}

static uint32_t rd_u(Bits *b, int n) {
    uint32_t v = 0;
    for (int i = 0; i < n; ++i) v = (v << 1) | rd_bit(b);
    return v;
}

static uint32_t rd_ue(Bits *b) {
    int zeros = 0;
    while (!b->err && rd_bit(b) == 0) {
        if (++zeros > 31) {
            b->err = 1;
            return 0;
        }
    }
    if (b->err) return 0;
    return (uint32_t)((1ull << zeros) - 1 + rd_u(b, zeros));
}

static int32_t rd_se(Bits *b) {
    uint32_t k = rd_ue(b);
    return (k & 1) ? (int32_t)((k + 1) / 2) : -(int32_t)(k / 2);
}

/* te(v) with range r (7.4.2.2): one inverted bit when r == 1, else ue(v). */
static uint32_t rd_te(Bits *b, uint32_t r) { return r == 1 ? 1u - rd_bit(b) : rd_ue(b); }

static int more_rbsp_data(const Bits *b) { return b->pos < b->bits; }

/* Emulation prevention removed (7.4.1: 0x000003 -> 0x0000); the stop bit located. */
static uint8_t *unescape(const uint8_t *nal, size_t n, Bits *b) {
    uint8_t *out = (uint8_t *)malloc(n ? n : 1);
    size_t m = 0, zeros = 0;
    for (size_t i = 1; i < n; ++i) {   /* skip the NAL header byte */
        if (zeros >= 2 && nal[i] == 3) {
            zeros = 0;
            continue;
        }
        out[m++] = nal[i];
        zeros = nal[i] == 0 ? zeros + 1 : 0;
    }
    while (m && out[m - 1] == 0) --m;
    size_t bits = m * 8;
    if (m) {
        int t = 0;
        while (!((out[m - 1] >> t) & 1)) ++t;
        bits = m * 8 - (size_t)t - 1;
    }
    b->d = out;
    b->bits = bits;
    b->pos = 0;
    b->err = 0;
    return out;
}

/* ------------------------------------------------------------------ parameter sets */
typedef struct {
    int valid, profile, chroma_format_idc, transform_bypass;
    int log2_max_frame_num, poc_type, log2_max_poc_lsb, delta_pic_order_always_zero;
    int offset_for_non_ref_pic, offset_for_top_to_bottom, cycle, offset_for_ref_frame[256];
    int max_num_ref_frames, width_mbs, height_mbs, frame_mbs_only, direct_8x8_inference;
} Sps;

typedef struct {
    int valid, sps_id, bottom_field_pic_order_present, num_ref_idx_default[2], weighted_pred, weighted_bipred_idc;
    int pic_init_qp, chroma_qp_offset[2], deblocking_filter_control_present, constrained_intra_pred;
    int redundant_pic_cnt_present, transform_8x8_mode;
} Pps;

static int parse_sps(Bits *b, Sps *all) {
    Sps s;
    memset(&s, 0, sizeof s);
    s.profile = (int)rd_u(b, 8);
    rd_u(b, 16);   /* constraint flags, level_idc */
    uint32_t id = rd_ue(b);
    if (id > 31) return -1;
    s.chroma_format_idc = 1;
    const int p = s.profile;
    if (p == 100 || p == 110 || p == 122 || p == 244 || p == 44 || p == 83 || p == 86 || p == 118 || p == 128 || p == 138 ||
        p == 139 || p == 134 || p == 135) {
        s.chroma_format_idc = (int)rd_ue(b);
        if (s.chroma_format_idc == 3) rd_bit(b);
        if (rd_ue(b) != 0 || rd_ue(b) != 0) return -2;   /* bit depths above 8 */
        s.transform_bypass = (int)rd_bit(b);
        if (rd_bit(b)) return -2;                         /* seq_scaling_matrix_present_flag */
    }
    if (s.chroma_format_idc != 1) return -2;
    s.log2_max_frame_num = (int)rd_ue(b) + 4;
    s.poc_type = (int)rd_ue(b);
    if (s.poc_type == 0) {
        s.log2_max_poc_lsb = (int)rd_ue(b) + 4;
    } else if (s.poc_type == 1) {
        s.delta_pic_order_always_zero = (int)rd_bit(b);
        s.offset_for_non_ref_pic = rd_se(b);
        s.offset_for_top_to_bottom = rd_se(b);
        s.cycle = (int)rd_ue(b);
        if (s.cycle > 255) return -1;
        for (int i = 0; i < s.cycle; ++i) s.offset_for_ref_frame[i] = rd_se(b);
    } else if (s.poc_type != 2) {
        return -1;
    }
    s.max_num_ref_frames = (int)rd_ue(b);
    rd_bit(b);   /* gaps_in_frame_num_value_allowed_flag */
    s.width_mbs = (int)rd_ue(b) + 1;
    s.height_mbs = (int)rd_ue(b) + 1;
    s.frame_mbs_only = (int)rd_bit(b);
    if (!s.frame_mbs_only) return -2;
    s.direct_8x8_inference = (int)rd_bit(b);
    if (rd_bit(b)) {   /* frame cropping: the records keep the coded size */
        rd_ue(b);
        rd_ue(b);
        rd_ue(b);
        rd_ue(b);
    }
    if (b->err || s.width_mbs > 1024 || s.height_mbs > 1024) return -1;
    s.valid = 1;
    all[id] = s;
    return 0;
}

static int parse_pps(Bits *b, Pps *all, const Sps *sps) {
    Pps q;
    memset(&q, 0, sizeof q);
    uint32_t id = rd_ue(b);
    q.sps_id = (int)rd_ue(b);
    if (id > 255 || q.sps_id > 31) return -1;
    if (rd_bit(b)) return -2;   /* entropy_coding_mode_flag: CABAC */
    q.bottom_field_pic_order_present = (int)rd_bit(b);
    if (rd_ue(b) != 0) return -2;   /* slice groups */
    q.num_ref_idx_default[0] = (int)rd_ue(b) + 1;
    q.num_ref_idx_default[1] = (int)rd_ue(b) + 1;
    q.weighted_pred = (int)rd_bit(b);
    q.weighted_bipred_idc = (int)rd_u(b, 2);
    q.pic_init_qp = 26 + rd_se(b);
    rd_se(b);   /* pic_init_qs */
    q.chroma_qp_offset[0] = rd_se(b);
    q.deblocking_filter_control_present = (int)rd_bit(b);
    q.constrained_intra_pred = (int)rd_bit(b);
    q.redundant_pic_cnt_present = (int)rd_bit(b);
    q.chroma_qp_offset[1] = q.chroma_qp_offset[0];
    if (more_rbsp_data(b)) {
        q.transform_8x8_mode = (int)rd_bit(b);
        if (rd_bit(b)) return -2;   /* pic_scaling_matrix_present_flag */
        q.chroma_qp_offset[1] = rd_se(b);
    }
    if (b->err || q.num_ref_idx_default[0] > 32 || q.num_ref_idx_default[1] > 32) return -1;
    (void)sps;
    q.valid = 1;
    all[id] = q;
    return 0;
}

/* ------------------------------------------------------------------ decoder state */
#define MAX_DPB 17

typedef struct {
    int used;           /* holds a reference picture */
    int long_term, long_term_idx;
    int frame_num, frame_num_wrap, pic_num;
    int poc, slot;
} Ref;

enum { MBK_NONE = 0, MBK_P_SKIP, MBK_INTER, MBK_I4, MBK_I8, MBK_I16, MBK_PCM };

typedef struct {
    uint8_t kind, slice_ok, transform8, cbp;
    int slice;
    int qp;
    uint8_t tc[16];             /* luma total_coeff per 4x4 (block index order) */
    uint8_t tc_c[2][4];         /* chroma AC total_coeff */
    uint8_t imode[16];          /* Intra4x4 / Intra8x8 modes per 4x4 (2 for other macroblocks) */
    int16_t mv[2][16][2];
    int8_t ref[2][16];
    int16_t mvd[2][16][2];
} MbState;

struct oracle_h264_decoder {
    Sps sps[32];
    Pps pps[256];
    Ref dpb[MAX_DPB];
    int next_slot;
    int prev_poc_msb, prev_poc_lsb, prev_frame_num, prev_frame_num_offset, prev_ref_has_mmco5;
    /* current picture */
    const Sps *s;
    const Pps *p;
    int width_mbs, height_mbs, mbs;
    MbState *mb;
    mmr_recon_h264_macroblock *rec;
    mmr_recon_h264_block *blk;
    mmr_recon_h264_intra *intra;
    int16_t *levels;
    uint32_t coded, levels_cap;
    mmr_recon_h264_weight weights[2][MMR_H264_MAX_REFS][3];
    mmr_recon_h264_bi_weight bi[MMR_H264_MAX_REFS][MMR_H264_MAX_REFS][3];
    oracle_h264_picture_info info;
    int have_picture;
};

oracle_h264_decoder *oracle_h264_decoder_new(void) {
    oracle_h264_decoder *d = (oracle_h264_decoder *)calloc(1, sizeof *d);
    return d;
}

static void free_picture(oracle_h264_decoder *d) {
    free(d->mb);
    free(d->rec);
    free(d->blk);
    free(d->intra);
    free(d->levels);
    d->mb = NULL;
    d->rec = NULL;
    d->blk = NULL;
    d->intra = NULL;
    d->levels = NULL;
    d->coded = d->levels_cap = 0;
}

void oracle_h264_decoder_free(oracle_h264_decoder *d) {
    if (!d) return;
    free_picture(d);
    free(d);
}

/* ------------------------------------------------------------------ slice header */
typedef struct {
    int first_mb, type, pps_id, frame_num, idr, idr_pic_id, nal_ref_idc;
    int poc_lsb, delta_poc_bottom, delta_poc[2];
    int direct_spatial, num_ref_idx_active[2];
    int mod_count[2], mod_idc[2][33], mod_val[2][33];
    int luma_log2_denom, chroma_log2_denom;
    int lw[2][32], lo[2][32], cw[2][32][2], co[2][32][2];
    int long_term_reference_flag, adaptive_marking;
    int mmco_count, mmco[66][3];
    int qp, disable_deblocking, alpha_off, beta_off;
} SliceHeader;

static int parse_slice_header(oracle_h264_decoder *d, Bits *b, int nal_type, int nal_ref_idc, SliceHeader *h) {
    memset(h, 0, sizeof *h);
    h->nal_ref_idc = nal_ref_idc;
    h->idr = nal_type == 5;
    h->first_mb = (int)rd_ue(b);
    h->type = (int)(rd_ue(b) % 5);
    if (h->type > 2) return -2;   /* SP / SI */
    h->pps_id = (int)rd_ue(b);
    if (h->pps_id > 255 || !d->pps[h->pps_id].valid || !d->sps[d->pps[h->pps_id].sps_id].valid) return -1;
    const Pps *p = &d->pps[h->pps_id];
    const Sps *s = &d->sps[p->sps_id];
    h->frame_num = (int)rd_u(b, s->log2_max_frame_num);
    if (h->idr) h->idr_pic_id = (int)rd_ue(b);
    if (s->poc_type == 0) {
        h->poc_lsb = (int)rd_u(b, s->log2_max_poc_lsb);
        if (p->bottom_field_pic_order_present) h->delta_poc_bottom = rd_se(b);
    } else if (s->poc_type == 1 && !s->delta_pic_order_always_zero) {
        h->delta_poc[0] = rd_se(b);
        if (p->bottom_field_pic_order_present) h->delta_poc[1] = rd_se(b);
    }
    if (p->redundant_pic_cnt_present && rd_ue(b) != 0) return -2;
    if (h->type == 1) h->direct_spatial = (int)rd_bit(b);
    h->num_ref_idx_active[0] = p->num_ref_idx_default[0];
    h->num_ref_idx_active[1] = p->num_ref_idx_default[1];
    if (h->type != 2) {
        if (rd_bit(b)) {
            h->num_ref_idx_active[0] = (int)rd_ue(b) + 1;
            if (h->type == 1) h->num_ref_idx_active[1] = (int)rd_ue(b) + 1;
        }
        if (h->num_ref_idx_active[0] > 16 || h->num_ref_idx_active[1] > 16) return -1;
        for (int list = 0; list < (h->type == 1 ? 2 : 1); ++list) {
            if (!rd_bit(b)) continue;
            for (;;) {
                int idc = (int)rd_ue(b);
                if (idc == 3) break;
                if (idc > 3 || h->mod_count[list] >= 33 || b->err) return -1;
                h->mod_idc[list][h->mod_count[list]] = idc;
                h->mod_val[list][h->mod_count[list]] = (int)rd_ue(b);
                ++h->mod_count[list];
            }
        }
    }
    if (h->type != 1) h->num_ref_idx_active[1] = 0;
    if (h->type == 2) h->num_ref_idx_active[0] = 0;
    if ((p->weighted_pred && h->type == 0) || (p->weighted_bipred_idc == 1 && h->type == 1)) {
        h->luma_log2_denom = (int)rd_ue(b);
        h->chroma_log2_denom = (int)rd_ue(b);
        if (h->luma_log2_denom > 7 || h->chroma_log2_denom > 7) return -1;
        for (int list = 0; list < (h->type == 1 ? 2 : 1); ++list)
            for (int i = 0; i < h->num_ref_idx_active[list]; ++i) {
                h->lw[list][i] = 1 << h->luma_log2_denom;
                h->lo[list][i] = 0;
                if (rd_bit(b)) {
                    h->lw[list][i] = rd_se(b);
                    h->lo[list][i] = rd_se(b);
                }
                for (int c = 0; c < 2; ++c) {
                    h->cw[list][i][c] = 1 << h->chroma_log2_denom;
                    h->co[list][i][c] = 0;
                }
                if (rd_bit(b))
                    for (int c = 0; c < 2; ++c) {
                        h->cw[list][i][c] = rd_se(b);
                        h->co[list][i][c] = rd_se(b);
                    }
            }
    }
    if (nal_ref_idc) {
        if (h->idr) {
            rd_bit(b);   /* no_output_of_prior_pics_flag */
            h->long_term_reference_flag = (int)rd_bit(b);
        } else {
            h->adaptive_marking = (int)rd_bit(b);
            if (h->adaptive_marking)
                for (;;) {
                    int op = (int)rd_ue(b);
                    if (op == 0) break;
                    if (op > 6 || h->mmco_count >= 66 || b->err) return -1;
                    if (op == 5) return -2;
                    int *m = h->mmco[h->mmco_count++];
                    m[0] = op;
                    m[1] = (op == 1 || op == 3) ? (int)rd_ue(b) : 0;   /* difference_of_pic_nums_minus1 */
                    if (op == 2) m[1] = (int)rd_ue(b);                 /* long_term_pic_num */
                    m[2] = (op == 3 || op == 6 || op == 4) ? (int)rd_ue(b) : 0;
                }
        }
    }
    h->qp = p->pic_init_qp + rd_se(b);
    if (h->qp < 0 || h->qp > 51) return -1;
    if (p->deblocking_filter_control_present) {
        h->disable_deblocking = (int)rd_ue(b);
        if (h->disable_deblocking > 2) return -1;
        if (h->disable_deblocking != 1) {
            h->alpha_off = rd_se(b) * 2;
            h->beta_off = rd_se(b) * 2;
            if (h->alpha_off < -12 || h->alpha_off > 12 || h->beta_off < -12 || h->beta_off > 12) return -1;
        }
    }
    return b->err ? -1 : 0;
}

/* ------------------------------------------------------------------ POC (8.2.1) */
static int picture_order_count(oracle_h264_decoder *d, const SliceHeader *h, int *frame_num_offset) {
    const Sps *s = d->s;
    const int max_frame_num = 1 << s->log2_max_frame_num;
    *frame_num_offset = 0;
    if (s->poc_type == 0) {
        const int max_lsb = 1 << s->log2_max_poc_lsb;
        int prev_msb = h->idr ? 0 : d->prev_poc_msb, prev_lsb = h->idr ? 0 : d->prev_poc_lsb;
        int msb;
        if (h->poc_lsb < prev_lsb && prev_lsb - h->poc_lsb >= max_lsb / 2) msb = prev_msb + max_lsb;
        else if (h->poc_lsb > prev_lsb && h->poc_lsb - prev_lsb > max_lsb / 2) msb = prev_msb - max_lsb;
        else msb = prev_msb;
        const int top = msb + h->poc_lsb, bottom = top + h->delta_poc_bottom;
        if (h->nal_ref_idc) {
            d->prev_poc_msb = msb;
            d->prev_poc_lsb = h->poc_lsb;
        }
        return top < bottom ? top : bottom;
    }
    int offset = 0;
    if (!h->idr) offset = d->prev_frame_num > h->frame_num ? d->prev_frame_num_offset + max_frame_num : d->prev_frame_num_offset;
    *frame_num_offset = offset;
    if (s->poc_type == 2) {
        if (h->idr) return 0;
        return h->nal_ref_idc ? 2 * (offset + h->frame_num) : 2 * (offset + h->frame_num) - 1;
    }
    /* type 1 */
    int abs_frame_num = s->cycle ? offset + h->frame_num : 0;
    if (!h->nal_ref_idc && abs_frame_num > 0) --abs_frame_num;
    int expected = 0;
    if (abs_frame_num > 0) {
        int delta_cycle = 0;
        for (int i = 0; i < s->cycle; ++i) delta_cycle += s->offset_for_ref_frame[i];
        const int cycle_cnt = (abs_frame_num - 1) / s->cycle, in_cycle = (abs_frame_num - 1) % s->cycle;
        expected = cycle_cnt * delta_cycle;
        for (int i = 0; i <= in_cycle; ++i) expected += s->offset_for_ref_frame[i];
    }
    if (!h->nal_ref_idc) expected += s->offset_for_non_ref_pic;
    const int top = expected + h->delta_poc[0], bottom = top + s->offset_for_top_to_bottom + h->delta_poc[1];
    return top < bottom ? top : bottom;
}

/* ------------------------------------------------------------------ reference lists (8.2.4) */
static int cmp_desc_pic_num(const void *a, const void *b) { return (*(Ref *const *)b)->pic_num - (*(Ref *const *)a)->pic_num; }
static int cmp_asc_lt(const void *a, const void *b) { return (*(Ref *const *)a)->long_term_idx - (*(Ref *const *)b)->long_term_idx; }
static int cmp_desc_poc(const void *a, const void *b) { return (*(Ref *const *)b)->poc - (*(Ref *const *)a)->poc; }
static int cmp_asc_poc(const void *a, const void *b) { return (*(Ref *const *)a)->poc - (*(Ref *const *)b)->poc; }

static int build_lists(oracle_h264_decoder *d, const SliceHeader *h, int cur_poc, Ref *lists[2][33], int counts[2]) {
    const int max_frame_num = 1 << d->s->log2_max_frame_num;
    Ref *st[MAX_DPB], *lt[MAX_DPB];
    int ns = 0, nl = 0;
    for (int i = 0; i < MAX_DPB; ++i) {
        Ref *r = &d->dpb[i];
        if (!r->used) continue;
        if (r->long_term) {
            r->pic_num = r->long_term_idx;
            lt[nl++] = r;
        } else {
            r->frame_num_wrap = r->frame_num > h->frame_num ? r->frame_num - max_frame_num : r->frame_num;
            r->pic_num = r->frame_num_wrap;
            st[ns++] = r;
        }
    }
    qsort(lt, (size_t)nl, sizeof lt[0], cmp_asc_lt);
    Ref *init[2][34];
    int n[2] = {0, 0};
    if (h->type == 0) {
        qsort(st, (size_t)ns, sizeof st[0], cmp_desc_pic_num);
        for (int i = 0; i < ns; ++i) init[0][n[0]++] = st[i];
        for (int i = 0; i < nl; ++i) init[0][n[0]++] = lt[i];
    } else if (h->type == 1) {
        Ref *before[MAX_DPB], *after[MAX_DPB];
        int nb = 0, na = 0;
        for (int i = 0; i < ns; ++i) {
            if (st[i]->poc < cur_poc) before[nb++] = st[i];
            else after[na++] = st[i];
        }
        qsort(before, (size_t)nb, sizeof before[0], cmp_desc_poc);
        qsort(after, (size_t)na, sizeof after[0], cmp_asc_poc);
        for (int i = 0; i < nb; ++i) init[0][n[0]++] = before[i];
        for (int i = 0; i < na; ++i) init[0][n[0]++] = after[i];
        for (int i = 0; i < nl; ++i) init[0][n[0]++] = lt[i];
        for (int i = 0; i < na; ++i) init[1][n[1]++] = after[i];
        for (int i = 0; i < nb; ++i) init[1][n[1]++] = before[i];
        for (int i = 0; i < nl; ++i) init[1][n[1]++] = lt[i];
        if (n[1] > 1 && n[0] == n[1]) {
            int same = 1;
            for (int i = 0; i < n[0]; ++i) same &= init[0][i] == init[1][i];
            if (same) {
                Ref *t = init[1][0];
                init[1][0] = init[1][1];
                init[1][1] = t;
            }
        }
    }
    for (int list = 0; list < 2; ++list) {
        const int active = h->num_ref_idx_active[list];
        int len = n[list] < active ? n[list] : active;
        Ref *l[34];
        for (int i = 0; i < len; ++i) l[i] = init[list][i];
        for (int i = len; i < active; ++i) l[i] = NULL;
        /* modification (8.2.4.3): each command moves one picture to position refIdx */
        int pred = h->frame_num, ref_idx = 0;
        for (int k = 0; k < h->mod_count[list]; ++k) {
            const int idc = h->mod_idc[list][k];
            Ref *pick = NULL;
            if (idc < 2) {
                const int diff = h->mod_val[list][k] + 1;
                int no_wrap = idc == 0 ? pred - diff : pred + diff;
                if (no_wrap < 0) no_wrap += max_frame_num;
                if (no_wrap >= max_frame_num) no_wrap -= max_frame_num;
                pred = no_wrap;
                const int pic_num = no_wrap > h->frame_num ? no_wrap - max_frame_num : no_wrap;
                for (int i = 0; i < MAX_DPB; ++i)
                    if (d->dpb[i].used && !d->dpb[i].long_term && d->dpb[i].pic_num == pic_num) pick = &d->dpb[i];
            } else {
                for (int i = 0; i < MAX_DPB; ++i)
                    if (d->dpb[i].used && d->dpb[i].long_term && d->dpb[i].long_term_idx == h->mod_val[list][k]) pick = &d->dpb[i];
            }
            if (!pick || ref_idx >= active) return -1;
            for (int c = active; c > ref_idx; --c) l[c] = l[c - 1];
            l[ref_idx++] = pick;
            int nidx = ref_idx;
            for (int c = ref_idx; c <= active; ++c)
                if (l[c] != pick) l[nidx++] = l[c];
        }
        for (int i = 0; i < active; ++i)
            if (!l[i]) return -1;   /* a reference index with no picture */
        for (int i = 0; i < active; ++i) lists[list][i] = l[i];
        counts[list] = active;
    }
    return 0;
}

/* ------------------------------------------------------------------ marking (8.2.5) */
static void mark_current(oracle_h264_decoder *d, const SliceHeader *h, int poc, int slot) {
    if (!h->nal_ref_idc) return;
    const int max_frame_num = 1 << d->s->log2_max_frame_num;
    int long_term = 0, lt_idx = 0;
    if (h->idr) {
        for (int i = 0; i < MAX_DPB; ++i) d->dpb[i].used = 0;
        long_term = h->long_term_reference_flag;
    } else if (h->adaptive_marking) {
        for (int k = 0; k < h->mmco_count; ++k) {
            const int *m = h->mmco[k];
            for (int i = 0; i < MAX_DPB; ++i) {
                Ref *r = &d->dpb[i];
                if (!r->used) continue;
                const int wrap = r->frame_num > h->frame_num ? r->frame_num - max_frame_num : r->frame_num;
                const int pic_num_x = h->frame_num - (m[1] + 1);
                if (m[0] == 1 && !r->long_term && wrap == pic_num_x) r->used = 0;
                if (m[0] == 2 && r->long_term && r->long_term_idx == m[1]) r->used = 0;
                if (m[0] == 3 && r->long_term && r->long_term_idx == m[2]) r->used = 0;
                if (m[0] == 4 && r->long_term && r->long_term_idx >= m[2]) r->used = 0;
                if (m[0] == 6 && r->long_term && r->long_term_idx == m[2]) r->used = 0;
            }
            if (m[0] == 3)
                for (int i = 0; i < MAX_DPB; ++i) {
                    Ref *r = &d->dpb[i];
                    const int wrap = r->frame_num > h->frame_num ? r->frame_num - max_frame_num : r->frame_num;
                    if (r->used && !r->long_term && wrap == h->frame_num - (m[1] + 1)) {
                        r->long_term = 1;
                        r->long_term_idx = m[2];
                    }
                }
            if (m[0] == 6) {
                long_term = 1;
                lt_idx = m[2];
            }
        }
    } else {
        int n = 0, oldest = -1;
        for (int i = 0; i < MAX_DPB; ++i) {
            Ref *r = &d->dpb[i];
            if (!r->used) continue;
            ++n;
            if (!r->long_term) {
                const int wrap = r->frame_num > h->frame_num ? r->frame_num - max_frame_num : r->frame_num;
                if (oldest < 0 || wrap < d->dpb[oldest].frame_num_wrap) {
                    oldest = i;
                    d->dpb[i].frame_num_wrap = wrap;
                }
            }
        }
        const int cap = d->s->max_num_ref_frames > 1 ? d->s->max_num_ref_frames : 1;
        if (n >= cap && oldest >= 0) d->dpb[oldest].used = 0;
    }
    for (int i = 0; i < MAX_DPB; ++i)
        if (!d->dpb[i].used) {
            Ref *r = &d->dpb[i];
            memset(r, 0, sizeof *r);
            r->used = 1;
            r->long_term = long_term;
            r->long_term_idx = lt_idx;
            r->frame_num = h->frame_num;
            r->poc = poc;
            r->slot = slot;
            return;
        }
}

/* ------------------------------------------------------------------ CAVLC tables
 * ITU-T H.264 Tables 9-5 (coeff_token: lengths and codes per nC class, index 4 * TotalCoeff +
 * TrailingOnes; chroma DC nC = -1), 9-7 / 9-8 (total_zeros), 9-9 (chroma DC total_zeros) and
 * 9-10 (run_before); the same numbers as crates/codecs/h264/src/cavlc.rs:4-115. */
static const uint8_t CT_LEN[4][68] = {
    {1, 0, 0, 0, 6, 2, 0, 0, 8, 6, 3, 0, 9, 8, 7, 5, 10, 9, 8, 6, 11, 10, 9, 7, 13, 11, 10, 8, 13, 13, 11, 9, 13, 13, 13, 10, 14,
     14, 13, 11, 14, 14, 14, 13, 15, 15, 14, 14, 15, 15, 15, 14, 16, 15, 15, 15, 16, 16, 16, 15, 16, 16, 16, 16, 16, 16, 16, 16},
    {2, 0, 0, 0, 6, 2, 0, 0, 6, 5, 3, 0, 7, 6, 6, 4, 8, 6, 6, 4, 8, 7, 7, 5, 9, 8, 8, 6, 11, 9, 9, 6, 11, 11, 11, 7, 12, 11, 11,
     9, 12, 12, 12, 11, 12, 12, 12, 11, 13, 13, 13, 12, 13, 13, 13, 13, 13, 14, 13, 13, 14, 14, 14, 13, 14, 14, 14, 14},
    {4, 0, 0, 0, 6, 4, 0, 0, 6, 5, 4, 0, 6, 5, 5, 4, 7, 5, 5, 4, 7, 5, 5, 4, 7, 6, 6, 4, 7, 6, 6, 4, 8, 7, 7, 5, 8, 8, 7, 6, 9,
     8, 8, 7, 9, 9, 8, 8, 9, 9, 9, 8, 10, 9, 9, 9, 10, 10, 10, 10, 10, 10, 10, 10, 10, 10, 10, 10},
    {6, 0, 0, 0, 6, 6, 0, 0, 6, 6, 6, 0, 6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 6,
     6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 6}};
static const uint8_t CT_CODE[4][68] = {
    {1, 0, 0, 0, 5, 1, 0, 0, 7, 4, 1, 0, 7, 6, 5, 3, 7, 6, 5, 3, 7, 6, 5, 4, 15, 6, 5, 4, 11, 14, 5, 4, 8, 10, 13, 4, 15, 14,
     9, 4, 11, 10, 13, 12, 15, 14, 9, 12, 11, 10, 13, 8, 15, 1, 9, 12, 11, 14, 13, 8, 7, 10, 9, 12, 4, 6, 5, 8},
    {3, 0, 0, 0, 11, 2, 0, 0, 7, 7, 3, 0, 7, 10, 9, 5, 7, 6, 5, 4, 4, 6, 5, 6, 7, 6, 5, 8, 15, 6, 5, 4, 11, 14, 13, 4, 15, 10,
     9, 4, 11, 14, 13, 12, 8, 10, 9, 8, 15, 14, 13, 12, 11, 10, 9, 12, 7, 11, 6, 8, 9, 8, 10, 1, 7, 6, 5, 4},
    {15, 0, 0, 0, 15, 14, 0, 0, 11, 15, 13, 0, 8, 12, 14, 12, 15, 10, 11, 11, 11, 8, 9, 10, 9, 14, 13, 9, 8, 10, 9, 8, 15, 14,
     13, 13, 11, 14, 10, 12, 15, 10, 13, 12, 11, 14, 9, 12, 8, 10, 13, 8, 13, 7, 9, 12, 9, 12, 11, 10, 5, 8, 7, 6, 1, 4, 3, 2},
    {3, 0, 0, 0, 0, 1, 0, 0, 4, 5, 6, 0, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28, 29,
     30, 31, 32, 33, 34, 35, 36, 37, 38, 39, 40, 41, 42, 43, 44, 45, 46, 47, 48, 49, 50, 51, 52, 53, 54, 55, 56, 57, 58, 59, 60,
     61, 62, 63}};
static const uint8_t CDC_LEN[20] = {2, 0, 0, 0, 6, 1, 0, 0, 6, 6, 3, 0, 6, 7, 7, 6, 6, 8, 8, 7};
static const uint8_t CDC_CODE[20] = {1, 0, 0, 0, 7, 1, 0, 0, 4, 6, 1, 0, 3, 3, 2, 5, 2, 3, 2, 0};
static const uint8_t TZ_LEN[15][16] = {
    {1, 3, 3, 4, 4, 5, 5, 6, 6, 7, 7, 8, 8, 9, 9, 9}, {3, 3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 6, 6, 6, 6, 0},
    {4, 3, 3, 3, 4, 4, 3, 3, 4, 5, 5, 6, 5, 6, 0, 0}, {5, 3, 4, 4, 3, 3, 3, 4, 3, 4, 5, 5, 5, 0, 0, 0},
    {4, 4, 4, 3, 3, 3, 3, 3, 4, 5, 4, 5, 0, 0, 0, 0}, {6, 5, 3, 3, 3, 3, 3, 3, 4, 3, 6, 0, 0, 0, 0, 0},
    {6, 5, 3, 3, 3, 2, 3, 4, 3, 6, 0, 0, 0, 0, 0, 0}, {6, 4, 5, 3, 2, 2, 3, 3, 6, 0, 0, 0, 0, 0, 0, 0},
    {6, 6, 4, 2, 2, 3, 2, 5, 0, 0, 0, 0, 0, 0, 0, 0}, {5, 5, 3, 2, 2, 2, 4, 0, 0, 0, 0, 0, 0, 0, 0, 0},
    {4, 4, 3, 3, 1, 3, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0}, {4, 4, 2, 1, 3, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
    {3, 3, 1, 2, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0}, {2, 2, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
    {1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0}};
static const uint8_t TZ_CODE[15][16] = {
    {1, 3, 2, 3, 2, 3, 2, 3, 2, 3, 2, 3, 2, 3, 2, 1}, {7, 6, 5, 4, 3, 5, 4, 3, 2, 3, 2, 3, 2, 1, 0, 0},
    {5, 7, 6, 5, 4, 3, 4, 3, 2, 3, 2, 1, 1, 0, 0, 0}, {3, 7, 5, 4, 6, 5, 4, 3, 3, 2, 2, 1, 0, 0, 0, 0},
    {5, 4, 3, 7, 6, 5, 4, 3, 2, 1, 1, 0, 0, 0, 0, 0}, {1, 1, 7, 6, 5, 4, 3, 2, 1, 1, 0, 0, 0, 0, 0, 0},
    {1, 1, 5, 4, 3, 3, 2, 1, 1, 0, 0, 0, 0, 0, 0, 0}, {1, 1, 1, 3, 3, 2, 2, 1, 0, 0, 0, 0, 0, 0, 0, 0},
    {1, 0, 1, 3, 2, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0, 0}, {1, 0, 1, 3, 2, 1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0},
    {0, 1, 1, 2, 1, 3, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0}, {0, 1, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
    {0, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0}, {0, 1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
    {0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0}};
static const uint8_t CTZ_LEN[3][4] = {{1, 2, 3, 3}, {1, 2, 2, 0}, {1, 1, 0, 0}};
static const uint8_t CTZ_CODE[3][4] = {{1, 1, 1, 0}, {1, 1, 0, 0}, {1, 0, 0, 0}};
static const uint8_t RUN_LEN[7][16] = {
    {1, 1, 0}, {1, 2, 2, 0}, {2, 2, 2, 2, 0}, {2, 2, 2, 3, 3, 0}, {2, 2, 3, 3, 3, 3, 0}, {2, 3, 3, 3, 3, 3, 3, 0},
    {3, 3, 3, 3, 3, 3, 3, 4, 5, 6, 7, 8, 9, 10, 11, 0}};
static const uint8_t RUN_CODE[7][16] = {
    {1, 0}, {1, 1, 0}, {3, 2, 1, 0}, {3, 2, 1, 1, 0}, {3, 2, 3, 2, 1, 0}, {3, 0, 1, 3, 2, 5, 4},
    {7, 6, 5, 4, 3, 2, 1, 1, 1, 1, 1, 1, 1, 1, 1}};

/* The entry of a prefix code (lens / codes, n entries) the next bits spell; -1 when none does. */
static int read_vlc(Bits *b, const uint8_t *lens, const uint8_t *codes, int n) {
    uint32_t v = 0;
    for (int len = 1; len <= 16; ++len) {
        v = (v << 1) | rd_bit(b);
        if (b->err) return -1;
        for (int i = 0; i < n; ++i)
            if (lens[i] == len && codes[i] == v) return i;
    }
    return -1;
}

/* residual_block_cavlc (7.3.5.3.3, 9.2; cavlc.rs:119-221): coefficients in scan order. Returns
 * TotalCoeff or -1. */
static int residual_block(Bits *b, int nc, int max_coeff, int32_t *coef) {
    for (int i = 0; i < max_coeff; ++i) coef[i] = 0;
    int token;
    if (nc == -1) token = read_vlc(b, CDC_LEN, CDC_CODE, 20);
    else token = read_vlc(b, CT_LEN[nc < 2 ? 0 : nc < 4 ? 1 : nc < 8 ? 2 : 3], CT_CODE[nc < 2 ? 0 : nc < 4 ? 1 : nc < 8 ? 2 : 3], 68);
    if (token < 0) return -1;
    const int total = token / 4, t1 = token % 4;
    if (total > max_coeff || t1 > (total < 3 ? total : 3)) return -1;
    if (total == 0) return 0;
    int32_t level[16];
    for (int i = 0; i < t1; ++i) level[i] = rd_bit(b) ? -1 : 1;
    int suffix_length = total > 10 && t1 < 3 ? 1 : 0;
    for (int i = t1; i < total; ++i) {
        int prefix = 0;
        while (!rd_bit(b)) {
            if (b->err || ++prefix > 31) return -1;
        }
        const int suffix_size = (prefix == 14 && suffix_length == 0) ? 4 : prefix >= 15 ? prefix - 3 : suffix_length;
        int64_t code = ((int64_t)(prefix < 15 ? prefix : 15) << suffix_length) + (suffix_size ? rd_u(b, suffix_size) : 0);
        if (prefix >= 15 && suffix_length == 0) code += 15;
        if (prefix >= 16) code += ((int64_t)1 << (prefix - 3)) - 4096;
        if (i == t1 && t1 < 3) code += 2;
        const int64_t lv = code % 2 == 0 ? (code + 2) >> 1 : (-code - 1) >> 1;
        if (lv > INT32_MAX || lv < INT32_MIN) return -1;
        level[i] = (int32_t)lv;
        if (suffix_length == 0) suffix_length = 1;
        if ((lv < 0 ? -lv : lv) > (3 << (suffix_length - 1)) && suffix_length < 6) ++suffix_length;
    }
    int zeros_left = 0;
    if (total < max_coeff) {
        if (nc == -1) zeros_left = read_vlc(b, CTZ_LEN[total - 1], CTZ_CODE[total - 1], 4 - total + 1);
        else zeros_left = read_vlc(b, TZ_LEN[total - 1], TZ_CODE[total - 1], 16 - total + 1);
        if (zeros_left < 0) return -1;
    }
    int run[16];
    for (int i = 0; i < total - 1; ++i) {
        int r = 0;
        if (zeros_left > 0) {
            const int t = zeros_left > 6 ? 6 : zeros_left - 1;
            r = read_vlc(b, RUN_LEN[t], RUN_CODE[t], zeros_left > 6 ? 15 : zeros_left + 1);
            if (r < 0 || r > zeros_left) return -1;
        }
        run[i] = r;
        zeros_left -= r;
    }
    run[total - 1] = zeros_left;
    int pos = -1;
    for (int i = total - 1; i >= 0; --i) {
        pos += run[i] + 1;
        if (pos >= max_coeff) return -1;
        coef[pos] = level[i];
    }
    return b->err ? -1 : total;
}

/* ------------------------------------------------------------------ macroblock tables */
static const uint8_t CBP_INTRA[48] = {47, 31, 15, 0,  23, 27, 29, 30, 7,  11, 13, 14, 39, 43, 45, 46,
                                      16, 3,  5,  10, 12, 19, 21, 26, 28, 35, 37, 42, 44, 1,  2,  4,
                                      8,  17, 18, 20, 24, 6,  9,  22, 25, 32, 33, 34, 36, 40, 38, 41};
static const uint8_t CBP_INTER[48] = {0,  16, 1,  2,  4,  8,  32, 3,  5,  10, 12, 15, 47, 7,  11, 13,
                                      14, 6,  9,  31, 35, 37, 42, 44, 33, 34, 36, 40, 39, 43, 45, 46,
                                      17, 18, 20, 24, 19, 21, 26, 28, 23, 27, 29, 30, 22, 25, 38, 41};

/* Block index of the 4x4 block at (x, y) in 4x4 units inside a macroblock (6.4.3). */
static int blk_idx(int x, int y) { return (y / 2) * 8 + (x / 2) * 4 + (y % 2) * 2 + x % 2; }

/* Partition prediction: 0 L0, 1 L1, 2 Bi. */
enum { PRED_L0 = 0, PRED_L1 = 1, PRED_BI = 2 };

/* B macroblock types 1..21 (Table 7-14): partition shape (0 16x16, 1 16x8, 2 8x16) and the two
 * partitions' predictions. */
static const int8_t B_SHAPE[22] = {-1, 0, 0, 0, 1, 2, 1, 2, 1, 2, 1, 2, 1, 2, 1, 2, 1, 2, 1, 2, 1, 2};
static const int8_t B_PRED[22][2] = {{-1, -1}, {0, -1}, {1, -1}, {2, -1}, {0, 0}, {0, 0}, {1, 1}, {1, 1}, {0, 1}, {0, 1}, {1, 0},
                                     {1, 0},   {0, 2},  {0, 2},  {1, 2},  {1, 2}, {2, 0}, {2, 0}, {2, 1}, {2, 1}, {2, 2}, {2, 2}};
/* P / B sub-macroblock types (Tables 7-17, 7-18): shape (0 8x8, 1 8x4, 2 4x8, 3 4x4) and prediction. */
static const int8_t PSUB_SHAPE[4] = {0, 1, 2, 3};
static const int8_t BSUB_SHAPE[13] = {-1, 0, 0, 0, 1, 2, 1, 2, 1, 2, 3, 3, 3};
static const int8_t BSUB_PRED[13] = {-1, 0, 1, 2, 0, 0, 1, 1, 2, 2, 0, 1, 2};

/* ------------------------------------------------------------------ neighbours (6.4.11) */
typedef struct {
    int avail;   /* the 4x4 block exists and is already decoded in this slice */
    int mb, blk;
} Nb;

/* The 4x4 block at (x, y) in 4x4 units relative to macroblock `mb` (x in -1..4, y in -1..3);
 * `cur` marks the current macroblock's blocks already decoded. */
static Nb nb4(const oracle_h264_decoder *d, int mb, int x, int y, const uint8_t *cur) {
    Nb n = {0, -1, -1};
    if (x >= 0 && x < 4 && y >= 0 && y < 4) {
        n.avail = cur ? cur[blk_idx(x, y)] : 1;
        n.mb = mb;
        n.blk = blk_idx(x, y);
        return n;
    }
    if (x > 3 && y >= 0) return n;   /* right of the macroblock, not above: not yet decoded */
    const int W = d->width_mbs, mx = mb % W + (x < 0 ? -1 : x > 3 ? 1 : 0), my = mb / W + (y < 0 ? -1 : 0);
    if (mx < 0 || mx >= W || my < 0) return n;
    const int a = my * W + mx;
    if (a >= mb || d->mb[a].kind == MBK_NONE || d->mb[a].slice != d->mb[mb].slice) return n;
    n.avail = 1;
    n.mb = a;
    n.blk = blk_idx((x + 4) % 4, (y + 4) % 4);
    return n;
}

static int is_intra_kind(int k) { return k == MBK_I4 || k == MBK_I8 || k == MBK_I16 || k == MBK_PCM; }

/* refIdx and mv of list `list` at a neighbour (-1 / 0 when unavailable, intra or not predicted). */
static void nb_motion(const oracle_h264_decoder *d, Nb n, int list, int *ref, int mv[2]) {
    *ref = -1;
    mv[0] = mv[1] = 0;
    if (!n.avail || is_intra_kind(d->mb[n.mb].kind)) return;
    *ref = d->mb[n.mb].ref[list][n.blk];
    if (*ref >= 0) {
        mv[0] = d->mb[n.mb].mv[list][n.blk][0];
        mv[1] = d->mb[n.mb].mv[list][n.blk][1];
    }
}

static int median3(int a, int b, int c) {
    const int mx = a > b ? (a > c ? a : c) : (b > c ? b : c), mn = a < b ? (a < c ? a : c) : (b < c ? b : c);
    return a + b + c - mx - mn;
}

/* Luma motion vector prediction (8.4.1.3) for the partition at (x, y), w x h in 4x4 units.
 * shape: 1 16x8, 2 8x16 (directional rules), 0 otherwise. */
static void predict_mv(const oracle_h264_decoder *d, int mb, int list, int ref, int x, int y, int w, int shape, const uint8_t *cur,
                       int out[2]) {
    Nb A = nb4(d, mb, x - 1, y, cur), B = nb4(d, mb, x, y - 1, cur), C = nb4(d, mb, x + w, y - 1, cur);
    if (!C.avail) C = nb4(d, mb, x - 1, y - 1, cur);
    int ra, rb, rc, ma[2], mbv[2], mc[2];
    nb_motion(d, A, list, &ra, ma);
    nb_motion(d, B, list, &rb, mbv);
    nb_motion(d, C, list, &rc, mc);
    if (!B.avail && !C.avail && A.avail) {
        rb = rc = ra;
        mbv[0] = mc[0] = ma[0];
        mbv[1] = mc[1] = ma[1];
    }
    if (shape == 1 && y == 0 && rb == ref) { out[0] = mbv[0]; out[1] = mbv[1]; return; }
    if (shape == 1 && y != 0 && ra == ref) { out[0] = ma[0]; out[1] = ma[1]; return; }
    if (shape == 2 && x == 0 && ra == ref) { out[0] = ma[0]; out[1] = ma[1]; return; }
    if (shape == 2 && x != 0 && rc == ref) { out[0] = mc[0]; out[1] = mc[1]; return; }
    const int same = (ra == ref) + (rb == ref) + (rc == ref);
    if (same == 1) {
        const int *m = ra == ref ? ma : rb == ref ? mbv : mc;
        out[0] = m[0];
        out[1] = m[1];
        return;
    }
    out[0] = median3(ma[0], mbv[0], mc[0]);
    out[1] = median3(ma[1], mbv[1], mc[1]);
}

/* nC of a luma 4x4 block (9.2.1). */
static int luma_nc(const oracle_h264_decoder *d, int mb, int blk) {
    const int bx = (blk / 4 % 2) * 2 + blk % 2, by = (blk / 8) * 2 + (blk / 2) % 2;
    Nb A = nb4(d, mb, bx - 1, by, NULL), B = nb4(d, mb, bx, by - 1, NULL);
    const int na = A.avail ? d->mb[A.mb].tc[A.blk] : 0, nbv = B.avail ? d->mb[B.mb].tc[B.blk] : 0;
    if (A.avail && B.avail) return (na + nbv + 1) >> 1;
    return A.avail ? na : B.avail ? nbv : 0;
}

/* nC of chroma AC block `blk` (2x2 grid) of component c. */
static int chroma_nc(const oracle_h264_decoder *d, int mb, int c, int blk) {
    const int bx = blk % 2, by = blk / 2;
    int na = 0, nbv = 0, ha = 0, hb = 0;
    if (bx > 0) { na = d->mb[mb].tc_c[c][blk - 1]; ha = 1; }
    else {
        Nb A = nb4(d, mb, -1, 0, NULL);
        if (A.avail) { na = d->mb[A.mb].tc_c[c][by * 2 + 1]; ha = 1; }
    }
    if (by > 0) { nbv = d->mb[mb].tc_c[c][blk - 2]; hb = 1; }
    else {
        Nb B = nb4(d, mb, 0, -1, NULL);
        if (B.avail) { nbv = d->mb[B.mb].tc_c[c][2 + bx]; hb = 1; }
    }
    if (ha && hb) return (na + nbv + 1) >> 1;
    return ha ? na : hb ? nbv : 0;
}

/* Intra4x4 / Intra8x8 mode predictor of the 4x4 block blk (8.3.1.1; neighbours of other
 * macroblock types count as DC, an unavailable neighbour makes the predictor DC). */
static int intra_mode_pred(const oracle_h264_decoder *d, int mb, int blk) {
    const int bx = (blk / 4 % 2) * 2 + blk % 2, by = (blk / 8) * 2 + (blk / 2) % 2;
    Nb A = nb4(d, mb, bx - 1, by, NULL), B = nb4(d, mb, bx, by - 1, NULL);
    if (!A.avail || !B.avail) return 2;
    const int ma = d->mb[A.mb].imode[A.blk], mbv = d->mb[B.mb].imode[B.blk];
    return ma < mbv ? ma : mbv;
}

/* ------------------------------------------------------------------ macroblock layer (7.3.5) */
typedef struct {
    const SliceHeader *h;
    int slice_id;
    int qp;                 /* QP_Y,PRED for the next macroblock */
    Ref *lists[2][33];
    int counts[2];
} SliceCtx;

static int16_t *level_record(oracle_h264_decoder *d, int mb) {
    if (d->coded == d->levels_cap) {
        d->levels_cap = d->levels_cap ? d->levels_cap * 2 : 64;
        d->levels = (int16_t *)realloc(d->levels, (size_t)d->levels_cap * MMR_H264_LEVELS_PER_MB * sizeof(int16_t));
    }
    int16_t *r = d->levels + (size_t)d->coded * MMR_H264_LEVELS_PER_MB;
    memset(r, 0, MMR_H264_LEVELS_PER_MB * sizeof(int16_t));
    d->rec[mb].coef_offset = d->coded++;
    return r;
}

static int clamp_level(int32_t v, int16_t *out) {
    if (v < -32768 || v > 32767) return -1;
    *out = (int16_t)v;
    return 0;
}

/* Residual of a non-PCM macroblock (7.3.5.3, CAVLC) into its level record and nonzero counts. */
static int residual(oracle_h264_decoder *d, Bits *b, int mb, int i16, int cbp_luma, int cbp_chroma, int16_t *rec) {
    MbState *m = &d->mb[mb];
    int32_t c[16];
    if (i16) {
        const int t = residual_block(b, luma_nc(d, mb, 0), 16, c);
        if (t < 0) return -1;
        for (int k = 0; k < 16; ++k)
            if (clamp_level(c[k], &rec[16 * k])) return -1;
    }
    for (int b8 = 0; b8 < 4; ++b8)
        for (int i4 = 0; i4 < 4; ++i4) {
            const int blk = b8 * 4 + i4;
            if (!(cbp_luma & (1 << b8))) {
                m->tc[blk] = 0;
                continue;
            }
            int t;
            if (i16) {
                t = residual_block(b, luma_nc(d, mb, blk), 15, c);
                if (t < 0) return -1;
                for (int k = 0; k < 15; ++k)
                    if (clamp_level(c[k], &rec[16 * blk + 1 + k])) return -1;
            } else if (m->transform8) {
                t = residual_block(b, luma_nc(d, mb, blk), 16, c);
                if (t < 0) return -1;
                for (int k = 0; k < 16; ++k)
                    if (clamp_level(c[k], &rec[64 * b8 + 4 * k + i4])) return -1;
            } else {
                t = residual_block(b, luma_nc(d, mb, blk), 16, c);
                if (t < 0) return -1;
                for (int k = 0; k < 16; ++k)
                    if (clamp_level(c[k], &rec[16 * blk + k])) return -1;
            }
            m->tc[blk] = (uint8_t)t;
        }
    for (int comp = 0; comp < 2; ++comp)
        if (cbp_chroma & 3) {
            if (residual_block(b, -1, 4, c) < 0) return -1;
            for (int k = 0; k < 4; ++k)
                if (clamp_level(c[k], &rec[256 + 4 * comp + k])) return -1;
        }
    for (int comp = 0; comp < 2; ++comp)
        for (int blk = 0; blk < 4; ++blk) {
            if (!(cbp_chroma & 2)) {
                m->tc_c[comp][blk] = 0;
                continue;
            }
            const int t = residual_block(b, chroma_nc(d, mb, comp, blk), 15, c);
            if (t < 0) return -1;
            for (int k = 0; k < 15; ++k)
                if (clamp_level(c[k], &rec[264 + 15 * (4 * comp + blk) + k])) return -1;
            m->tc_c[comp][blk] = (uint8_t)t;
        }
    return 0;
}

/* Motion of an inter partition: refIdx, mvd and mv into the 4x4 blocks it covers. */
static void set_partition(MbState *m, int list, int x, int y, int w, int h, int ref, const int mv[2], const int mvd[2], uint8_t *cur) {
    for (int yy = y; yy < y + h; ++yy)
        for (int xx = x; xx < x + w; ++xx) {
            const int k = blk_idx(xx, yy);
            m->ref[list][k] = (int8_t)ref;
            m->mv[list][k][0] = (int16_t)mv[0];
            m->mv[list][k][1] = (int16_t)mv[1];
            m->mvd[list][k][0] = (int16_t)mvd[0];
            m->mvd[list][k][1] = (int16_t)mvd[1];
            if (cur) cur[k] = 1;
        }
}

/* The coded part of an inter macroblock after mb_type: partitions (shape 0 16x16, 1 16x8, 2 8x16)
 * with their predictions, or sub-macroblocks (sub != 0). */
static int inter_mb(oracle_h264_decoder *d, Bits *b, SliceCtx *sc, int mb, int p_slice, int shape, const int pred[2], int sub,
                    int ref0_only) {
    MbState *m = &d->mb[mb];
    const SliceHeader *h = sc->h;
    for (int l = 0; l < 2; ++l)
        for (int k = 0; k < 16; ++k) {
            m->ref[l][k] = -1;
            m->mv[l][k][0] = m->mv[l][k][1] = 0;
            m->mvd[l][k][0] = m->mvd[l][k][1] = 0;
        }
    uint8_t cur[16];
    memset(cur, 0, sizeof cur);
    if (!sub) {
        const int parts = shape == 0 ? 1 : 2;
        int ref[2][2] = {{0, 0}, {0, 0}}, mvd[2][2][2];
        for (int l = 0; l < 2; ++l)
            for (int i = 0; i < parts; ++i)
                if (pred[i] == l || pred[i] == PRED_BI) {
                    if (h->num_ref_idx_active[l] > 1) ref[l][i] = (int)rd_te(b, (uint32_t)h->num_ref_idx_active[l] - 1);
                    if (ref[l][i] >= h->num_ref_idx_active[l]) return -1;
                }
        for (int l = 0; l < 2; ++l)
            for (int i = 0; i < parts; ++i)
                if (pred[i] == l || pred[i] == PRED_BI) {
                    mvd[l][i][0] = rd_se(b);
                    mvd[l][i][1] = rd_se(b);
                }
        /* predictions in partition order, each list in turn (the neighbours of partition 1
         * include partition 0) */
        for (int i = 0; i < parts; ++i) {
            const int x = shape == 2 ? 2 * i : 0, y = shape == 1 ? 2 * i : 0, w = shape == 2 ? 2 : 4, hh = shape == 1 ? 2 : 4;
            for (int l = 0; l < 2; ++l) {
                if (!(pred[i] == l || pred[i] == PRED_BI)) continue;
                int mvp[2];
                predict_mv(d, mb, l, ref[l][i], x, y, w, shape, cur, mvp);
                const int mv[2] = {mvp[0] + mvd[l][i][0], mvp[1] + mvd[l][i][1]};
                set_partition(m, l, x, y, w, hh, ref[l][i], mv, mvd[l][i], NULL);
            }
            for (int yy = y; yy < y + hh; ++yy)
                for (int xx = x; xx < x + w; ++xx) cur[blk_idx(xx, yy)] = 1;
        }
        return b->err ? -1 : 0;
    }
    int st[4], sshape[4], spred[4];
    for (int i = 0; i < 4; ++i) {
        st[i] = (int)rd_ue(b);
        if (p_slice) {
            if (st[i] > 3) return -1;
            sshape[i] = PSUB_SHAPE[st[i]];
            spred[i] = PRED_L0;
        } else {
            if (st[i] > 12) return -1;
            if (st[i] == 0) return -2;   /* B_Direct_8x8 */
            sshape[i] = BSUB_SHAPE[st[i]];
            spred[i] = BSUB_PRED[st[i]];
        }
    }
    int ref[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
    for (int l = 0; l < 2; ++l)
        for (int i = 0; i < 4; ++i)
            if (spred[i] == l || spred[i] == PRED_BI) {
                if (h->num_ref_idx_active[l] > 1 && !ref0_only) ref[l][i] = (int)rd_te(b, (uint32_t)h->num_ref_idx_active[l] - 1);
                if (ref[l][i] >= h->num_ref_idx_active[l]) return -1;
            }
    int mvd[2][4][4][2];
    for (int l = 0; l < 2; ++l)
        for (int i = 0; i < 4; ++i) {
            if (!(spred[i] == l || spred[i] == PRED_BI)) continue;
            const int n = sshape[i] == 0 ? 1 : sshape[i] == 3 ? 4 : 2;
            for (int j = 0; j < n; ++j) {
                mvd[l][i][j][0] = rd_se(b);
                mvd[l][i][j][1] = rd_se(b);
            }
        }
    for (int i = 0; i < 4; ++i) {
        const int x8 = (i % 2) * 2, y8 = (i / 2) * 2;
        const int n = sshape[i] == 0 ? 1 : sshape[i] == 3 ? 4 : 2;
        const int w = (sshape[i] == 0 || sshape[i] == 1) ? 2 : 1, hh = (sshape[i] == 0 || sshape[i] == 2) ? 2 : 1;
        for (int j = 0; j < n; ++j) {
            const int x = x8 + (sshape[i] == 2 || sshape[i] == 3 ? j % 2 : 0);
            const int y = y8 + (sshape[i] == 1 ? j : sshape[i] == 3 ? j / 2 : 0);
            for (int l = 0; l < 2; ++l) {
                if (!(spred[i] == l || spred[i] == PRED_BI)) continue;
                int mvp[2];
                predict_mv(d, mb, l, ref[l][i], x, y, w, 0, cur, mvp);
                const int mv[2] = {mvp[0] + mvd[l][i][j][0], mvp[1] + mvd[l][i][j][1]};
                set_partition(m, l, x, y, w, hh, ref[l][i], mv, mvd[l][i][j], NULL);
            }
            for (int yy = y; yy < y + hh; ++yy)
                for (int xx = x; xx < x + w; ++xx) cur[blk_idx(xx, yy)] = 1;
        }
    }
    /* transform_size_8x8_flag needs every sub-partition at least 8x8 */
    int all8 = 1;
    for (int i = 0; i < 4; ++i) all8 &= sshape[i] == 0;
    return b->err ? -1 : (all8 ? 1 : 0);
}

static void write_records(oracle_h264_decoder *d, SliceCtx *sc, int mb) {
    const MbState *m = &d->mb[mb];
    mmr_recon_h264_macroblock *r = &d->rec[mb];
    const SliceHeader *h = sc->h;
    r->flags = (uint8_t)((is_intra_kind(m->kind) ? MMR_H264_MB_INTRA : 0) | (m->transform8 ? MMR_H264_MB_TRANSFORM_8X8 : 0) |
                         (h->disable_deblocking == 1 ? MMR_H264_MB_NO_DEBLOCK : 0) |
                         (h->disable_deblocking == 2 ? MMR_H264_MB_SLICE_EDGE_OFF : 0));
    r->qp = (uint8_t)m->qp;
    r->cbp = m->cbp;
    r->slice_id = (uint16_t)sc->slice_id;
    r->offset_a = (int8_t)h->alpha_off;
    r->offset_b = (int8_t)h->beta_off;
    for (int k = 0; k < 16; ++k) {
        mmr_recon_h264_block *bk = &d->blk[(size_t)mb * 16 + k];
        for (int l = 0; l < 2; ++l) {
            bk->ref_idx[l] = is_intra_kind(m->kind) ? -1 : m->ref[l][k];
            bk->mv[l][0] = bk->ref_idx[l] >= 0 ? m->mv[l][k][0] : 0;
            bk->mv[l][1] = bk->ref_idx[l] >= 0 ? m->mv[l][k][1] : 0;
        }
        bk->total_coeff = m->tc[k];
    }
}

/* One coded macroblock (after mb_skip_run). Returns 0, -1 invalid, -2 unsupported. */
static int macroblock(oracle_h264_decoder *d, Bits *b, SliceCtx *sc, int mb) {
    MbState *m = &d->mb[mb];
    const SliceHeader *h = sc->h;
    const Pps *p = d->p;
    memset(m->tc, 0, sizeof m->tc);
    memset(m->tc_c, 0, sizeof m->tc_c);
    memset(m->imode, 2, sizeof m->imode);
    m->transform8 = 0;
    int t = (int)rd_ue(b);
    if (b->err) return -1;
    int intra_type = -1;   /* I slice mb_type numbering */
    if (h->type == 2) intra_type = t;
    else if (h->type == 0 && t >= 5) intra_type = t - 5;
    else if (h->type == 1 && t >= 23) intra_type = t - 23;
    if (intra_type > 25) return -1;
    mmr_recon_h264_intra *ir = &d->intra[mb];
    memset(ir, 0, sizeof *ir);
    if (intra_type == 25) {   /* I_PCM */
        m->kind = MBK_PCM;
        while (b->pos & 7) rd_bit(b);   /* pcm_alignment_zero_bit */
        int16_t *rec = level_record(d, mb);
        for (int i = 0; i < 384; ++i) rec[i] = (int16_t)rd_u(b, 8);
        memset(m->tc, 16, sizeof m->tc);
        memset(m->tc_c, 16, sizeof m->tc_c);
        m->cbp = 0;
        m->qp = 0;              /* deblocking takes qP 0 for I_PCM; QP_Y,PRED stays */
        ir->type = MMR_H264_INTRA_PCM;
        for (int l = 0; l < 2; ++l) memset(m->ref[l], -1, sizeof m->ref[l]);
        write_records(d, sc, mb);
        return b->err ? -1 : 0;
    }
    int cbp_luma = 0, cbp_chroma = 0, i16 = 0, allow8 = 0;
    if (intra_type >= 0) {
        for (int l = 0; l < 2; ++l) memset(m->ref[l], -1, sizeof m->ref[l]);
        if (intra_type == 0) {   /* I_NxN */
            m->transform8 = p->transform_8x8_mode ? (uint8_t)rd_bit(b) : 0;
            m->kind = m->transform8 ? MBK_I8 : MBK_I4;
            ir->type = m->transform8 ? MMR_H264_INTRA_8X8 : MMR_H264_INTRA_4X4;
            for (int i = 0; i < (m->transform8 ? 4 : 16); ++i) {
                const int blk = m->transform8 ? 4 * i : i;
                const int pred = intra_mode_pred(d, mb, blk);
                int mode = pred;
                if (!rd_bit(b)) {
                    const int rem = (int)rd_u(b, 3);
                    mode = rem < pred ? rem : rem + 1;
                }
                if (mode > 8) return -1;
                if (m->transform8) {
                    for (int k = 0; k < 4; ++k) m->imode[blk + k] = (uint8_t)mode;
                    ir->modes[i] = (uint8_t)mode;
                } else {
                    m->imode[blk] = (uint8_t)mode;
                    ir->modes[blk] = (uint8_t)mode;
                }
            }
        } else {                 /* Intra16x16 */
            m->kind = MBK_I16;
            i16 = 1;
            ir->type = MMR_H264_INTRA_16X16;
            ir->luma_mode = (uint8_t)((intra_type - 1) % 4);
            cbp_chroma = ((intra_type - 1) / 4) % 3;
            cbp_luma = intra_type >= 13 ? 15 : 0;
        }
        const int chroma_mode = (int)rd_ue(b);
        if (chroma_mode > 3) return -1;
        ir->chroma_mode = (uint8_t)chroma_mode;
        if (!i16) {
            const uint32_t code = rd_ue(b);
            if (code > 47) return -1;
            cbp_luma = CBP_INTRA[code] & 15;
            cbp_chroma = CBP_INTRA[code] >> 4;
        }
    } else {
        m->kind = MBK_INTER;
        int r;
        if (h->type == 0) {
            if (t <= 2) {
                const int pred[2] = {PRED_L0, PRED_L0};
                r = inter_mb(d, b, sc, mb, 1, t, pred, 0, 0);
                allow8 = 1;
            } else {
                r = inter_mb(d, b, sc, mb, 1, 0, NULL, 1, t == 4);
                allow8 = r == 1;
            }
        } else {
            if (t == 0) return -2;   /* B_Direct_16x16 */
            if (t <= 21) {
                const int pred[2] = {B_PRED[t][0], B_PRED[t][1]};
                r = inter_mb(d, b, sc, mb, 0, B_SHAPE[t], pred, 0, 0);
                allow8 = 1;
            } else {
                r = inter_mb(d, b, sc, mb, 0, 0, NULL, 1, 0);
                allow8 = r == 1;
            }
        }
        if (r < 0) return r;
        const uint32_t code = rd_ue(b);
        if (code > 47) return -1;
        cbp_luma = CBP_INTER[code] & 15;
        cbp_chroma = CBP_INTER[code] >> 4;
        if (cbp_luma && p->transform_8x8_mode && allow8) m->transform8 = (uint8_t)rd_bit(b);
    }
    m->cbp = (uint8_t)(cbp_luma | (cbp_chroma << 4));
    if (cbp_luma || cbp_chroma || i16) {
        const int delta = rd_se(b);
        if (delta < -26 || delta > 25) return -1;
        sc->qp = (sc->qp + delta + 52) % 52;
    }
    m->qp = sc->qp;
    if (m->cbp || i16) {
        int16_t *rec = level_record(d, mb);
        if (residual(d, b, mb, i16, cbp_luma, cbp_chroma, rec)) return -1;
    } else {
        d->rec[mb].coef_offset = 0;
    }
    write_records(d, sc, mb);
    return b->err ? -1 : 0;
}

/* P_Skip (8.4.1.1): reference 0, motion predicted unless a neighbour says zero. */
static void p_skip(oracle_h264_decoder *d, SliceCtx *sc, int mb) {
    MbState *m = &d->mb[mb];
    m->kind = MBK_P_SKIP;
    m->transform8 = 0;
    m->cbp = 0;
    m->qp = sc->qp;
    memset(m->tc, 0, sizeof m->tc);
    memset(m->tc_c, 0, sizeof m->tc_c);
    memset(m->imode, 2, sizeof m->imode);
    for (int l = 0; l < 2; ++l)
        for (int k = 0; k < 16; ++k) {
            m->ref[l][k] = -1;
            m->mv[l][k][0] = m->mv[l][k][1] = 0;
            m->mvd[l][k][0] = m->mvd[l][k][1] = 0;
        }
    Nb A = nb4(d, mb, -1, 0, NULL), B = nb4(d, mb, 0, -1, NULL);
    int mv[2] = {0, 0};
    if (A.avail && B.avail) {
        int ra, rb, ma[2], mbv[2];
        nb_motion(d, A, 0, &ra, ma);
        nb_motion(d, B, 0, &rb, mbv);
        if (!((ra == 0 && ma[0] == 0 && ma[1] == 0) || (rb == 0 && mbv[0] == 0 && mbv[1] == 0))) predict_mv(d, mb, 0, 0, 0, 0, 4, 0, NULL, mv);
    }
    const int zero[2] = {0, 0};
    set_partition(m, 0, 0, 0, 4, 4, 0, mv, zero, NULL);
    d->intra[mb].type = 0;
    d->rec[mb].coef_offset = 0;
    write_records(d, sc, mb);
}

/* ------------------------------------------------------------------ slices and pictures */
static int start_picture(oracle_h264_decoder *d, const SliceHeader *h) {
    const Pps *p = &d->pps[h->pps_id];
    const Sps *s = &d->sps[p->sps_id];
    free_picture(d);
    d->s = s;
    d->p = p;
    d->width_mbs = s->width_mbs;
    d->height_mbs = s->height_mbs;
    d->mbs = s->width_mbs * s->height_mbs;
    d->mb = (MbState *)calloc((size_t)d->mbs, sizeof(MbState));
    d->rec = (mmr_recon_h264_macroblock *)calloc((size_t)d->mbs, sizeof(mmr_recon_h264_macroblock));
    d->blk = (mmr_recon_h264_block *)calloc((size_t)d->mbs * 16, sizeof(mmr_recon_h264_block));
    d->intra = (mmr_recon_h264_intra *)calloc((size_t)d->mbs, sizeof(mmr_recon_h264_intra));
    return d->mb && d->rec && d->blk && d->intra ? 0 : -1;
}

/* Weight tables of the picture from its (first) slice header: explicit tables, implicit
 * bi-prediction weights from the pictures' order counts (8.4.2.3), or none. */
static void weights_of(oracle_h264_decoder *d, const SliceHeader *h, SliceCtx *sc, int poc,
                       mmr_recon_h264_weight (*weights)[MMR_H264_MAX_REFS][3],
                       mmr_recon_h264_bi_weight (*bi)[MMR_H264_MAX_REFS][3], uint32_t *mode) {
    memset(weights, 0, sizeof d->weights);
    memset(bi, 0, sizeof d->bi);
    *mode = MMR_H264_WEIGHTED_NONE;
    const int explicit_w = (d->p->weighted_pred && h->type == 0) || (d->p->weighted_bipred_idc == 1 && h->type == 1);
    const int implicit_w = d->p->weighted_bipred_idc == 2 && h->type == 1;
    if (!explicit_w && !implicit_w) return;
    *mode = MMR_H264_WEIGHTED_TABLE;
    for (int l = 0; l < 2; ++l)
        for (int i = 0; i < sc->counts[l]; ++i)
            for (int c = 0; c < 3; ++c) {
                mmr_recon_h264_weight *w = &weights[l][i][c];
                if (explicit_w) {
                    w->log2_denom = (uint8_t)(c ? h->chroma_log2_denom : h->luma_log2_denom);
                    w->weight = (int16_t)(c ? h->cw[l][i][c - 1] : h->lw[l][i]);
                    w->offset = (int16_t)(c ? h->co[l][i][c - 1] : h->lo[l][i]);
                } else {
                    w->log2_denom = 0;   /* single-list prediction stays unweighted */
                    w->weight = 1;
                    w->offset = 0;
                }
            }
    if (h->type != 1) return;
    for (int i = 0; i < sc->counts[0]; ++i)
        for (int j = 0; j < sc->counts[1]; ++j)
            for (int c = 0; c < 3; ++c) {
                mmr_recon_h264_bi_weight *w = &bi[i][j][c];
                if (explicit_w) {
                    w->log2_denom = (uint8_t)(c ? h->chroma_log2_denom : h->luma_log2_denom);
                    w->w0 = (int16_t)(c ? h->cw[0][i][c - 1] : h->lw[0][i]);
                    w->w1 = (int16_t)(c ? h->cw[1][j][c - 1] : h->lw[1][j]);
                    w->o0 = (int16_t)(c ? h->co[0][i][c - 1] : h->lo[0][i]);
                    w->o1 = (int16_t)(c ? h->co[1][j][c - 1] : h->lo[1][j]);
                } else {
                    int32_t w0, w1;
                    oracle_h264_implicit_weights(sc->lists[0][i]->poc, sc->lists[1][j]->poc, poc, sc->lists[0][i]->long_term,
                                                 sc->lists[1][j]->long_term, &w0, &w1);
                    w->log2_denom = 5;
                    w->w0 = (int16_t)w0;
                    w->w1 = (int16_t)w1;
                }
            }
}

/* The picture's reference lists and weight tables are those of its slices: a slice may use a
 * longer list than an earlier one, but where two slices name index i of a list they must name
 * the same picture with the same weights (the records carry one list and table per picture). */
static int merge_slice(oracle_h264_decoder *d, const SliceHeader *h, SliceCtx *sc, int poc, int first) {
    mmr_recon_h264_weight w[2][MMR_H264_MAX_REFS][3];
    mmr_recon_h264_bi_weight bi[MMR_H264_MAX_REFS][MMR_H264_MAX_REFS][3];
    uint32_t mode;
    weights_of(d, h, sc, poc, w, bi, &mode);
    oracle_h264_picture_info *info = &d->info;
    if (first) {
        memset(info, 0, sizeof *info);
        memset(d->weights, 0, sizeof d->weights);
        memset(d->bi, 0, sizeof d->bi);
        info->weighted_mode = mode;
    } else if (mode != info->weighted_mode) {
        return -2;
    }
    for (int l = 0; l < 2; ++l)
        for (int i = 0; i < sc->counts[l]; ++i) {
            if ((uint32_t)i < info->ref_count[l]) {
                if (info->ref_slot[l][i] != sc->lists[l][i]->slot || memcmp(d->weights[l][i], w[l][i], sizeof w[l][i])) return -2;
                continue;
            }
            info->ref_slot[l][i] = sc->lists[l][i]->slot;
            info->ref_poc[l][i] = sc->lists[l][i]->poc;
            memcpy(d->weights[l][i], w[l][i], sizeof w[l][i]);
            info->ref_count[l] = (uint32_t)i + 1;
        }
    for (int i = 0; i < sc->counts[0]; ++i)
        for (int j = 0; j < sc->counts[1]; ++j) {
            mmr_recon_h264_bi_weight *have = d->bi[i][j];
            int empty = 1;
            for (size_t k = 0; k < sizeof d->bi[i][j]; ++k) empty &= ((const uint8_t *)have)[k] == 0;
            if (empty) memcpy(have, bi[i][j], sizeof bi[i][j]);
            else if (memcmp(have, bi[i][j], sizeof bi[i][j])) return -2;
        }
    return 0;
}

/* slice_data (7.3.4, CAVLC). */
static int slice_data(oracle_h264_decoder *d, Bits *b, SliceCtx *sc) {
    const SliceHeader *h = sc->h;
    int mb = h->first_mb;
    if (mb >= d->mbs) return -1;
    for (;;) {
        if (h->type != 2) {
            uint32_t run = rd_ue(b);
            if (b->err) return -1;
            for (; run > 0; --run) {
                if (mb >= d->mbs) return -1;
                if (h->type == 1) return -2;   /* B_Skip: direct prediction */
                d->mb[mb].slice = sc->slice_id;
                p_skip(d, sc, mb);
                ++mb;
            }
            if (!more_rbsp_data(b)) break;
        }
        if (mb >= d->mbs) return -1;
        d->mb[mb].slice = sc->slice_id;
        const int r = macroblock(d, b, sc, mb);
        if (r) return r;
        ++mb;
        if (!more_rbsp_data(b)) break;
    }
    return b->err ? -1 : 0;
}

int oracle_h264_decode_access_unit(oracle_h264_decoder *d, const uint8_t *data, size_t size, oracle_h264_picture_info *info) {
    int slices = 0, status = 0, poc = 0, frame_num_offset = 0;
    SliceHeader first;
    SliceHeader h;
    d->have_picture = 0;
    for (size_t at = 0; at + 4 <= size && status == 0;) {
        const size_t n = ((size_t)data[at] << 24) | ((size_t)data[at + 1] << 16) | ((size_t)data[at + 2] << 8) | data[at + 3];
        at += 4;
        if (n == 0 || at + n > size) return -1;
        const uint8_t *nal = data + at;
        at += n;
        const int nal_ref_idc = (nal[0] >> 5) & 3, type = nal[0] & 31;
        Bits b;
        uint8_t *buf = unescape(nal, n, &b);
        if (type == 7) status = parse_sps(&b, d->sps);
        else if (type == 8) status = parse_pps(&b, d->pps, d->sps);
        else if (type == 1 || type == 5) {
            status = parse_slice_header(d, &b, type, nal_ref_idc, &h);
            if (status == 0 && slices == 0) {
                first = h;
                status = start_picture(d, &h);
                if (status == 0) poc = picture_order_count(d, &h, &frame_num_offset);
            } else if (status == 0 && (h.frame_num != first.frame_num || h.pps_id != first.pps_id || h.idr != first.idr)) {
                status = -1;   /* slices of one access unit belong to one picture */
            }
            if (status == 0) {
                SliceCtx sc;
                memset(&sc, 0, sizeof sc);
                sc.h = &h;
                sc.slice_id = slices;
                sc.qp = h.qp;
                status = build_lists(d, &h, poc, sc.lists, sc.counts);
                if (status == 0) status = merge_slice(d, &h, &sc, poc, slices == 0);
                if (status == 0) status = slice_data(d, &b, &sc);
                ++slices;
            }
        }
        free(buf);
    }
    if (status) return status;
    if (!slices) return 0;
    for (int i = 0; i < d->mbs; ++i)
        if (d->mb[i].kind == MBK_NONE) return -1;   /* the slices do not cover the picture */
    oracle_h264_picture_info *o = &d->info;
    o->width_mbs = (uint32_t)d->width_mbs;
    o->height_mbs = (uint32_t)d->height_mbs;
    o->slot = d->next_slot++;
    o->poc = poc;
    o->is_reference = first.nal_ref_idc != 0;
    o->coded_mbs = d->coded;
    o->deblock = 1;
    o->chroma_qp_offset[0] = d->p->chroma_qp_offset[0];
    o->chroma_qp_offset[1] = d->p->chroma_qp_offset[1];
    o->transform_bypass = (uint32_t)d->s->transform_bypass;
    o->slice_type = (uint32_t)first.type;
    mark_current(d, &first, poc, o->slot);
    d->prev_frame_num = first.frame_num;
    d->prev_frame_num_offset = frame_num_offset;
    d->have_picture = 1;
    if (info) *info = *o;
    return 1;
}

int oracle_h264_picture_records(const oracle_h264_decoder *d, mmr_recon_h264_macroblock *mbs, mmr_recon_h264_block *blocks,
                                int16_t *levels, mmr_recon_h264_intra *intra, mmr_recon_h264_weight *weights,
                                mmr_recon_h264_bi_weight *bi_weights, int16_t *mvd) {
    if (!d->have_picture) return -1;
    if (mbs) memcpy(mbs, d->rec, (size_t)d->mbs * sizeof *mbs);
    if (blocks) memcpy(blocks, d->blk, (size_t)d->mbs * 16 * sizeof *blocks);
    if (levels && d->coded) memcpy(levels, d->levels, (size_t)d->coded * MMR_H264_LEVELS_PER_MB * sizeof *levels);
    if (intra) memcpy(intra, d->intra, (size_t)d->mbs * sizeof *intra);
    if (weights) memcpy(weights, d->weights, sizeof d->weights);
    if (bi_weights) memcpy(bi_weights, d->bi, sizeof d->bi);
    if (mvd)
        for (int i = 0; i < d->mbs; ++i)
            for (int l = 0; l < 2; ++l)
                for (int k = 0; k < 16; ++k)
                    for (int c = 0; c < 2; ++c) mvd[(((size_t)i * 2 + l) * 16 + k) * 2 + c] = d->mb[i].mvd[l][k][c];
    return 0;
}

/* One residual block from a byte string (tests): TotalCoeff, or -1. */
int oracle_h264_cavlc_block(const uint8_t *bytes, size_t n, int nc, int max_coeff, int32_t *coef, size_t *bits_used) {
    Bits b = {bytes, n * 8, 0, 0};
    const int t = residual_block(&b, nc, max_coeff, coef);
    if (bits_used) *bits_used = b.pos;
    return t;
}
