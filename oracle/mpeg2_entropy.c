/*
 * CPU MPEG-2 entropy front end -- TEST INFRASTRUCTURE.
 *
 * Stands in for the reference's CPU VLC side so that real MPEG-2 elementary streams can be fed to
 * the reconstruction path: it parses the stream and decodes every frame picture's slices into
 * the post-entropy buffers of mmr_recon_mpeg2_picture -- one macroblock record per macroblock
 * (skipped macroblocks resolved to explicit records, motion vectors after prediction) and the
 * levels of every coded block by scan position (intra slot 0 = the DC predictor value). Only
 * tests/ use it; the product library never links it.
 *
 * Behaviour follows the reference:
 *   start codes / units / picture assembly  crates/codecs/mpeg2/src/parser.rs:135-390
 *   sequence header + extension, matrices   parser.rs:392-540, syntax.rs:229-306 (loaded
 *                                           matrices arrive in zig-zag order; chroma matrices
 *                                           default to the luma ones)
 *   picture header / coding extension       parser.rs:567-665
 *   slice header                            parser.rs:667-712, decoder.rs:349-371
 *   macroblock loop                         decoder.rs:340-606 (types, quantiser scale, intra /
 *                                           non-intra blocks, motion types, predictor resets)
 *   skipped macroblocks                     decoder.rs:608-679 (P: zero forward vector, B: the
 *                                           previous macroblock's prediction)
 *   blocks / run-level / motion components  decoder.rs:681-942
 *   reference selection                     decoder.rs:253-307, :953-957 (P: newer anchor, B:
 *                                           older and newer anchors)
 * Field-predicted macroblocks carry MMR_MPEG2_MB_FIELD_MOTION, the parity-0 vectors in the record
 * and the parity-1 vectors and field selects in the picture's field-motion array.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mpeg2_vlc.h"
#include "oracle.h"

static const uint8_t kZz[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,  12, 19, 26, 33, 40, 48,
                                41, 34, 27, 20, 13, 6,  7,  14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23,
                                30, 37, 44, 51, 58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

enum { MB_INTRA = 1, MB_PATTERN = 2, MB_FORWARD = 4, MB_BACKWARD = 8, MB_QUANT = 16 };
enum { SYM_ESCAPE = 111, SYM_EOB = 112 };

typedef struct {
    const uint8_t *p;
    size_t nbits, pos;
} bits_t;

typedef struct {
    char *err;
    size_t cap;
} errctx;

static int fail(errctx *e, const char *msg) {
    if (e->err && e->cap) snprintf(e->err, e->cap, "%s", msg);
    return -1;
}

static size_t bits_left(const bits_t *b) { return b->nbits - b->pos; }

static int get_bit(bits_t *b, int *v, errctx *e) {
    if (b->pos >= b->nbits) return fail(e, "unexpected end of slice data");
    *v = (b->p[b->pos >> 3] >> (7 - (b->pos & 7))) & 1;
    ++b->pos;
    return 0;
}

static int get_bits(bits_t *b, int n, uint32_t *v, errctx *e) {
    uint32_t x = 0;
    for (int i = 0; i < n; ++i) {
        int bit;
        if (get_bit(b, &bit, e)) return -1;
        x = (x << 1) | (uint32_t)bit;
    }
    *v = x;
    return 0;
}

static int remaining_zero(const bits_t *b) {
    for (size_t i = b->pos; i < b->nbits; ++i)
        if ((b->p[i >> 3] >> (7 - (i & 7))) & 1) return 0;
    return 1;
}

/* Bit by bit, the first entry whose length and code match (the reference's VlcTable::decode). */
static int vlc(bits_t *b, const m2v_vlc *t, int n, int *sym, errctx *e) {
    uint32_t code = 0;
    for (int len = 1; len <= 16; ++len) {
        int bit;
        if (get_bit(b, &bit, e)) return -1;
        code = (code << 1) | (uint32_t)bit;
        for (int s = 0; s < n; ++s)
            if (t[s].length == len && t[s].code == code) {
                *sym = s;
                return 0;
            }
    }
    return fail(e, "no VLC table entry matches the input bits");
}

static int vlc_sym(bits_t *b, const m2v_vlc_sym *t, int n, int *sym, errctx *e) {
    uint32_t code = 0;
    for (int len = 1; len <= 16; ++len) {
        int bit;
        if (get_bit(b, &bit, e)) return -1;
        code = (code << 1) | (uint32_t)bit;
        for (int s = 0; s < n; ++s)
            if (t[s].length == len && t[s].code == code) {
                *sym = t[s].symbol;
                return 0;
            }
    }
    return fail(e, "no VLC table entry matches the input bits");
}

/* ------------------------------------------------------------------ stream state */

typedef struct {
    int have_seq, have_ext;
    uint32_t width, height, chroma_format;
    int seq_intra_loaded, seq_non_intra_loaded;
    uint8_t seq_intra[64], seq_non_intra[64];      /* natural order */
    int qm_loaded[4];                              /* merged quant-matrix extensions */
    uint8_t qm[4][64];
} seq_state;

typedef struct {
    uint32_t width, height, picture_type;
    uint8_t f_code[2][2];
    uint32_t intra_dc_precision, picture_structure, frame_pred_frame_dct, q_scale_type, intra_vlc_format, alternate_scan;
    uint32_t progressive_frame, display_index;
    int have_coding_ext;
    uint8_t matrices[4][64];   /* intra, non-intra, chroma intra, chroma non-intra (natural) */
    size_t nslices, cap_slices;
    struct {
        uint32_t vertical_position;
        size_t begin, end;
    } * slices;
    /* outputs */
    int32_t forward_index, backward_index;
    uint32_t mb_count, coded_blocks, field_motion_mbs;
    mmr_recon_mpeg2_macroblock *mbs;
    mmr_recon_mpeg2_field_motion *field;
    int16_t *levels;
    size_t levels_cap;
} m2v_picture;

struct oracle_mpeg2_stream {
    uint32_t count, cap;
    m2v_picture *pics;
};

static void resolve_matrices(const seq_state *s, uint8_t m[4][64]) {
    for (int i = 0; i < 64; ++i) {
        m[0][i] = s->seq_intra_loaded ? s->seq_intra[i] : m2v_default_intra[i];
        m[1][i] = s->seq_non_intra_loaded ? s->seq_non_intra[i] : 16;
    }
    memcpy(m[2], m[0], 64);
    memcpy(m[3], m[1], 64);
    for (int k = 0; k < 4; ++k)
        if (s->qm_loaded[k]) memcpy(m[k], s->qm[k], 64);
}

/* A transmitted matrix: 64 bytes in zig-zag order -> natural order. */
static int read_matrix(bits_t *b, uint8_t out[64], errctx *e) {
    for (int k = 0; k < 64; ++k) {
        uint32_t v;
        if (get_bits(b, 8, &v, e)) return -1;
        out[kZz[k]] = (uint8_t)v;
    }
    return 0;
}

/* ------------------------------------------------------------------ picture decoding */

static int quantiser_scale(uint32_t code, uint32_t non_linear, int *q, errctx *e) {
    if (code == 0) return fail(e, "zero MPEG-2 quantiser_scale_code");
    *q = non_linear ? m2v_nonlinear_q[code] : (int)code * 2;
    return 0;
}

static int skip_extra(bits_t *b, errctx *e) {
    for (;;) {
        int bit;
        if (get_bit(b, &bit, e)) return -1;
        if (!bit) return 0;
        if (bits_left(b) < 8) return fail(e, "truncated MPEG-2 extra information");
        b->pos += 8;
    }
}

static int mb_increment(bits_t *b, uint32_t *inc, errctx *e) {
    uint32_t add = 0;
    for (;;) {
        int s;
        if (vlc(b, m2v_mbai, 36, &s, e)) return -1;
        if (s <= 32) {
            *inc = add + (uint32_t)s + 1;
            return 0;
        }
        if (s == 33) add += 33;
        else if (s == 35) return fail(e, "unexpected slice-end code");
        /* 34: stuffing */
    }
}

static int receive_signed(bits_t *b, int size, int32_t *v, errctx *e) {
    if (size == 0) {
        *v = 0;
        return 0;
    }
    uint32_t x;
    if (get_bits(b, size, &x, e)) return -1;
    const int32_t threshold = 1 << (size - 1);
    *v = (int32_t)x < threshold ? (int32_t)x + 1 - (1 << size) : (int32_t)x;
    return 0;
}

static int run_level(bits_t *b, int sym, int *run, int32_t *level, errctx *e) {
    if (sym == SYM_ESCAPE) {
        uint32_t r, raw;
        if (get_bits(b, 6, &r, e) || get_bits(b, 12, &raw, e)) return -1;
        const int32_t lv = (raw & 0x800) ? (int32_t)raw - 0x1000 : (int32_t)raw;
        if (lv == 0) return fail(e, "zero escaped MPEG-2 coefficient");
        *run = (int)r;
        *level = lv;
        return 0;
    }
    int sign;
    if (get_bit(b, &sign, e)) return -1;
    *run = m2v_run[sym];
    *level = sign ? -(int32_t)m2v_level[sym] : (int32_t)m2v_level[sym];
    return 0;
}

/* Levels of an intra block by scan position; slot 0 = the DC predictor after the differential. */
static int intra_block(bits_t *b, int block, int32_t dc_pred[3], const m2v_picture *p, int16_t out[64], errctx *e) {
    memset(out, 0, 64 * sizeof(int16_t));
    const int comp = block < 4 ? 0 : block - 3;
    int size;
    if (comp == 0 ? vlc(b, m2v_dc_luma, 12, &size, e) : vlc(b, m2v_dc_chroma, 12, &size, e)) return -1;
    int32_t diff;
    if (receive_signed(b, size, &diff, e)) return -1;
    dc_pred[comp] += diff;
    if (dc_pred[comp] < -32768 || dc_pred[comp] > 32767) return fail(e, "DC predictor outside 16 bits");
    out[0] = (int16_t)dc_pred[comp];
    const m2v_vlc *t = p->intra_vlc_format ? m2v_dct1 : m2v_dct0;
    int pos = 0;
    for (;;) {
        int sym;
        if (vlc(b, t, 113, &sym, e)) return -1;
        if (sym == SYM_EOB) return 0;
        int run;
        int32_t level;
        if (run_level(b, sym, &run, &level, e)) return -1;
        pos += run + 1;
        if (pos >= 64) return fail(e, "intra coefficient exceeds block");
        out[pos] = (int16_t)level;
    }
}

static int non_intra_block(bits_t *b, int16_t out[64], errctx *e) {
    memset(out, 0, 64 * sizeof(int16_t));
    int pos = -1;
    if (bits_left(b) < 1) return fail(e, "unexpected end of slice data");
    if ((b->p[b->pos >> 3] >> (7 - (b->pos & 7))) & 1) {   /* first coefficient "1s" */
        ++b->pos;
        int sign;
        if (get_bit(b, &sign, e)) return -1;
        out[0] = sign ? -1 : 1;
        pos = 0;
    }
    for (;;) {
        int sym;
        if (vlc(b, m2v_dct0, 113, &sym, e)) return -1;
        if (sym == SYM_EOB) return 0;
        int run;
        int32_t level;
        if (run_level(b, sym, &run, &level, e)) return -1;
        pos += run + 1;
        if (pos < 0 || pos >= 64) return fail(e, "non-intra coefficient exceeds block");
        out[pos] = (int16_t)level;
    }
}

static int motion_component(bits_t *b, int f_code, int32_t predictor, int32_t *out, errctx *e) {
    if (f_code == 0 || f_code > 7) return fail(e, "invalid MPEG-2 f_code");
    int code;
    if (vlc(b, m2v_motion, 17, &code, e)) return -1;
    if (code == 0) {
        *out = predictor;
        return 0;
    }
    int negative;
    if (get_bit(b, &negative, e)) return -1;
    const int shift = f_code - 1;
    uint32_t residual = 0;
    if (shift && get_bits(b, shift, &residual, e)) return -1;
    int32_t delta = ((code - 1) << shift) + (int32_t)residual + 1;
    if (negative) delta = -delta;
    const int32_t value = predictor + delta;
    const int nb = 5 + shift;
    const int32_t modulus = 1 << nb, sign = 1 << (nb - 1);
    const int32_t wrapped = value & (modulus - 1);
    *out = (wrapped & sign) ? wrapped - modulus : wrapped;
    return 0;
}

typedef struct {
    int has_f, has_b, field;
    int32_t f[2], bk[2];       /* frame vectors, or the parity-0 field vectors */
    int32_t f1[2], bk1[2];     /* parity-1 field vectors */
    uint8_t select[2][2];      /* [direction][parity] */
} motion_t;

static int grow_levels(m2v_picture *p, uint32_t need, errctx *e) {
    if (need <= p->levels_cap) return 0;
    size_t cap = p->levels_cap ? p->levels_cap : 1024;
    while (cap < need) cap *= 2;
    int16_t *n = (int16_t *)realloc(p->levels, cap * 64 * sizeof(int16_t));
    if (!n) return fail(e, "out of memory");
    p->levels = n;
    p->levels_cap = cap;
    return 0;
}

static void put_record(m2v_picture *p, uint32_t a, uint32_t flags, uint32_t cbp, int q, const motion_t *m, uint32_t offset) {
    mmr_recon_mpeg2_macroblock *r = &p->mbs[a];
    memset(r, 0, sizeof *r);
    r->flags = (uint8_t)flags;
    r->cbp = (uint8_t)cbp;
    r->qscale = (uint8_t)(q > 255 ? 255 : q);
    if (m) {
        r->mv[0][0] = (int16_t)m->f[0];
        r->mv[0][1] = (int16_t)m->f[1];
        r->mv[1][0] = (int16_t)m->bk[0];
        r->mv[1][1] = (int16_t)m->bk[1];
    }
    r->block_offset = offset;
    if (flags & MMR_MPEG2_MB_FIELD_MOTION) {
        ++p->field_motion_mbs;
        mmr_recon_mpeg2_field_motion *f = &p->field[a];
        f->mv[0][0] = (int16_t)m->f1[0];
        f->mv[0][1] = (int16_t)m->f1[1];
        f->mv[1][0] = (int16_t)m->bk1[0];
        f->mv[1][1] = (int16_t)m->bk1[1];
        memcpy(f->field_select, m->select, sizeof f->field_select);
    }
}

static uint32_t motion_flags(const motion_t *m) {
    return (m->has_f ? MMR_MPEG2_MB_FORWARD : 0u) | (m->has_b ? MMR_MPEG2_MB_BACKWARD : 0u) |
           (m->field ? MMR_MPEG2_MB_FIELD_MOTION : 0u);
}

static int decode_picture(const uint8_t *data, m2v_picture *p, errctx *e) {
    const uint32_t mb_w = (p->width + 15) / 16, mb_h = (p->height + 15) / 16;
    p->mb_count = mb_w * mb_h;
    p->mbs = (mmr_recon_mpeg2_macroblock *)calloc(p->mb_count, sizeof(mmr_recon_mpeg2_macroblock));
    p->field = (mmr_recon_mpeg2_field_motion *)calloc(p->mb_count, sizeof(mmr_recon_mpeg2_field_motion));
    uint8_t *done = (uint8_t *)calloc(p->mb_count, 1);
    if (!p->mbs || !p->field || !done) {
        free(done);
        return fail(e, "out of memory");
    }
    const int type = (int)p->picture_type;
    int rc = -1;
    for (size_t si = 0; si < p->nslices; ++si) {
        bits_t b = {data + p->slices[si].begin, (p->slices[si].end - p->slices[si].begin) * 8, 0};
        if (p->height > 2800) b.pos += 3;
        uint32_t qcode;
        int q;
        if (get_bits(&b, 5, &qcode, e) || quantiser_scale(qcode, p->q_scale_type, &q, e) || skip_extra(&b, e)) goto out;
        if (p->slices[si].vertical_position == 0) {
            fail(e, "zero MPEG-2 slice row");
            goto out;
        }
        uint32_t inc;
        if (mb_increment(&b, &inc, e)) goto out;
        uint32_t address = (p->slices[si].vertical_position - 1) * mb_w + inc - 1;
        const int32_t dc_reset = 128 << p->intra_dc_precision;
        int32_t dc_pred[3] = {dc_reset, dc_reset, dc_reset};
        int32_t pmv[2][2][2];
        memset(pmv, 0, sizeof pmv);
        motion_t prev_b;
        memset(&prev_b, 0, sizeof prev_b);
        for (;;) {
            if (address >= p->mb_count) {
                fail(e, "macroblock address exceeds the picture");
                goto out;
            }
            if (done[address]) {
                fail(e, "duplicate MPEG-2 macroblock address");
                goto out;
            }
            int mbt;
            if (type == MMR_MPEG2_PICTURE_I) {
                int bit;
                if (get_bit(&b, &bit, e)) goto out;
                if (bit) {
                    mbt = MB_INTRA;
                } else {
                    if (get_bit(&b, &bit, e)) goto out;
                    if (!bit) {
                        fail(e, "invalid I-picture macroblock type");
                        goto out;
                    }
                    mbt = MB_INTRA | MB_QUANT;
                }
            } else if (type == MMR_MPEG2_PICTURE_P) {
                if (vlc_sym(&b, m2v_mb_p, 7, &mbt, e)) goto out;
            } else {
                if (vlc_sym(&b, m2v_mb_b, 11, &mbt, e)) goto out;
            }
            int interlaced = 0;
            if (mbt & MB_INTRA) {
                if (!p->frame_pred_frame_dct && get_bit(&b, &interlaced, e)) goto out;
                if (mbt & MB_QUANT) {
                    if (get_bits(&b, 5, &qcode, e) || quantiser_scale(qcode, p->q_scale_type, &q, e)) goto out;
                }
                memset(pmv, 0, sizeof pmv);
                if (grow_levels(p, p->coded_blocks + 6, e)) goto out;
                const uint32_t first = p->coded_blocks;
                for (int blk = 0; blk < 6; ++blk)
                    if (intra_block(&b, blk, dc_pred, p, p->levels + (size_t)(p->coded_blocks++) * 64, e)) goto out;
                put_record(p, address, MMR_MPEG2_MB_INTRA | (interlaced ? MMR_MPEG2_MB_INTERLACED_DCT : 0u), 0x3f, q, NULL, first);
            } else {
                dc_pred[0] = dc_pred[1] = dc_pred[2] = dc_reset;
                motion_t m;
                memset(&m, 0, sizeof m);
                int motion_type = 2;   /* 1 field, 2 frame */
                if (mbt & (MB_FORWARD | MB_BACKWARD)) {
                    if (!p->frame_pred_frame_dct) {
                        uint32_t t;
                        if (get_bits(&b, 2, &t, e)) goto out;
                        if (t == 0) {
                            fail(e, "zero MPEG-2 frame_motion_type");
                            goto out;
                        }
                        if (t == 3) {
                            fail(e, "MPEG-2 dual-prime prediction is unsupported");
                            goto out;
                        }
                        motion_type = (int)t;
                    }
                } else if (type == MMR_MPEG2_PICTURE_P && (mbt & MB_PATTERN)) {
                    m.has_f = 1;   /* no motion compensation: zero forward vector */
                    memset(pmv[0], 0, sizeof pmv[0]);
                }
                if (!p->frame_pred_frame_dct && (mbt & MB_PATTERN) && get_bit(&b, &interlaced, e)) goto out;
                if (mbt & MB_QUANT) {
                    if (get_bits(&b, 5, &qcode, e) || quantiser_scale(qcode, p->q_scale_type, &q, e)) goto out;
                }
                for (int dir = 0; dir < 2; ++dir) {
                    if (!(mbt & (dir ? MB_BACKWARD : MB_FORWARD))) continue;
                    int32_t *vec = dir ? m.bk : m.f;
                    int32_t *vec1 = dir ? m.bk1 : m.f1;
                    const uint8_t *fc = p->f_code[dir];
                    if (motion_type == 1) {   /* field: a vector and a field select per parity */
                        for (int fld = 0; fld < 2; ++fld) {
                            int select;
                            int32_t x, y;
                            if (get_bit(&b, &select, e) || motion_component(&b, fc[0], pmv[dir][fld][0], &x, e) ||
                                motion_component(&b, fc[1], pmv[dir][fld][1] >> 1, &y, e))
                                goto out;
                            pmv[dir][fld][0] = x;
                            pmv[dir][fld][1] = y * 2;
                            int32_t *v = fld ? vec1 : vec;
                            v[0] = x;
                            v[1] = y;
                            m.select[dir][fld] = (uint8_t)select;
                        }
                        m.field = 1;
                    } else {
                        int32_t x, y;
                        if (motion_component(&b, fc[0], pmv[dir][0][0], &x, e) || motion_component(&b, fc[1], pmv[dir][0][1], &y, e))
                            goto out;
                        pmv[dir][0][0] = pmv[dir][1][0] = x;
                        pmv[dir][0][1] = pmv[dir][1][1] = y;
                        vec[0] = x;
                        vec[1] = y;
                    }
                    if (dir) m.has_b = 1;
                    else m.has_f = 1;
                }
                uint32_t cbp = 0;
                if (mbt & MB_PATTERN) {
                    int s;
                    if (vlc(&b, m2v_cbp, 64, &s, e)) goto out;
                    cbp = (uint32_t)s;
                }
                const uint32_t first = p->coded_blocks;
                if (grow_levels(p, p->coded_blocks + 6, e)) goto out;
                for (int blk = 0; blk < 6; ++blk)
                    if (cbp & (1u << (5 - blk)))
                        if (non_intra_block(&b, p->levels + (size_t)(p->coded_blocks++) * 64, e)) goto out;
                put_record(p, address, motion_flags(&m) | (interlaced ? MMR_MPEG2_MB_INTERLACED_DCT : 0u), cbp, q, &m, first);
                prev_b = m;
            }
            done[address] = 1;
            if (remaining_zero(&b)) break;
            if (mb_increment(&b, &inc, e)) goto out;
            for (uint32_t a = address + 1; a < address + inc; ++a) {   /* skipped macroblocks */
                if (a >= p->mb_count || done[a]) {
                    fail(e, "invalid skipped macroblock address");
                    goto out;
                }
                motion_t m;
                memset(&m, 0, sizeof m);
                if (type == MMR_MPEG2_PICTURE_P) {
                    memset(pmv, 0, sizeof pmv);
                    m.has_f = 1;
                } else if (type == MMR_MPEG2_PICTURE_B) {
                    m = prev_b;
                } else {
                    fail(e, "skipped macroblock in intra picture");
                    goto out;
                }
                dc_pred[0] = dc_pred[1] = dc_pred[2] = dc_reset;
                if (!m.has_f && !m.has_b) {
                    fail(e, "skipped B macroblock has no previous prediction mode");
                    goto out;
                }
                put_record(p, a, motion_flags(&m), 0, q, &m, p->coded_blocks);
                done[a] = 1;
            }
            address += inc;
        }
    }
    for (uint32_t a = 0; a < p->mb_count; ++a)
        if (!done[a]) {
            fail(e, "MPEG-2 picture contains undecoded macroblocks");
            goto out;
        }
    rc = 0;
out:
    free(done);
    return rc;
}

/* ------------------------------------------------------------------ stream parsing */

static int push_picture(oracle_mpeg2_stream *s, m2v_picture **out, errctx *e) {
    if (s->count == s->cap) {
        const uint32_t cap = s->cap ? s->cap * 2 : 16;
        m2v_picture *n = (m2v_picture *)realloc(s->pics, cap * sizeof(m2v_picture));
        if (!n) return fail(e, "out of memory");
        s->pics = n;
        s->cap = cap;
    }
    *out = &s->pics[s->count++];
    memset(*out, 0, sizeof **out);
    return 0;
}

static size_t next_start(const uint8_t *d, size_t len, size_t from) {
    for (size_t i = from; i + 3 < len; ++i)
        if (d[i] == 0 && d[i + 1] == 0 && d[i + 2] == 1) return i;
    return len;
}

int oracle_mpeg2_parse(const uint8_t *data, size_t len, oracle_mpeg2_stream **out, char *err, size_t errcap) {
    errctx e = {err, errcap};
    *out = NULL;
    oracle_mpeg2_stream *s = (oracle_mpeg2_stream *)calloc(1, sizeof *s);
    if (!s) return fail(&e, "out of memory");
    seq_state seq;
    memset(&seq, 0, sizeof seq);
    m2v_picture *cur = NULL;
    uint32_t gop_base = 0;   /* pictures before the current group of pictures */
    int32_t older = -1, newer = -1;
    size_t at = next_start(data, len, 0);
    if (at >= len) {
        oracle_mpeg2_free(s);
        return fail(&e, "MPEG-2 elementary stream contains no start code");
    }
#define FINISH()                                                                                                  \
    do {                                                                                                          \
        if (cur) {                                                                                                \
            if (!cur->have_coding_ext) goto bad_noext;                                                            \
            if (cur->picture_structure != 3) goto bad_field;                                                      \
            cur->forward_index = cur->picture_type == MMR_MPEG2_PICTURE_B ? older : (cur->picture_type == MMR_MPEG2_PICTURE_P ? newer : -1); \
            cur->backward_index = cur->picture_type == MMR_MPEG2_PICTURE_B ? newer : -1;                          \
            if (decode_picture(data, cur, &e)) goto bad;                                                          \
            if (cur->picture_type != MMR_MPEG2_PICTURE_B) {                                                       \
                older = newer;                                                                                    \
                newer = (int32_t)(cur - s->pics);                                                                 \
            }                                                                                                     \
            cur = NULL;                                                                                           \
        }                                                                                                         \
    } while (0)
    while (at < len) {
        if (at + 3 >= len) {
            fail(&e, "truncated MPEG start code");
            goto bad;
        }
        const uint8_t code = data[at + 3];
        const size_t begin = at + 4, end = next_start(data, len, begin);
        bits_t b = {data + begin, (end - begin) * 8, 0};
        if (code == 0xb3) {   /* sequence header */
            FINISH();
            uint32_t w, h, skip;
            if (get_bits(&b, 12, &w, &e) || get_bits(&b, 12, &h, &e) || get_bits(&b, 4 + 4 + 18 + 1 + 10 + 1, &skip, &e)) goto bad;
            memset(&seq, 0, sizeof seq);
            seq.have_seq = 1;
            seq.width = w;
            seq.height = h;
            int load;
            if (get_bit(&b, &load, &e)) goto bad;
            if (load) {
                if (read_matrix(&b, seq.seq_intra, &e)) goto bad;
                seq.seq_intra_loaded = 1;
            }
            if (get_bit(&b, &load, &e)) goto bad;
            if (load) {
                if (read_matrix(&b, seq.seq_non_intra, &e)) goto bad;
                seq.seq_non_intra_loaded = 1;
            }
        } else if (code == 0xb5) {   /* extension */
            uint32_t id;
            if (get_bits(&b, 4, &id, &e)) goto bad;
            if (id == 1) {   /* sequence extension */
                if (cur) {
                    fail(&e, "sequence extension inside picture");
                    goto bad;
                }
                uint32_t pl, prog, cf, hx, vx;
                if (get_bits(&b, 8, &pl, &e) || get_bits(&b, 1, &prog, &e) || get_bits(&b, 2, &cf, &e) || get_bits(&b, 2, &hx, &e) ||
                    get_bits(&b, 2, &vx, &e))
                    goto bad;
                seq.chroma_format = cf;
                seq.width |= hx << 12;
                seq.height |= vx << 12;
                seq.have_ext = 1;
            } else if (id == 3) {   /* quant matrix extension: merged into the sequence state */
                if (!seq.have_seq) {
                    fail(&e, "quant-matrix extension without sequence");
                    goto bad;
                }
                for (int k = 0; k < 4; ++k) {
                    int load;
                    if (get_bit(&b, &load, &e)) goto bad;
                    if (load) {
                        if (read_matrix(&b, seq.qm[k], &e)) goto bad;
                        seq.qm_loaded[k] = 1;
                        if (cur) memcpy(cur->matrices[k], seq.qm[k], 64);
                    }
                }
            } else if (id == 8) {   /* picture coding extension */
                if (!cur) {
                    fail(&e, "picture coding extension without picture");
                    goto bad;
                }
                uint32_t f[4], v[11];
                for (int k = 0; k < 4; ++k)
                    if (get_bits(&b, 4, &f[k], &e)) goto bad;
                if (get_bits(&b, 2, &v[0], &e) || get_bits(&b, 2, &v[1], &e)) goto bad;
                for (int k = 2; k < 11; ++k)
                    if (get_bits(&b, 1, &v[k], &e)) goto bad;
                cur->f_code[0][0] = (uint8_t)f[0];
                cur->f_code[0][1] = (uint8_t)f[1];
                cur->f_code[1][0] = (uint8_t)f[2];
                cur->f_code[1][1] = (uint8_t)f[3];
                cur->intra_dc_precision = v[0];
                cur->picture_structure = v[1];
                cur->frame_pred_frame_dct = v[3];
                cur->q_scale_type = v[5];
                cur->intra_vlc_format = v[6];
                cur->alternate_scan = v[7];
                cur->progressive_frame = v[10];
                cur->have_coding_ext = 1;
            }
        } else if (code == 0xb8) {   /* group of pictures: temporal references restart */
            FINISH();
            gop_base = s->count;
        } else if (code == 0x00) {   /* picture */
            FINISH();
            if (!seq.have_seq || !seq.have_ext) {
                fail(&e, "picture without an MPEG-2 sequence header and extension");
                goto bad;
            }
            if (seq.chroma_format != 1) {
                fail(&e, "only 4:2:0 MPEG-2 video is supported");
                goto bad;
            }
            if (push_picture(s, &cur, &e)) goto bad;
            uint32_t tr, pt;
            if (get_bits(&b, 10, &tr, &e) || get_bits(&b, 3, &pt, &e)) goto bad;
            if (pt < 1 || pt > 3) {
                fail(&e, "unsupported picture coding type");
                goto bad;
            }
            cur->picture_type = pt;   /* 1 I, 2 P, 3 B = MMR_MPEG2_PICTURE_* */
            cur->display_index = gop_base + tr;
            cur->width = seq.width;
            cur->height = seq.height;
            resolve_matrices(&seq, cur->matrices);
        } else if (code >= 0x01 && code <= 0xaf) {   /* slice */
            if (!cur) {
                fail(&e, "slice outside a picture");
                goto bad;
            }
            if (cur->nslices == cur->cap_slices) {
                const size_t cap = cur->cap_slices ? cur->cap_slices * 2 : 64;
                void *n = realloc(cur->slices, cap * sizeof(*cur->slices));
                if (!n) {
                    fail(&e, "out of memory");
                    goto bad;
                }
                cur->slices = n;
                cur->cap_slices = cap;
            }
            uint32_t vpos = code;
            if (seq.height > 2800) {
                uint32_t ext;
                if (get_bits(&b, 3, &ext, &e)) goto bad;
                vpos |= ext << 7;
            }
            cur->slices[cur->nslices].vertical_position = vpos;
            cur->slices[cur->nslices].begin = begin;
            cur->slices[cur->nslices].end = end;
            ++cur->nslices;
        } else if (code == 0xb7) {   /* sequence end */
            FINISH();
            seq.have_seq = seq.have_ext = 0;
        }
        at = end;
    }
    FINISH();
#undef FINISH
    if (s->count == 0) {
        fail(&e, "MPEG-2 stream contains no coded pictures");
        goto bad;
    }
    *out = s;
    return 0;
bad_noext:
    fail(&e, "picture has no coding extension");
    goto bad;
bad_field:
    fail(&e, "MPEG-2 field pictures are not reconstructed");
bad:
    oracle_mpeg2_free(s);
    return -1;
}

uint32_t oracle_mpeg2_picture_count(const oracle_mpeg2_stream *s) { return s->count; }

int oracle_mpeg2_picture_info(const oracle_mpeg2_stream *s, uint32_t index, oracle_mpeg2_picture *out) {
    if (index >= s->count) return -1;
    const m2v_picture *p = &s->pics[index];
    out->width = p->width;
    out->height = p->height;
    out->picture_type = p->picture_type;
    out->display_index = p->display_index;
    out->picture_structure = p->picture_structure;
    out->frame_pred_frame_dct = p->frame_pred_frame_dct;
    out->progressive_frame = p->progressive_frame;
    out->alternate_scan = p->alternate_scan;
    out->intra_dc_precision = p->intra_dc_precision;
    out->forward_index = p->forward_index;
    out->backward_index = p->backward_index;
    out->mb_count = p->mb_count;
    out->coded_blocks = p->coded_blocks;
    out->field_motion_mbs = p->field_motion_mbs;
    out->matrices = &p->matrices[0][0];
    out->macroblocks = p->mbs;
    out->field_motion = p->field;
    out->levels = p->levels;
    return 0;
}

void oracle_mpeg2_free(oracle_mpeg2_stream *s) {
    if (!s) return;
    for (uint32_t i = 0; i < s->count; ++i) {
        free(s->pics[i].slices);
        free(s->pics[i].mbs);
        free(s->pics[i].field);
        free(s->pics[i].levels);
    }
    free(s->pics);
    free(s);
}
