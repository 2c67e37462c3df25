/*
 * CPU DV25 entropy front end -- TEST INFRASTRUCTURE.
 *
 * Stands in for the reference's CPU VLC side so that real DV25 frames can be fed to the
 * reconstruction path: it checks the DIF layout, decodes every video segment's coefficients with
 * the three-level bit spill and emits the post-entropy buffers of mmr_recon_dv_frame (per block
 * the signed 9-bit DC and the levels by scan position, plus dct_mode / class / quantisation
 * number). Only tests/ use it; the product library never links it.
 *
 * Behaviour follows the reference:
 *   profile detection                 crates/codecs/dv/src/profile.rs:87-128 (header block first,
 *                                     DSF bit of byte 3, exact frame size)
 *   DIF identifiers / expected layout crates/codecs/dv/src/dif.rs:56-64, :180-258 (section,
 *                                     sequence and block number of every video block)
 *   segment / macroblock decoding     crates/codecs/dv/src/decoder.rs:150-248 (five macroblocks
 *                                     per segment at sequence * 150 + 7 + 16 * group + index,
 *                                     6 blocks of 112 / 112 / 112 / 112 / 80 / 80 bits, DC 9 bits
 *                                     signed, dct_mode, class; spill inside the macroblock,
 *                                     then inside the segment; every block must end complete)
 *   coefficient VLCs                  decoder.rs:250-288, :312-352 (canonical codes assigned in
 *                                     table order, a sign bit after codes with a level, the
 *                                     run / level escapes, EOB = run 127; fewer than 16 bits
 *                                     left without a match wait for the next pool)
 *   concealment                       decoder.rs:100-146 (a segment that fails is left at the
 *                                     planes' initial Y = 16, C = 128)
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* ------------------------------------------------------------------ coefficient VLC table */

/* IEC 61834 coefficient codes by length (the values of crates/codecs/dv/src/tables.rs:51-105):
 * (run, level) per code, lengths 2..12, in canonical order. Level 0 with run 127 is EOB; level 0
 * otherwise is a run escape. Lengths 13 (run escapes, run 0..63) and 15 (level escapes, level
 * 0..255) follow as loops. */
typedef struct { uint8_t run, level; } dv_rl;
static const dv_rl kLen2[] = {{0, 1}};
static const dv_rl kLen3[] = {{0, 2}};
static const dv_rl kLen4[] = {{127, 0}, {1, 1}, {0, 3}, {0, 4}};
static const dv_rl kLen5[] = {{2, 1}, {1, 2}, {0, 5}, {0, 6}};
static const dv_rl kLen6[] = {{3, 1}, {4, 1}, {0, 7}, {0, 8}};
static const dv_rl kLen7[] = {{5, 1}, {6, 1}, {2, 2}, {1, 3}, {1, 4}, {0, 9}, {0, 10}, {0, 11}};
static const dv_rl kLen8[] = {{7, 1}, {8, 1}, {9, 1}, {10, 1}, {3, 2}, {4, 2}, {2, 3}, {1, 5},
                              {1, 6}, {1, 7}, {0, 12}, {0, 13}, {0, 14}, {0, 15}, {0, 16}, {0, 17}};
static const dv_rl kLen9[] = {{11, 1}, {12, 1}, {13, 1}, {14, 1}, {5, 2}, {6, 2}, {3, 3}, {4, 3},
                              {2, 4}, {2, 5}, {1, 8}, {0, 18}, {0, 19}, {0, 20}, {0, 21}, {0, 22}};
static const dv_rl kLen10[] = {{5, 3}, {3, 4}, {3, 5}, {2, 6}, {1, 9}, {1, 10}, {1, 11}};
static const dv_rl kLen11[] = {{0, 0}, {1, 0}, {6, 3}, {4, 4}, {3, 6}, {1, 12}, {1, 13}, {1, 14}};
static const dv_rl kLen12[] = {{2, 0}, {3, 0}, {4, 0}, {5, 0}, {7, 2}, {8, 2}, {9, 2}, {10, 2}, {7, 3}, {8, 3},
                               {4, 5}, {3, 7}, {2, 7}, {2, 8}, {2, 9}, {2, 10}, {2, 11}, {1, 15}, {1, 16}, {1, 17}};

typedef struct {
    uint16_t code;
    uint8_t len, run;
    int16_t level;
} dv_vlc;

#define DV_VLC_MAX 746
static dv_vlc g_vlc[DV_VLC_MAX];
static int g_nvlc;
static uint32_t g_canonical;
static int16_t g_lut[1 << 16];   /* next 16 bits -> the entry whose code prefixes them */

static void add_code(int len, int run, int level) {
    const uint32_t code = g_canonical >> (32 - len);
    g_canonical += 1u << (32 - len);
    if (level == 0) {
        g_vlc[g_nvlc++] = (dv_vlc){(uint16_t)code, (uint8_t)len, (uint8_t)run, 0};
    } else {   /* sign bit: 0 positive, 1 negative */
        g_vlc[g_nvlc++] = (dv_vlc){(uint16_t)(code << 1), (uint8_t)(len + 1), (uint8_t)run, (int16_t)level};
        g_vlc[g_nvlc++] = (dv_vlc){(uint16_t)((code << 1) | 1u), (uint8_t)(len + 1), (uint8_t)run, (int16_t)-level};
    }
}

static void build_vlc(void) {
    if (g_nvlc) return;
    static const struct { const dv_rl *e; int n; } by_len[11] = {
        {kLen2, 1}, {kLen3, 1}, {kLen4, 4}, {kLen5, 4}, {kLen6, 4}, {kLen7, 8},
        {kLen8, 16}, {kLen9, 16}, {kLen10, 7}, {kLen11, 8}, {kLen12, 20}};
    g_canonical = 0;
    for (int l = 0; l < 11; ++l)
        for (int i = 0; i < by_len[l].n; ++i) add_code(l + 2, by_len[l].e[i].run, by_len[l].e[i].level);
    for (int run = 0; run < 64; ++run) add_code(13, run, 0);
    for (int level = 0; level < 256; ++level) add_code(15, 0, level);
    for (int i = 0; i < (1 << 16); ++i) g_lut[i] = -1;
    for (int i = 0; i < g_nvlc; ++i) {
        const int shift = 16 - g_vlc[i].len;
        for (uint32_t s = 0; s < (1u << shift); ++s) g_lut[((uint32_t)g_vlc[i].code << shift) | s] = (int16_t)i;
    }
}

int oracle_dv_vlc_table(uint16_t *code, uint8_t *len, uint8_t *run, int16_t *level, int capacity) {
    build_vlc();
    if (capacity < g_nvlc) return -1;
    for (int i = 0; i < g_nvlc; ++i) {
        code[i] = g_vlc[i].code;
        len[i] = g_vlc[i].len;
        run[i] = g_vlc[i].run;
        level[i] = g_vlc[i].level;
    }
    return g_nvlc;
}

/* ------------------------------------------------------------------ bit vectors */

/* The reference keeps bits one per byte and moves them between pools; so does this front end. */
typedef struct {
    uint8_t *b;
    size_t n, cap;
} bitvec;

static int bv_append(bitvec *v, const uint8_t *bits, size_t n) {
    if (v->n + n > v->cap) {
        size_t cap = v->cap ? v->cap : 256;
        while (cap < v->n + n) cap *= 2;
        uint8_t *p = (uint8_t *)realloc(v->b, cap);
        if (!p) return -1;
        v->b = p;
        v->cap = cap;
    }
    if (n) memcpy(v->b + v->n, bits, n);
    v->n += n;
    return 0;
}

static uint32_t bits_value(const uint8_t *bits, int n) {
    uint32_t v = 0;
    for (int i = 0; i < n; ++i) v = (v << 1) | bits[i];
    return v;
}

/* decode_vlc (decoder.rs:312-317): the entry whose code matches the next len bits, or -1 when
 * none fits in the n bits given. The codes form a prefix code of at most 16 bits, so at most one
 * entry matches; it is the one the zero-padded next 16 bits select, if its length is <= n. */
static int decode_vlc(const uint8_t *bits, size_t n) {
    const int avail = n < 16 ? (int)n : 16;
    const int k = g_lut[bits_value(bits, avail) << (16 - avail)];
    return k >= 0 && (size_t)g_vlc[k].len <= n ? k : -1;
}

/* ------------------------------------------------------------------ blocks */

typedef struct {
    int16_t *levels;   /* [64] by scan position, [0] = DC */
    int position;
    int complete;
    bitvec partial;
} dv_block;

typedef struct {
    char *err;
    size_t cap;
} errctx;

static int fail(errctx *e, const char *msg) {
    if (e->err && e->cap) snprintf(e->err, e->cap, "%s", msg);
    return -1;
}

/* consume_ac (decoder.rs:250-288): leftover bits after the block's EOB go to *out. */
static int consume_ac(dv_block *blk, const uint8_t *input, size_t n, bitvec *out, errctx *e) {
    out->n = 0;
    if (blk->complete) return bv_append(out, input, n);
    bitvec bits = blk->partial;
    blk->partial = (bitvec){0, 0, 0};
    if (bv_append(&bits, input, n)) return fail(e, "out of memory");
    size_t cursor = 0;
    for (;;) {
        const size_t left = bits.n - cursor;
        const int k = decode_vlc(bits.b + cursor, left);
        if (k >= 0) {
            cursor += g_vlc[k].len;
            blk->position += g_vlc[k].run + 1;
            if (blk->position >= 64) {
                blk->complete = 1;
                const int r = bv_append(out, bits.b + cursor, bits.n - cursor);
                free(bits.b);
                return r;
            }
            if (g_vlc[k].level != 0) blk->levels[blk->position] = g_vlc[k].level;
        } else if (left < 16) {
            const int r = bv_append(&blk->partial, bits.b + cursor, left);
            free(bits.b);
            return r;
        } else {
            free(bits.b);
            return fail(e, "invalid DV variable-length coefficient code");
        }
    }
}

/* consume_pool (decoder.rs:236-248): the pool feeds the incomplete blocks in order. */
static int consume_pool(dv_block *blocks, int count, bitvec *pool, errctx *e) {
    bitvec next = {0, 0, 0};
    for (int i = 0; i < count; ++i) {
        if (blocks[i].complete || pool->n == 0) continue;
        if (consume_ac(&blocks[i], pool->b, pool->n, &next, e)) {
            free(next.b);
            return -1;
        }
        bitvec t = *pool;
        *pool = next;
        next = t;
    }
    free(next.b);
    return 0;
}

/* decode_macroblock (decoder.rs:201-234) into blocks[6]; the bits the macroblock leaves go to
 * *overflow. */
static int decode_macroblock(const uint8_t *bytes, dv_block *blocks, uint8_t *info, bitvec *overflow, errctx *e) {
    static const int kBlockBits[6] = {112, 112, 112, 112, 80, 80};
    uint8_t payload[76 * 8];
    for (int i = 0; i < 76; ++i)
        for (int s = 0; s < 8; ++s) payload[i * 8 + s] = (bytes[4 + i] >> (7 - s)) & 1;
    const uint32_t q = bytes[3] & 0x0f;
    bitvec pool = {0, 0, 0}, left = {0, 0, 0};
    int cursor = 0, rc = 0;
    for (int b = 0; b < 6 && !rc; ++b) {
        const uint8_t *region = payload + cursor;
        cursor += kBlockBits[b];
        int32_t dc = (int32_t)bits_value(region, 9);
        if (region[0]) dc -= 1 << 9;
        const uint32_t mode = region[9], cls = bits_value(region + 10, 2);
        blocks[b].levels[0] = (int16_t)dc;
        info[b] = (uint8_t)(mode | cls << 1 | q << 3);
        rc = consume_ac(&blocks[b], region + 12, (size_t)(kBlockBits[b] - 12), &left, e);
        if (!rc && blocks[b].complete) rc = bv_append(&pool, left.b, left.n) ? fail(e, "out of memory") : 0;
    }
    if (!rc) rc = consume_pool(blocks, 6, &pool, e);
    if (!rc) rc = bv_append(overflow, pool.b, pool.n) ? fail(e, "out of memory") : 0;
    free(pool.b);
    free(left.b);
    return rc;
}

/* DIF identifier of the video block at physical index `index` (dif.rs:180-258, :235-258). */
static int video_block_damaged(const uint8_t *block, uint32_t index) {
    const uint32_t sequence = index / 150, position = index % 150;
    const uint32_t av = position - 6;
    const uint32_t number = av - av / 16 - 1;
    return (block[0] >> 5) != 4 || (uint32_t)(block[1] >> 4) != sequence || block[2] != number;
}

int oracle_dv_entropy(const uint8_t *data, size_t len, int conceal, uint32_t *system, int16_t *levels, uint8_t *info,
                      uint32_t *concealed, uint32_t concealed_cap, uint32_t *concealed_count, char *err, size_t errcap) {
    errctx e = {err, errcap};
    build_vlc();
    if (len < 4) return fail(&e, "DV profile detection needs 4 header bytes");
    if (data[0] >> 5 != 0) return fail(&e, "first DIF block is not a header");
    const int pal = (data[3] & 0x80) != 0;
    const uint32_t sequences = pal ? 12 : 10;
    const size_t frame_size = (size_t)sequences * 150 * 80;
    if (len != frame_size) return fail(&e, "DV header identifies a frame of another size");
    *system = pal ? MMR_DV_SYSTEM_625_50 : MMR_DV_SYSTEM_525_60;
    *concealed_count = 0;
    const size_t blocks_total = (size_t)sequences * 27 * 5 * 6;
    memset(levels, 0, blocks_total * 64 * sizeof(int16_t));
    memset(info, 0, blocks_total);
    dv_block blocks[30];
    char reason[160];
    for (uint32_t seq = 0; seq < sequences; ++seq)
        for (uint32_t slot = 0; slot < 27; ++slot) {
            const size_t first = ((size_t)seq * 27 + slot) * 30;
            errctx se = {reason, sizeof reason};
            int rc = 0;
            for (int i = 0; i < 30; ++i) {
                blocks[i] = (dv_block){levels + (first + i) * 64, 0, 0, {0, 0, 0}};
                memset(blocks[i].levels, 0, 64 * sizeof(int16_t));
            }
            bitvec segment_pool = {0, 0, 0};
            for (int m = 0; m < 5 && !rc; ++m) {
                const uint32_t video = slot * 5 + (uint32_t)m;
                const uint32_t index = seq * 150 + 7 + (video / 15) * 16 + video % 15;
                const uint8_t *bytes = data + (size_t)index * 80;
                if (video_block_damaged(bytes, index)) {
                    rc = fail(&se, "DV video segment has a damaged DIF identifier");
                    break;
                }
                rc = decode_macroblock(bytes, blocks + m * 6, info + first + m * 6, &segment_pool, &se);
            }
            if (!rc) rc = consume_pool(blocks, 30, &segment_pool, &se);
            for (int i = 0; i < 30 && !rc; ++i)
                if (!blocks[i].complete) rc = fail(&se, "DV video segment has incomplete coefficients");
            for (int i = 0; i < 30; ++i) free(blocks[i].partial.b);
            free(segment_pool.b);
            if (!rc) continue;
            if (!conceal) {
                if (err && errcap) snprintf(err, errcap, "sequence %u slot %u: %s", seq, slot, reason);
                return -1;
            }
            /* Left at the initial planes: DC-only blocks of Y 16 (DC -224: (4 * -224 + 1024) / 8) and
             * C 128 (DC 0), frame DCT. */
            for (int i = 0; i < 30; ++i) {
                int16_t *lv = levels + (first + i) * 64;
                memset(lv, 0, 64 * sizeof(int16_t));
                lv[0] = (int16_t)(i % 6 < 4 ? -224 : 0);
                info[first + i] = 0;
            }
            if (*concealed_count < concealed_cap) concealed[*concealed_count] = seq << 8 | slot;
            ++*concealed_count;
        }
    return 0;
}
