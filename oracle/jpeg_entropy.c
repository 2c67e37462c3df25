/*
 * CPU JPEG entropy front end -- TEST INFRASTRUCTURE.
 *
 * Stands in for the reference's CPU entropy side so that real baseline JPEG / MJPEG files can be
 * fed to the reconstruction path: it parses the markers, decodes the Huffman-coded scan and emits
 * the post-entropy buffers of one frame in the layout of mmr_recon_mjpeg_frame (levels by zig-zag
 * position, slot 0 = the DC prediction value, in the MCU walk's block order; quantisation tables
 * in zig-zag order). Only tests/ use it; the product library never links it.
 *
 * Behaviour follows the reference:
 *   marker walk               crates/codecs/mjpeg/src/parser.rs:18-81 (SOI first, segments with
 *                             lengths, SOS followed by entropy data up to the next non-RST marker)
 *   DQT / SOF0 / DHT / SOS    parser.rs:149-345 (8- or 16-bit DQT values in zig-zag order; SOF0
 *                             8-bit precision, sampling 1..4, table selectors 0..3)
 *   DRI                       decoder.rs:259-297 (tables collected from segments before the scan)
 *   canonical Huffman codes   entropy.rs:19-57 (codes assigned by length, oversubscription error)
 *   bit reader                entropy.rs:73-170 (MSB first, FF00 stuffing, fill bytes FF..FF,
 *                             any other marker inside coded data is an error)
 *   receive_extend            entropy.rs:172-182
 *   restart handling          decoder.rs:361-371 (RSTn expected every restart_interval MCUs,
 *                             numbers modulo 8, DC predictions reset)
 *   block decoding            decoder.rs:502-548 (DC category <= 11, AC run/size symbols, ZRL,
 *                             EOB, category <= 10, runs past the block end rejected)
 *   scan validation           decoder.rs:215-252 (baseline sequential parameters, every frame
 *                             component selected once)
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

typedef struct {
    uint8_t present;
    uint16_t code[256];
    uint8_t length[256];
    uint8_t symbol[256];
    int count;
} huff_table;

typedef struct {
    const uint8_t *data;
    size_t len, pos;
    uint8_t byte;
    int bits_left;
    char *err;
    size_t errcap;
} bit_reader;

static int fail(char *err, size_t cap, const char *msg, size_t offset) {
    if (err && cap) snprintf(err, cap, "%s (byte 0x%08zx)", msg, offset);
    return -1;
}

/* Canonical code assignment (entropy.rs:19-57). */
static int build_huffman(const uint8_t counts[16], const uint8_t *symbols, int nsym, huff_table *t) {
    int total = 0;
    for (int i = 0; i < 16; ++i) total += counts[i];
    if (total == 0 || total != nsym || total > 256) return -1;
    uint32_t code = 0;
    int k = 0;
    for (int len = 1; len <= 16; ++len) {
        const uint32_t limit = 1u << len;
        if (code + counts[len - 1] > limit) return -1;   /* oversubscribed */
        for (int i = 0; i < counts[len - 1]; ++i) {
            t->code[k] = (uint16_t)code;
            t->length[k] = (uint8_t)len;
            t->symbol[k] = symbols[k];
            ++k;
            ++code;
        }
        code <<= 1;
    }
    t->count = total;
    t->present = 1;
    return 0;
}

/* Next entropy-coded byte: FF00 is a stuffed FF, FF fill bytes are skipped, other markers end
 * the data with an error (entropy.rs:143-170). */
static int read_data_byte(bit_reader *r, uint8_t *out) {
    if (r->pos >= r->len) return fail(r->err, r->errcap, "unexpected end of entropy-coded data", r->pos);
    const uint8_t b = r->data[r->pos++];
    if (b != 0xff) {
        *out = b;
        return 0;
    }
    while (r->pos < r->len && r->data[r->pos] == 0xff) ++r->pos;
    if (r->pos >= r->len) return fail(r->err, r->errcap, "truncated entropy marker", r->pos);
    const uint8_t code = r->data[r->pos++];
    if (code != 0x00) return fail(r->err, r->errcap, "unexpected marker inside coded block", r->pos);
    *out = 0xff;
    return 0;
}

static int read_bit(bit_reader *r, int *bit) {
    if (r->bits_left == 0) {
        if (read_data_byte(r, &r->byte)) return -1;
        r->bits_left = 8;
    }
    --r->bits_left;
    *bit = (r->byte >> r->bits_left) & 1;
    return 0;
}

static int read_bits(bit_reader *r, int n, uint32_t *v) {
    uint32_t x = 0;
    for (int i = 0; i < n; ++i) {
        int b;
        if (read_bit(r, &b)) return -1;
        x = (x << 1) | (uint32_t)b;
    }
    *v = x;
    return 0;
}

/* Bit by bit, the first code of matching length and value (entropy.rs:59-71). */
static int huff_decode(bit_reader *r, const huff_table *t, int *symbol) {
    uint32_t bits = 0;
    for (int len = 1; len <= 16; ++len) {
        int b;
        if (read_bit(r, &b)) return -1;
        bits = (bits << 1) | (uint32_t)b;
        for (int k = 0; k < t->count; ++k)
            if (t->length[k] == len && t->code[k] == bits) {
                *symbol = t->symbol[k];
                return 0;
            }
    }
    return fail(r->err, r->errcap, "no Huffman code matches the next 16 bits", r->pos);
}

static int receive_extend(bit_reader *r, int size, int32_t *value) {
    if (size == 0) {
        *value = 0;
        return 0;
    }
    if (size > 15) return fail(r->err, r->errcap, "coefficient magnitude contains more than 15 bits", r->pos);
    uint32_t bits;
    if (read_bits(r, size, &bits)) return -1;
    const int32_t threshold = 1 << (size - 1);
    *value = (int32_t)bits < threshold ? (int32_t)bits + 1 - (1 << size) : (int32_t)bits;
    return 0;
}

/* RSTn at an MCU boundary (entropy.rs:116-139). */
static int consume_restart(bit_reader *r, int expected) {
    r->bits_left = 0;
    if (r->pos >= r->len || r->data[r->pos] != 0xff) return fail(r->err, r->errcap, "expected restart marker at MCU boundary", r->pos);
    ++r->pos;
    while (r->pos < r->len && r->data[r->pos] == 0xff) ++r->pos;
    if (r->pos >= r->len) return fail(r->err, r->errcap, "truncated restart marker", r->pos);
    const uint8_t code = r->data[r->pos++];
    if (code != 0xd0 + expected) return fail(r->err, r->errcap, "unexpected restart marker number", r->pos - 1);
    return 0;
}

/* One block: DC difference added to the prediction (slot 0 holds the prediction value), AC
 * levels at their zig-zag positions (decoder.rs:502-548 before dequantisation). */
static int decode_block(bit_reader *r, const huff_table *dc, const huff_table *ac, int32_t *prediction, int16_t out[64]) {
    memset(out, 0, 64 * sizeof(int16_t));
    int size;
    if (huff_decode(r, dc, &size)) return -1;
    if (size > 11) return fail(r->err, r->errcap, "baseline DC coefficient category exceeds 11", r->pos);
    int32_t diff;
    if (receive_extend(r, size, &diff)) return -1;
    *prediction += diff;
    if (*prediction < -32768 || *prediction > 32767) return fail(r->err, r->errcap, "DC prediction outside 16 bits", r->pos);
    out[0] = (int16_t)*prediction;
    int k = 1;
    while (k < 64) {
        int sym;
        if (huff_decode(r, ac, &sym)) return -1;
        const int run = sym >> 4, sz = sym & 15;
        if (sz == 0) {
            if (run == 0) break;   /* EOB */
            if (run != 15) return fail(r->err, r->errcap, "invalid zero-size AC Huffman symbol", r->pos);
            k += 16;               /* ZRL */
            if (k > 64) return fail(r->err, r->errcap, "AC zero run exceeds the end of the block", r->pos);
            continue;
        }
        if (sz > 10) return fail(r->err, r->errcap, "baseline AC coefficient category exceeds 10", r->pos);
        k += run;
        if (k >= 64) return fail(r->err, r->errcap, "AC coefficient run exceeds the end of the block", r->pos);
        int32_t v;
        if (receive_extend(r, sz, &v)) return -1;
        out[k++] = (int16_t)v;   /* |v| < 2^10 */
    }
    return 0;
}

static uint16_t be16(const uint8_t *p) { return (uint16_t)((p[0] << 8) | p[1]); }

int oracle_jpeg_entropy(const uint8_t *data, size_t len, oracle_jpeg_levels *out, char *err, size_t errcap) {
    memset(out, 0, sizeof *out);
    huff_table dc[4], ac[4];
    memset(dc, 0, sizeof dc);
    memset(ac, 0, sizeof ac);
    uint8_t q_present[4] = {0, 0, 0, 0};
    uint8_t comp_id[4] = {0, 0, 0, 0};
    uint32_t restart_interval = 0;
    int have_frame = 0;
    if (len < 2 || data[0] != 0xff || data[1] != 0xd8) return fail(err, errcap, "expected SOI as the first marker", 0);
    size_t pos = 2;
    for (;;) {
        /* marker: FF fill bytes, then the code */
        if (pos >= len || data[pos] != 0xff) return fail(err, errcap, "expected a marker", pos);
        const size_t marker_at = pos;
        while (pos < len && data[pos] == 0xff) ++pos;
        if (pos >= len) return fail(err, errcap, "truncated marker", pos);
        const uint8_t m = data[pos++];
        if (m == 0xd9) {   /* EOI */
            out->consumed = pos;
            if (!out->levels) return fail(err, errcap, "image has no scan", marker_at);
            return 0;
        }
        if (m == 0x01 || (m >= 0xd0 && m <= 0xd7)) continue;   /* no length */
        if (pos + 2 > len) return fail(err, errcap, "truncated segment length", pos);
        const uint16_t seg = be16(data + pos);
        if (seg < 2) return fail(err, errcap, "segment length must include its two length bytes", pos);
        const uint8_t *p = data + pos + 2;
        const size_t n = (size_t)seg - 2;
        if (pos + 2 + n > len) return fail(err, errcap, "truncated segment payload", pos + 2);
        const size_t payload_at = pos + 2;
        pos += 2 + n;
        if (m == 0xdb) {   /* DQT */
            size_t i = 0;
            while (i < n) {
                const int precision = p[i] >> 4, id = p[i] & 15;
                if (precision > 1 || id > 3) return fail(err, errcap, "invalid DQT table specification", payload_at + i);
                ++i;
                const size_t bytes = precision ? 128 : 64;
                if (n - i < bytes) return fail(err, errcap, "truncated DQT values", payload_at + i);
                if (precision) return fail(err, errcap, "16-bit quantization tables are not allowed in baseline JPEG", payload_at + i);
                for (int k = 0; k < 64; ++k) out->quant[id][k] = p[i + (size_t)k];
                i += bytes;
                q_present[id] = 1;
            }
        } else if (m == 0xc4) {   /* DHT */
            size_t i = 0;
            while (i < n) {
                if (n - i < 17) return fail(err, errcap, "truncated DHT", payload_at + i);
                const int cls = p[i] >> 4, id = p[i] & 15;
                if (cls > 1 || id > 3) return fail(err, errcap, "invalid DHT table class or id", payload_at + i);
                const uint8_t *counts = p + i + 1;
                int total = 0;
                for (int k = 0; k < 16; ++k) total += counts[k];
                if (n - i - 17 < (size_t)total) return fail(err, errcap, "truncated DHT symbols", payload_at + i);
                huff_table *t = cls ? &ac[id] : &dc[id];
                if (build_huffman(counts, p + i + 17, total, t)) return fail(err, errcap, "inconsistent Huffman table", payload_at + i);
                i += 17 + (size_t)total;
            }
        } else if (m == 0xdd) {   /* DRI */
            if (n != 2) return fail(err, errcap, "DRI payload must be 2 bytes", payload_at);
            restart_interval = be16(p);
        } else if (m == 0xc0) {   /* SOF0 */
            if (n < 6) return fail(err, errcap, "SOF0 payload is too short", payload_at);
            const uint32_t nc = p[5];
            if (nc == 0 || nc > 4 || n != 6 + 3 * nc) return fail(err, errcap, "invalid SOF0 component count", payload_at + 5);
            if (p[0] != 8) return fail(err, errcap, "baseline SOF0 sample precision must be 8 bits", payload_at);
            out->height = be16(p + 1);
            out->width = be16(p + 3);
            if (out->width == 0 || out->height == 0) return fail(err, errcap, "SOF0 dimensions must be non-zero", payload_at + 1);
            out->component_count = nc;
            for (uint32_t c = 0; c < nc; ++c) {
                comp_id[c] = p[6 + 3 * c];
                out->h[c] = p[7 + 3 * c] >> 4;
                out->v[c] = p[7 + 3 * c] & 15;
                out->qsel[c] = p[8 + 3 * c];
                if (out->h[c] < 1 || out->h[c] > 4 || out->v[c] < 1 || out->v[c] > 4 || out->qsel[c] > 3)
                    return fail(err, errcap, "invalid SOF0 component", payload_at + 6 + 3 * c);
            }
            have_frame = 1;
        } else if (m >= 0xc1 && m <= 0xcf && m != 0xc4 && m != 0xc8 && m != 0xcc) {
            return fail(err, errcap, "only baseline sequential JPEG (SOF0) is supported", marker_at);
        } else if (m == 0xda) {   /* SOS, then the entropy-coded scan */
            if (!have_frame) return fail(err, errcap, "scan before frame header", marker_at);
            if (out->levels) return fail(err, errcap, "more than one scan", marker_at);
            if (n < 1) return fail(err, errcap, "SOS payload is too short", payload_at);
            const uint32_t ns = p[0];
            if (ns != out->component_count || n != 4 + 2 * ns) return fail(err, errcap, "scan must select every frame component once", payload_at);
            int dct[4], act[4];
            for (uint32_t s = 0; s < ns; ++s) {
                /* blocks are emitted in frame component order, so the scan must use that order */
                if (p[1 + 2 * s] != comp_id[s]) return fail(err, errcap, "scan component order differs from the frame's", payload_at + 1 + 2 * s);
                dct[s] = p[2 + 2 * s] >> 4;
                act[s] = p[2 + 2 * s] & 15;
                if (dct[s] > 3 || act[s] > 3 || !dc[dct[s]].present || !ac[act[s]].present)
                    return fail(err, errcap, "scan references a missing Huffman table", payload_at + 2 + 2 * s);
                if (!q_present[out->qsel[s]]) return fail(err, errcap, "component references a missing quantization table", payload_at);
            }
            const uint8_t *tail = p + 1 + 2 * ns;
            if (tail[0] != 0 || tail[1] != 63 || tail[2] != 0) return fail(err, errcap, "scan parameters are not baseline sequential", payload_at);
            /* MCU walk (decoder.rs:346-401) */
            const uint32_t nc = out->component_count;
            uint32_t hmax = 1, vmax = 1;
            for (uint32_t c = 0; c < nc; ++c) {
                if (out->h[c] > hmax) hmax = out->h[c];
                if (out->v[c] > vmax) vmax = out->v[c];
            }
            const int interleaved = nc > 1;
            const uint32_t mcu_w = interleaved ? hmax * 8 : 8, mcu_h = interleaved ? vmax * 8 : 8;
            const uint32_t cols = (out->width + mcu_w - 1) / mcu_w, rows = (out->height + mcu_h - 1) / mcu_h;
            uint32_t per_mcu = 0;
            for (uint32_t c = 0; c < nc; ++c) per_mcu += interleaved ? (uint32_t)out->h[c] * out->v[c] : 1;
            out->block_count = cols * rows * per_mcu;
            out->levels = (int16_t *)calloc((size_t)out->block_count * 64, sizeof(int16_t));
            if (!out->levels) return fail(err, errcap, "out of memory", payload_at);
            bit_reader r = {data, len, pos, 0, 0, err, errcap};
            int32_t pred[4] = {0, 0, 0, 0};
            int rst = 0;
            uint32_t blk = 0;
            for (uint32_t mcu = 0; mcu < cols * rows; ++mcu) {
                if (mcu > 0 && restart_interval && mcu % restart_interval == 0) {
                    if (consume_restart(&r, rst)) return -1;
                    rst = (rst + 1) & 7;
                    memset(pred, 0, sizeof pred);
                }
                for (uint32_t c = 0; c < nc; ++c) {
                    const uint32_t nb = interleaved ? (uint32_t)out->h[c] * out->v[c] : 1;
                    for (uint32_t b = 0; b < nb; ++b, ++blk)
                        if (decode_block(&r, &dc[dct[c]], &ac[act[c]], &pred[c], out->levels + (size_t)blk * 64)) return -1;
                }
            }
            /* the scan's coded data ends at the next marker that is not RSTn or stuffing */
            pos = r.pos;
            while (pos < len) {
                if (data[pos] != 0xff) {
                    ++pos;
                    continue;
                }
                size_t q = pos + 1;
                while (q < len && data[q] == 0xff) ++q;
                if (q >= len) return fail(err, errcap, "truncated marker in entropy-coded data", q);
                if (data[q] == 0x00 || (data[q] >= 0xd0 && data[q] <= 0xd7)) {
                    pos = q + 1;
                    continue;
                }
                break;   /* pos is at the FF of the next marker */
            }
            if (pos >= len) return fail(err, errcap, "entropy-coded scan has no terminating marker", len);
        }
        /* APPn, COM and other segments with lengths are skipped */
    }
}

void oracle_jpeg_free(oracle_jpeg_levels *f) {
    free(f->levels);
    f->levels = NULL;
}

/* Known-answer helpers for the tests: decode `n` symbols with a table built from counts/symbols,
 * and receive_extend `n` values of the given sizes, from raw entropy-coded bytes. */
int oracle_jpeg_huffman_symbols(const uint8_t counts[16], const uint8_t *symbols, int nsym, const uint8_t *data, size_t len,
                                int *out, int n) {
    huff_table t;
    memset(&t, 0, sizeof t);
    if (build_huffman(counts, symbols, nsym, &t)) return -1;
    bit_reader r = {data, len, 0, 0, 0, NULL, 0};
    for (int i = 0; i < n; ++i)
        if (huff_decode(&r, &t, &out[i])) return -1;
    return 0;
}

int oracle_jpeg_read_bits(const uint8_t *data, size_t len, int nbits, uint32_t *out) {
    bit_reader r = {data, len, 0, 0, 0, NULL, 0};
    return read_bits(&r, nbits, out);
}

int oracle_jpeg_receive_extend(const uint8_t *data, size_t len, const int *sizes, int n, int32_t *out) {
    bit_reader r = {data, len, 0, 0, 0, NULL, 0};
    for (int i = 0; i < n; ++i)
        if (receive_extend(&r, sizes[i], &out[i])) return -1;
    return 0;
}
