/* Oracle: MJPEG reconstruction restated from crates/codecs/mjpeg/src. Test infrastructure only. */
#include <math.h>
#include <string.h>

#include "oracle.h"
#include "tables.h"

/* inverse_dct, crates/codecs/mjpeg/src/transform.rs:86-109.
 * Row pass tmp[v*8+x] += M[x][u] * c[v*8+u] (u ascending, from 0.0), then column pass
 * sum += M[y][v] * tmp[v*8+x] (v ascending), (sum + 128).round().clamp(0, 255). */
void oracle_mjpeg_inverse_dct(const int32_t coefficients[64], uint8_t out[64]) {
    double temporary[64];
    for (int i = 0; i < 64; ++i) temporary[i] = 0.0;
    for (int v = 0; v < 8; ++v)
        for (int x = 0; x < 8; ++x)
            for (int u = 0; u < 8; ++u) {
                double product = ORACLE_IDCT_MATRIX[x][u] * (double)coefficients[v * 8 + u];
                temporary[v * 8 + x] = temporary[v * 8 + x] + product;
            }
    for (int y = 0; y < 8; ++y)
        for (int x = 0; x < 8; ++x) {
            double sum = 0.0;
            for (int v = 0; v < 8; ++v) {
                double product = ORACLE_IDCT_MATRIX[y][v] * temporary[v * 8 + x];
                sum = sum + product;
            }
            double r = round(sum + 128.0);
            if (r < 0.0) r = 0.0;
            if (r > 255.0) r = 255.0;
            out[y * 8 + x] = (uint8_t)r;
        }
}

/* Dequantisation of decode_block, crates/codecs/mjpeg/src/decoder.rs:502-546:
 * coefficients[0] = dc_prediction * Q[0]; AC at zigzag k: level * Q[k] stored at
 * ZIGZAG_TO_NATURAL[k]. levels[0] holds the DC predictor (prediction stays on the CPU). */
void oracle_mjpeg_dequantize(const int16_t levels[64], const uint16_t quant_zigzag[64],
                             int32_t coefficients[64]) {
    memset(coefficients, 0, 64 * sizeof(int32_t));
    coefficients[0] = (int32_t)levels[0] * (int32_t)quant_zigzag[0];
    for (int k = 1; k < 64; ++k) {
        if (levels[k] != 0) coefficients[ORACLE_ZIGZAG[k]] = (int32_t)levels[k] * (int32_t)quant_zigzag[k];
    }
}

static void component_geometry(const mmr_recon_mjpeg_frame *f, int c, int *w, int *h) {
    int hmax = 0, vmax = 0;
    for (uint32_t i = 0; i < f->component_count; ++i) {
        if (f->h_sampling[i] > hmax) hmax = f->h_sampling[i];
        if (f->v_sampling[i] > vmax) vmax = f->v_sampling[i];
    }
    /* make_component_states, decoder.rs:437-440: ceil(W * h / hmax) */
    *w = (int)((f->width * f->h_sampling[c] + hmax - 1) / hmax);
    *h = (int)((f->height * f->v_sampling[c] + vmax - 1) / vmax);
}

uint32_t oracle_mjpeg_block_count(const mmr_recon_mjpeg_frame *f) {
    int hmax = 0, vmax = 0, per_mcu = 0;
    for (uint32_t i = 0; i < f->component_count; ++i) {
        if (f->h_sampling[i] > hmax) hmax = f->h_sampling[i];
        if (f->v_sampling[i] > vmax) vmax = f->v_sampling[i];
        per_mcu += f->h_sampling[i] * f->v_sampling[i];
    }
    int interleaved = f->component_count > 1;
    int mcu_w = interleaved ? hmax * 8 : 8, mcu_h = interleaved ? vmax * 8 : 8;
    uint32_t cols = (f->width + mcu_w - 1) / mcu_w, rows = (f->height + mcu_h - 1) / mcu_h;
    return cols * rows * (interleaved ? (uint32_t)per_mcu : 1u);
}

/* write_block, decoder.rs:550-565: clip at the right and bottom plane edges. */
static void write_block(uint8_t *plane, int pw, int ph, int x, int y, const uint8_t block[64]) {
    for (int by = 0; by < 8; ++by) {
        int dy = y + by;
        if (dy >= ph) break;
        for (int bx = 0; bx < 8; ++bx) {
            int dx = x + bx;
            if (dx >= pw) break;
            plane[dy * pw + dx] = block[by * 8 + bx];
        }
    }
}

/* The MCU walk of reconstruct(), decoder.rs:313-402, with entropy decoding replaced by
 * reading the level buffer in the same block order. */
int oracle_mjpeg_reconstruct(const mmr_recon_mjpeg_frame *f) {
    if (f->component_count < 1 || f->component_count > 4) return -1;
    if (f->block_count != oracle_mjpeg_block_count(f)) return -1;
    int hmax = 0, vmax = 0;
    for (uint32_t i = 0; i < f->component_count; ++i) {
        if (f->h_sampling[i] > hmax) hmax = f->h_sampling[i];
        if (f->v_sampling[i] > vmax) vmax = f->v_sampling[i];
    }
    int interleaved = f->component_count > 1;
    int mcu_w = interleaved ? hmax * 8 : 8, mcu_h = interleaved ? vmax * 8 : 8;
    int cols = (int)((f->width + mcu_w - 1) / mcu_w), rows = (int)((f->height + mcu_h - 1) / mcu_h);
    const int16_t *levels = f->levels;
    int32_t coefficients[64];
    uint8_t samples[64];
    for (int mcu = 0; mcu < cols * rows; ++mcu) {
        int mcu_x = mcu % cols, mcu_y = mcu / cols;
        for (uint32_t c = 0; c < f->component_count; ++c) {
            int pw, ph;
            component_geometry(f, (int)c, &pw, &ph);
            int brows = interleaved ? f->v_sampling[c] : 1, bcols = interleaved ? f->h_sampling[c] : 1;
            const uint16_t *q = f->quant_tables + 64 * f->quant_table[c];
            for (int by = 0; by < brows; ++by)
                for (int bx = 0; bx < bcols; ++bx) {
                    oracle_mjpeg_dequantize(levels, q, coefficients);
                    levels += 64;
                    oracle_mjpeg_inverse_dct(coefficients, samples);
                    write_block(f->planes[c], pw, ph, (mcu_x * bcols + bx) * 8, (mcu_y * brows + by) * 8, samples);
                }
        }
    }
    if (f->rgb && f->component_count == 3) {
        mmr_recon_yuv_to_rgb_job job;
        memset(&job, 0, sizeof job);
        int cw, ch, lw, lh;
        component_geometry(f, 0, &lw, &lh);
        component_geometry(f, 1, &cw, &ch);
        job.width = f->width;
        job.height = f->height;
        job.chroma_width = (uint32_t)cw;
        job.chroma_height = (uint32_t)ch;
        job.luma_stride = (uint32_t)lw;
        job.chroma_stride = (uint32_t)cw;
        job.range = f->rgb_range ? f->rgb_range : MMR_RANGE_FULL;
        job.matrix = MMR_MATRIX_BT601;
        job.planes[0] = f->planes[0];
        job.planes[1] = f->planes[1];
        job.planes[2] = f->planes[2];
        job.rgb = f->rgb;
        return oracle_yuv_to_rgb(&job);
    }
    return 0;
}
