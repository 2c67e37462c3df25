/*
 * Oracle: a scalar C restatement of MMRecode's reconstruction functions.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product library links, loads or
 * calls this code; tests, the smoke check and the CPU-baseline leg of the
 * benchmark use it as the reference answer (and as the timed CPU baseline).
 *
 * Every function follows the operation order, integer widths, rounding and
 * non-fused f64 arithmetic of the Rust function it cites (paths relative to
 * the MMRecode repository). Descriptors use the layouts of include/mmr_recon.h
 * but all pointers are HOST pointers here.
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off, no -ffast-math).
 */
#ifndef MMR_ORACLE_H
#define MMR_ORACLE_H

#include <stdint.h>
#include "../include/mmr_recon.h"

#ifdef __cplusplus
extern "C" {
#endif

/* MJPEG (crates/codecs/mjpeg/src) */
void oracle_mjpeg_inverse_dct(const int32_t coefficients[64], uint8_t out[64]);
void oracle_mjpeg_dequantize(const int16_t levels[64], const uint16_t quant_zigzag[64],
                             int32_t coefficients[64]);
uint32_t oracle_mjpeg_block_count(const mmr_recon_mjpeg_frame *frame);
int oracle_mjpeg_reconstruct(const mmr_recon_mjpeg_frame *frame);

/* JPEG entropy front end (jpeg_entropy.c): one baseline frame's post-entropy buffers. */
typedef struct oracle_jpeg_levels {
    uint32_t width, height, component_count;
    uint8_t h[4], v[4], qsel[4];
    uint16_t quant[4][64];       /* zig-zag order, rows by table id */
    uint32_t block_count;
    int16_t *levels;             /* [block_count][64] zig-zag order, [0] = DC prediction value */
    size_t consumed;             /* input bytes up to and including EOI */
} oracle_jpeg_levels;
int oracle_jpeg_entropy(const uint8_t *data, size_t len, oracle_jpeg_levels *out, char *err, size_t errcap);
void oracle_jpeg_free(oracle_jpeg_levels *frame);
int oracle_jpeg_huffman_symbols(const uint8_t counts[16], const uint8_t *symbols, int nsym, const uint8_t *data,
                                size_t len, int *out, int n);
int oracle_jpeg_read_bits(const uint8_t *data, size_t len, int nbits, uint32_t *out);
int oracle_jpeg_receive_extend(const uint8_t *data, size_t len, const int *sizes, int n, int32_t *out);

/* MPEG-2 (crates/codecs/mpeg2/src) */
void oracle_mpeg2_inverse_dct(const int32_t coefficients[64], int16_t out[64]);
void oracle_mpeg2_dequantize_intra(const int16_t levels[64], int component, int qscale,
                                   int intra_dc_precision, int alternate_scan,
                                   const uint8_t matrix[64], int32_t coefficients[64]);
void oracle_mpeg2_dequantize_non_intra(const int16_t levels[64], int qscale, int alternate_scan,
                                       const uint8_t matrix[64], int32_t coefficients[64]);
uint8_t oracle_mpeg2_halfpel_sample(const uint8_t *reference, int stride, int height, int x,
                                    int y, int mv_x, int mv_y);
uint8_t oracle_mpeg2_halfpel_field_sample(const uint8_t *reference, int stride, int height, int x, int dest_y,
                                          int mv_x, int mv_y, int select);
int oracle_mpeg2_reconstruct(const mmr_recon_mpeg2_picture *picture);

/* MPEG-2 entropy front end (mpeg2_entropy.c): a parsed stream's frame pictures in decode order. */
typedef struct oracle_mpeg2_stream oracle_mpeg2_stream;
typedef struct oracle_mpeg2_picture {
    uint32_t width, height, picture_type, alternate_scan, intra_dc_precision;
    int32_t forward_index, backward_index;   /* decode-order indices of the references, -1 none */
    uint32_t mb_count, coded_blocks, field_motion_mbs;
    const uint8_t *matrices;                 /* [4][64] natural: intra, non-intra, chroma intra, chroma non-intra */
    const mmr_recon_mpeg2_macroblock *macroblocks;
    const mmr_recon_mpeg2_field_motion *field_motion;   /* [mb_count] */
    const int16_t *levels;                   /* [coded_blocks][64] by scan position */
    uint32_t display_index;                  /* pictures before its GOP + temporal_reference */
    uint32_t picture_structure, frame_pred_frame_dct, progressive_frame;   /* picture coding extension */
} oracle_mpeg2_picture;
int oracle_mpeg2_parse(const uint8_t *data, size_t len, oracle_mpeg2_stream **out, char *err, size_t errcap);
uint32_t oracle_mpeg2_picture_count(const oracle_mpeg2_stream *stream);
int oracle_mpeg2_picture_info(const oracle_mpeg2_stream *stream, uint32_t index, oracle_mpeg2_picture *out);
void oracle_mpeg2_free(oracle_mpeg2_stream *stream);

/* DV25 entropy front end (dv_entropy.c): one raw DIF frame -> mmr_recon_dv_frame buffers.
 * levels [sequences * 810][64] and info [sequences * 810]; concealed segments as seq << 8 | slot. */
int oracle_dv_vlc_table(uint16_t *code, uint8_t *len, uint8_t *run, int16_t *level, int capacity);
int oracle_dv_entropy(const uint8_t *data, size_t len, int conceal, uint32_t *system, int16_t *levels, uint8_t *info,
                      uint32_t *concealed, uint32_t concealed_cap, uint32_t *concealed_count, char *err, size_t errcap);

/* DV (crates/codecs/dv/src) */
void oracle_dv_inverse_88(const int32_t coefficients[64], uint8_t out[64]);
void oracle_dv_inverse_248(const int32_t coefficients[64], uint8_t out[64]);
int32_t oracle_dv_inverse_factor(int quantization, int class_number, int dct_mode, int position);
void oracle_dv_dequantize(const int16_t levels[64], uint8_t block_info, int32_t coefficients[64]);
void oracle_dv_macroblock_coordinates(uint32_t system, int sequence, int slot, int out_xy[10]);
uint32_t oracle_dv_block_count(uint32_t system);
int oracle_dv_reconstruct(const mmr_recon_dv_frame *frame);
const double *oracle_dv_cos16_table(void); /* [x*8+u] = cos((2x+1)u pi/16) */
const double *oracle_dv_cos8_table(void);  /* [y*4+v] = cos((2y+1)v pi/8) */

/* H.264 (crates/codecs/h264/src/decoder.rs, deblock.rs) */
void oracle_h264_bypass_residual_4x4(const int32_t levels[16], int intra_mode, int32_t out[16]);
int oracle_h264_transform_residual_4x4(const int32_t levels[16], int qp, int dc_already_scaled,
                                       const uint8_t scaling[16], int32_t out[16]);
int oracle_h264_transform_residual_8x8(const int32_t levels[64], int qp,
                                       const uint8_t scaling[64], int32_t out[64]);
int oracle_h264_transform_chroma_dc(const int32_t levels[4], int qp, const uint8_t scaling[16],
                                    int32_t out[4]);
int oracle_h264_chroma_qp(int luma_qp, int offset);
uint8_t oracle_h264_luma_qpel(const uint8_t *plane, int width, int height, int x_q4, int y_q4);
uint8_t oracle_h264_chroma_epel(const uint8_t *plane, int width, int height, int x_q8, int y_q8);
uint8_t oracle_h264_apply_weight(uint8_t sample, int log2_denom, int weight, int offset);
uint8_t oracle_h264_apply_bi_weight(uint8_t s0, uint8_t s1, int log2_denom, int w0, int w1,
                                    int o0, int o1, int explicit_weights);
void oracle_h264_implicit_weights(int32_t poc0, int32_t poc1, int32_t current_poc, int lt0,
                                  int lt1, int32_t *w0, int32_t *w1);
/* H.264 intra prediction (h264_intra.c); 0 or -1 when a needed neighbour is missing / bad mode. */
int oracle_h264_intra4_predict(const uint8_t *plane, int stride, int ox, int oy, int block_index, int mode,
                               int top_mb, int left_mb, uint8_t out[16]);
int oracle_h264_intra8_predict(const uint8_t *plane, int stride, int ox, int oy, int group, int mode, int top_mb,
                               int left_mb, uint8_t out[64]);
int oracle_h264_intra_block_predict(const uint8_t *plane, int stride, int ox, int oy, int size, int mode, int has_top,
                                    int has_left, uint8_t *out);
int oracle_h264_chroma_predict(const uint8_t *plane, int stride, int ox, int oy, int mode, int has_top, int has_left,
                               uint8_t out[64]);
int oracle_h264_intra16_dc(const int32_t levels[16], int qp, const uint8_t scaling[16], int32_t out[16]);
int oracle_h264_reconstruct(const mmr_recon_h264_picture *picture);
int oracle_h264_deblock(const mmr_recon_h264_picture *picture);
void oracle_h264_filter_samples(uint8_t *plane, const int p_idx[4], const int q_idx[4], int qp,
                                int bs, int chroma, int offset_a, int offset_b);

/* colour + blend (crates/cli/src/terminal_preview.rs, crates/render/src/frame_graph.rs,
 * crates/mmfx/src/render.rs) */
void oracle_ycbcr_to_rgb(uint8_t y, uint8_t cb, uint8_t cr, uint32_t range, uint32_t matrix,
                         uint8_t rgb[3]);
int oracle_yuv_to_rgb(const mmr_recon_yuv_to_rgb_job *job);
void oracle_blend_rgba_pixel(uint8_t destination[4], const uint8_t source[4]);
void oracle_premultiplied_over_pixel(uint8_t destination[4], const uint8_t source[4],
                                     uint32_t opacity);
uint8_t oracle_blend_channel(uint8_t destination, uint8_t source, uint8_t alpha);
/* H.264 CAVLC front end (h264_entropy.c): access units in, the library's records out. */
typedef struct oracle_h264_decoder oracle_h264_decoder;
typedef struct {
    uint32_t width_mbs, height_mbs;
    int32_t slot;               /* id of this picture; reference lists name pictures by it */
    int32_t poc;
    int32_t is_reference;
    uint32_t slice_type;        /* first slice: 0 P, 1 B, 2 I */
    uint32_t ref_count[2];
    int32_t ref_slot[2][16];
    int32_t ref_poc[2][16];
    uint32_t coded_mbs;         /* level records */
    uint32_t weighted_mode;
    uint32_t deblock;
    int32_t chroma_qp_offset[2];
    uint32_t transform_bypass;
} oracle_h264_picture_info;
oracle_h264_decoder *oracle_h264_decoder_new(void);
void oracle_h264_decoder_free(oracle_h264_decoder *d);
int oracle_h264_cavlc_block(const uint8_t *bytes, size_t n, int nc, int max_coeff, int32_t *coef, size_t *bits_used);
int oracle_h264_decode_access_unit(oracle_h264_decoder *d, const uint8_t *data, size_t size, oracle_h264_picture_info *info);
int oracle_h264_picture_records(const oracle_h264_decoder *d, mmr_recon_h264_macroblock *mbs, mmr_recon_h264_block *blocks,
                                int16_t *levels, mmr_recon_h264_intra *intra, mmr_recon_h264_weight *weights,
                                mmr_recon_h264_bi_weight *bi_weights, int16_t *mvd);
int oracle_composite(const mmr_recon_composite_job *job);
uint16_t oracle_linear_multiply_channel(uint16_t left, uint16_t right);
uint16_t oracle_srgb_to_linear_u16(uint8_t value);
uint8_t oracle_linear_u16_to_srgb(uint16_t value);
void oracle_linear_from_rgba(const uint8_t rgba[4], uint16_t out[4]);
void oracle_linear_over(const uint16_t src[4], uint16_t dst[4], uint16_t opacity);
void oracle_linear_to_rgba8(const uint16_t px[4], uint8_t out[4]);
int oracle_linear_composite(const mmr_recon_linear_composite_job *job);
int oracle_linear_convert(const mmr_recon_linear_convert_job *job);
int oracle_yuv_blend(const mmr_recon_yuv_blend_job *job);
void oracle_rgb_to_yuv(uint8_t red, uint8_t green, uint8_t blue, uint8_t limited[3], uint8_t full[3]);
int oracle_overlay_prepare(const mmr_recon_overlay_prepare_job *job);
int oracle_scale_taps(uint32_t filter, uint32_t n, uint32_t m, uint32_t o, int32_t *left, float *w, int cap);
void oracle_scaled_dimensions(uint32_t sw, uint32_t sh, uint32_t cw, uint32_t ch, uint32_t mode, uint32_t *ow, uint32_t *oh);
int oracle_scale_yuv420(const mmr_recon_scale_job *job);

#ifdef __cplusplus
}
#endif

#endif
