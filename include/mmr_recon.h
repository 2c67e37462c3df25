/*
 * mmr_recon.h — C ABI of the MI355X (gfx950) pixel-reconstruction library.
 *
 * The library reconstructs pictures from post-entropy buffers (coefficient
 * levels, macroblock records, motion vectors, prediction weights) that a CPU
 * entropy stage hands over, and composites / colour-converts the results. It
 * is the device-side replacement for the reconstruction loops of MMRecode's
 * codec crates; the entropy decoders stay on the CPU.
 *
 * Conventions follow crates/capi/include/mmrecode.h of MMRecode:
 *   - int32_t status codes with the same values (crates/capi/include/mmrecode.h:24-32),
 *     plus MMR_STATUS_BUSY for non-blocking back-pressure (DecodeExecutor::submit,
 *     crates/playback/src/executor.rs:235-245);
 *   - struct_size-prefixed descriptors (crates/capi/include/mmrecode.h:57-100);
 *   - a thread-local last error (crates/capi/include/mmrecode.h:113-118);
 *   - borrowed host inputs are only read during the call.
 *
 * Memory: every pointer inside a submit descriptor is a DEVICE pointer
 * (HBM-resident), obtained from mmr_recon_device_alloc or any other HIP
 * allocation on the context's device. Descriptors themselves are host memory
 * and are copied during the submit call. Device buffers named by a submission
 * must stay valid until its ticket completes.
 */
#ifndef MMR_RECON_H
#define MMR_RECON_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__GNUC__) && __GNUC__ >= 4
#define MMR_RECON_API __attribute__((visibility("default")))
#else
#define MMR_RECON_API
#endif

#define MMR_RECON_ABI_VERSION 1u

#ifndef MMR_ABI_VERSION
typedef int32_t mmr_status;
#define MMR_STATUS_OK ((mmr_status)0)
#define MMR_STATUS_INVALID_ARGUMENT ((mmr_status)1)
#define MMR_STATUS_INVALID_DATA ((mmr_status)2)
#define MMR_STATUS_UNSUPPORTED ((mmr_status)3)
#define MMR_STATUS_INVALID_STATE ((mmr_status)4)
#define MMR_STATUS_IO_ERROR ((mmr_status)5)
#define MMR_STATUS_INTERNAL_ERROR ((mmr_status)100)
#endif
/* Queue full or HBM budget exhausted: retry later (no work was queued). */
#define MMR_STATUS_BUSY ((mmr_status)6)

#define MMR_RANGE_FULL 1u
#define MMR_RANGE_LIMITED 2u
#define MMR_MATRIX_BT601 1u
#define MMR_MATRIX_BT709 2u

/* Submission flags. */
/* Plain path: one kernel launch per frame / picture, in submission order. */
#define MMR_SUBMIT_PER_FRAME_LAUNCH 1u

typedef struct mmr_recon_context mmr_recon_context;

typedef struct mmr_recon_config {
    size_t struct_size;
    int32_t device_ordinal;      /* HIP device ordinal */
    uint32_t queue_depth;        /* max in-flight tickets before MMR_STATUS_BUSY (0: 256) */
    uint64_t hbm_budget_bytes;   /* admission budget of the serving loop (0: unlimited) */
    uint32_t persistent_groups;  /* workgroups of a persistent launch (0: autotuned) */
    uint32_t flags;              /* reserved, 0 */
} mmr_recon_config;

/* ------------------------------------------------------------------ core */

/* Replaces: mmr_abi_version, crates/capi/include/mmrecode.h:102-103. */
MMR_RECON_API uint32_t mmr_recon_abi_version(void);

/* Replaces: mmr_last_error_message, crates/capi/include/mmrecode.h:113-118.
 * Copies the calling thread's last diagnostic (NUL-terminated when capacity
 * is nonzero) and returns the capacity required including the NUL. */
MMR_RECON_API size_t mmr_recon_last_error_message(char *buffer, size_t capacity);

/* New (no reference counterpart; the reference decoders are constructed per
 * stream, crates/core/src/codec.rs:102-130): a context owns one device, its
 * stream, descriptor arena and work-queue counters. */
MMR_RECON_API mmr_status mmr_recon_context_create(const mmr_recon_config *config,
                                                  mmr_recon_context **out_context);
MMR_RECON_API void mmr_recon_context_destroy(mmr_recon_context *context);

/* Device memory helpers for hosts without their own HIP allocator. */
MMR_RECON_API mmr_status mmr_recon_device_alloc(mmr_recon_context *context, size_t bytes,
                                                void **out_device_ptr);
MMR_RECON_API mmr_status mmr_recon_device_free(mmr_recon_context *context, void *device_ptr);
MMR_RECON_API mmr_status mmr_recon_copy_to_device(mmr_recon_context *context, void *device_dst,
                                                  const void *host_src, size_t bytes);
MMR_RECON_API mmr_status mmr_recon_copy_to_host(mmr_recon_context *context, void *host_dst,
                                                const void *device_src, size_t bytes);

/* Completion. Tickets are issued in increasing order per context. wait / poll return the
 * status of one submission; synchronize waits for every submission of the context and returns
 * the first error of a submission whose status was not yet returned by wait or poll. */
MMR_RECON_API mmr_status mmr_recon_wait(mmr_recon_context *context, uint64_t ticket);
MMR_RECON_API mmr_status mmr_recon_poll(mmr_recon_context *context, uint64_t ticket,
                                        int32_t *out_done);
MMR_RECON_API mmr_status mmr_recon_synchronize(mmr_recon_context *context);

/* Device time (milliseconds, HIP events around the launches) of the last
 * completed ticket's kernels, for roofline accounting. */
MMR_RECON_API mmr_status mmr_recon_last_kernel_ms(mmr_recon_context *context, uint64_t ticket,
                                                  double *out_ms);

/* Device time (milliseconds) of each launch phase of a completed ticket, from HIP events recorded
 * on the submission stream between the phases; the phases add up to mmr_recon_last_kernel_ms.
 * Batched submissions (flags without MMR_SUBMIT_PER_FRAME_LAUNCH) have these phases:
 *   MJPEG   [block kernel] or [block kernel, YUV->RGB] when a frame asks for RGB
 *   MPEG-2  [macroblock order kernel, all dependency-level launches of the block kernel]
 *   H.264   [boundary strengths (when deblocking), inter macroblocks, intra bands (when
 *           present), deblocking (when present)]
 *   others  one phase.
 * Per-frame launches report one phase. At most MMR_RECON_MAX_PHASES values are written to
 * out_ms (capacity entries available); *out_count receives the number of phases. No counterpart
 * in the reference (measurement of this library's own launches). */
#define MMR_RECON_MAX_PHASES 4u
MMR_RECON_API mmr_status mmr_recon_kernel_phases_ms(mmr_recon_context *context, uint64_t ticket,
                                                    double *out_ms, uint32_t capacity,
                                                    uint32_t *out_count);

/* ----------------------------------------------------------------- MJPEG */

/*
 * One baseline JPEG frame after Huffman decoding.
 * Replaces the per-block reconstruction of reconstruct(),
 * crates/codecs/mjpeg/src/decoder.rs:386-399: dequantisation
 * (decode_block :502-546), inverse_dct (crates/codecs/mjpeg/src/transform.rs:86-109)
 * and write_block (decoder.rs:550-565).
 *
 * Block order is the reference's MCU walk (decoder.rs:346-401): MCUs in raster
 * order; inside an interleaved MCU, components in order, then block rows, then
 * block columns. A single-component frame is non-interleaved (one block per
 * MCU, ceil(W/8) x ceil(H/8) MCUs).
 */
typedef struct mmr_recon_mjpeg_frame {
    size_t struct_size;
    uint32_t width;
    uint32_t height;
    uint32_t component_count;     /* 1..4 */
    uint8_t h_sampling[4];
    uint8_t v_sampling[4];
    uint8_t quant_table[4];       /* per component: row of quant_tables */
    uint32_t block_count;         /* must equal the MCU walk's block count */
    const int16_t *levels;        /* [block_count][64], zigzag order; [0] = DC predictor value */
    const uint16_t *quant_tables; /* [n][64], zigzag order (values_in_zigzag_order) */
    uint8_t *planes[4];           /* component planes, stride = component width */
    uint8_t *rgb;                 /* optional W*H*3 RGB8 (needs 3 components), NULL to skip */
    uint32_t rgb_range;           /* MMR_RANGE_FULL (JPEG default) or MMR_RANGE_LIMITED */
} mmr_recon_mjpeg_frame;

MMR_RECON_API mmr_status mmr_recon_mjpeg_submit(mmr_recon_context *context,
                                                const mmr_recon_mjpeg_frame *frames,
                                                uint32_t frame_count, uint32_t flags,
                                                uint64_t *out_ticket);

/* ---------------------------------------------------------------- MPEG-2 */

#define MMR_MPEG2_MB_INTRA 1u
#define MMR_MPEG2_MB_FORWARD 2u
#define MMR_MPEG2_MB_BACKWARD 4u
#define MMR_MPEG2_MB_INTERLACED_DCT 8u
#define MMR_MPEG2_MB_FIELD_MOTION 16u  /* field prediction: field vectors in the picture's field_motion */

/*
 * One macroblock after VLC decoding, skip resolution and motion-vector
 * prediction (crates/codecs/mpeg2/src/decoder.rs:340-679 produce these
 * values; skipped macroblocks are explicit records with cbp 0).
 */
typedef struct mmr_recon_mpeg2_macroblock {
    uint8_t flags;          /* MMR_MPEG2_MB_* */
    uint8_t cbp;            /* coded_block_pattern, bit (5-b) = block b; ignored for intra */
    uint8_t qscale;         /* quantiser_scale value (decode_quantiser_scale, decoder.rs:721-732) */
    uint8_t reserved;
    int16_t mv[2][2];       /* [forward, backward][x, y], half-sample units; with FIELD_MOTION the
                               first (top-row) field's vectors in field units */
    uint32_t block_offset;  /* index of the first coded block of this macroblock in levels
                               (below 2^24 - 6; larger values are reported INVALID_DATA) */
} mmr_recon_mpeg2_macroblock;

/*
 * Field prediction of a frame-picture macroblock (MMR_MPEG2_MB_FIELD_MOTION; the reference's
 * FieldMotion pairs, decode_field_motion decoder.rs:891-908, predicted by predict_field_plane /
 * halfpel_field_sample :1053-1148): rows of parity f are predicted from field field_select[d][f]
 * of the reference with vector f. The parity-0 vectors are the macroblock record's mv.
 */
typedef struct mmr_recon_mpeg2_field_motion {
    int16_t mv[2][2];             /* [forward, backward][x, y] of the parity-1 rows, field units */
    uint8_t field_select[2][2];   /* [forward, backward][parity]: 0 top field, 1 bottom field
                                     (other values of a FIELD_MOTION macroblock: INVALID_DATA) */
} mmr_recon_mpeg2_field_motion;

#define MMR_MPEG2_PICTURE_I 1u
#define MMR_MPEG2_PICTURE_P 2u
#define MMR_MPEG2_PICTURE_B 3u

/*
 * One frame picture. Replaces Mpeg2PictureDecoder::decode_picture's
 * reconstruction (crates/codecs/mpeg2/src/decoder.rs:263-304): intra and
 * non-intra dequantisation with mismatch control (:734-846), inverse_dct
 * (crates/codecs/mpeg2/src/transform.rs:90-110), half-sample prediction
 * (decoder.rs:944-1046, :1150-1222) and put_block (:1231-1298).
 *
 * Planes are coded-size (FrameBuffer::new, decoder.rs:103-124): luma stride
 * ceil16(width), chroma stride = luma stride / 2, heights ceil16(height) and half.
 * Coded blocks of a macroblock are stored consecutively in block order 0..5.
 */
typedef struct mmr_recon_mpeg2_picture {
    size_t struct_size;
    uint32_t width;
    uint32_t height;
    uint32_t picture_type;          /* MMR_MPEG2_PICTURE_* */
    uint32_t alternate_scan;        /* 0: ZIGZAG, 1: ALTERNATE_SCAN */
    uint32_t intra_dc_precision;    /* 0..3 */
    const uint8_t *matrices;        /* [4][64] natural order: intra, non-intra, chroma intra, chroma non-intra */
    const mmr_recon_mpeg2_macroblock *macroblocks;  /* raster order, mb_width * mb_height */
    const int16_t *levels;          /* [coded blocks][64] by scan position; intra [0] = DC predictor */
    uint8_t *planes[3];             /* output coded planes */
    const uint8_t *forward[3];      /* P: newer anchor, B: older anchor */
    const uint8_t *backward[3];     /* B: newer anchor */
    int32_t forward_source;         /* index in this submission producing forward[], or -1 */
    int32_t backward_source;        /* index in this submission producing backward[], or -1 */
    const mmr_recon_mpeg2_field_motion *field_motion;   /* [mb_width * mb_height], read for
                                       FIELD_MOTION macroblocks; NULL when there are none
                                       (a FIELD_MOTION macroblock then reports INVALID_DATA). A
                                       struct_size that ends before this member reads as NULL. */
} mmr_recon_mpeg2_picture;

MMR_RECON_API mmr_status mmr_recon_mpeg2_submit(mmr_recon_context *context,
                                                const mmr_recon_mpeg2_picture *pictures,
                                                uint32_t picture_count, uint32_t flags,
                                                uint64_t *out_ticket);

/* Device scratch of the allocation-free entry point below: 64 MiB, held by every GPU the library
 * runs on until the process ends. A submission needs 256-byte-aligned room for its job records and
 * tile table, its counters and status word, and 96 bytes of block order per macroblock (a 1080p
 * picture takes about 0.8 MB, so a closed GOP of 15 1080p pictures fits). */
#define MMR_RECON_NOALLOC_SCRATCH_BYTES (64u << 20)

/*
 * No counterpart in the reference. For callers that cannot allocate device memory (tight serving
 * loops, harnesses that forbid hipMalloc): the same reconstruction as mmr_recon_mpeg2_submit
 * (same descriptors, flags and results), run synchronously on the context's stream with its job
 * records, counters and block order in a fixed device array of MMR_RECON_NOALLOC_SCRATCH_BYTES
 * instead of a per-submission allocation. It allocates no device, pinned or event memory and
 * returns once the pictures are reconstructed, with their status (INVALID_DATA / UNSUPPORTED as
 * mmr_recon_wait reports them). A submission larger than the array fails with
 * MMR_STATUS_INVALID_ARGUMENT before anything runs. Every call on one GPU uses that GPU's one array:
 * calls from one context are serialised, but calls that overlap on one GPU from different contexts
 * race and must be serialised by the caller.
 */
MMR_RECON_API mmr_status mmr_recon_mpeg2_reconstruct_noalloc(mmr_recon_context *context,
                                                             const mmr_recon_mpeg2_picture *pictures,
                                                             uint32_t picture_count, uint32_t flags);

/* -------------------------------------------------------------------- DV */

#define MMR_DV_SYSTEM_525_60 0u  /* 720x480, 4:1:1, 10 DIF sequences */
#define MMR_DV_SYSTEM_625_50 1u  /* 720x576, 4:2:0, 12 DIF sequences */

/*
 * One DV25 frame after VLC decoding and the three-level spill.
 * Replaces decode_segment's reconstruction (crates/codecs/dv/src/decoder.rs:186-189):
 * dequantisation (consume_ac :262-271, inverse_factor_for_encoding :290-310),
 * inverse_88 / inverse_248 (:520-602) and place_macroblock / put_split_chroma
 * (:432-505) at macroblock_coordinates (:395-422).
 *
 * Blocks are in DIF order: sequence, segment slot (27), macroblock (5), block (6).
 * block_info bit 0 = dct_mode (2-4-8), bits 1-2 = class, bits 3-6 = quantisation number.
 */
typedef struct mmr_recon_dv_frame {
    size_t struct_size;
    uint32_t system;              /* MMR_DV_SYSTEM_* */
    const int16_t *levels;        /* [blocks][64] by scan position; [0] = signed 9-bit DC */
    const uint8_t *block_info;    /* [blocks] */
    uint8_t *planes[3];           /* Y 720xH, Cb, Cr tightly packed */
} mmr_recon_dv_frame;

MMR_RECON_API mmr_status mmr_recon_dv_submit(mmr_recon_context *context,
                                             const mmr_recon_dv_frame *frames,
                                             uint32_t frame_count, uint32_t flags,
                                             uint64_t *out_ticket);

/* ----------------------------------------------------------------- H.264 */

#define MMR_H264_MB_INTRA 1u          /* intra macroblock: the picture's intra record says how it is
                                         predicted (deblocking bS 3/4); without intra records UNSUPPORTED */
#define MMR_H264_MB_TRANSFORM_8X8 2u
#define MMR_H264_MB_NO_DEBLOCK 4u     /* disable_deblocking_filter_idc == 1 for this MB's slice */
#define MMR_H264_MB_SLICE_EDGE_OFF 8u /* disable_deblocking_filter_idc == 2: no filtering across slices */

/* One macroblock (inter reconstruction and deblocking inputs). */
typedef struct mmr_recon_h264_macroblock {
    uint8_t flags;           /* MMR_H264_MB_* */
    uint8_t qp;              /* luma QP after mb_qp_delta, 0..51 */
    uint8_t cbp;             /* bits 0-3 luma 8x8 groups, bits 4-5 chroma (0, 1 = DC, 2 = DC+AC) */
    uint8_t reserved;
    uint16_t slice_id;
    int8_t offset_a;         /* FilterOffsetA */
    int8_t offset_b;         /* FilterOffsetB */
    uint32_t coef_offset;    /* index of this MB's 384-level record in levels (if cbp != 0) */
} mmr_recon_h264_macroblock;

/* Motion of one 4x4 luma block (16 per macroblock, H.264 block-index order). */
typedef struct mmr_recon_h264_block {
    int16_t mv[2][2];        /* [L0, L1][x, y], quarter-sample luma units */
    int8_t ref_idx[2];       /* -1: list unused */
    uint8_t total_coeff;     /* nonzero coefficient count (CAVLC total_coeff / CABAC 8x8 count) */
    uint8_t reserved;
} mmr_recon_h264_block;

/* Explicit single-list weight (PredictionWeight, crates/codecs/h264/src/decoder.rs:9884-9907). */
typedef struct mmr_recon_h264_weight {
    int16_t weight;
    int16_t offset;
    uint8_t log2_denom;
    uint8_t reserved[3];
} mmr_recon_h264_weight;

/* Bi-prediction weight pair for (ref_idx L0, ref_idx L1). */
typedef struct mmr_recon_h264_bi_weight {
    int16_t w0;
    int16_t w1;
    int16_t o0;
    int16_t o1;
    uint8_t log2_denom;
    uint8_t reserved[3];
} mmr_recon_h264_bi_weight;

#define MMR_H264_MAX_REFS 16u

typedef struct mmr_recon_h264_reference {
    const uint8_t *planes[3];  /* coded-size reference planes, same geometry as the picture */
    int32_t poc;               /* picture order count (reference identity for bS) */
    int32_t reserved;
} mmr_recon_h264_reference;

/* Level record of one coded MB (384 x int16):
 *   transform 4x4: [0,256)  16 luma blocks x 16 levels, zig-zag scan, block-index order
 *   transform 8x8: [0,256)  4 luma 8x8 blocks x 64 levels, 8x8 zig-zag scan (de-interleaved)
 *   Intra16x16:    [16 b + 1, 16 b + 16) the 15 AC levels of luma block b (scan index 1..15),
 *                  [16 k] the luma DC level at scan index k of the 4x4 DC array
 *   [256,264) chroma DC: Cb c0..c3, Cr c0..c3
 *   [264,384) chroma AC: Cb blocks 0..3, Cr blocks 0..3, 15 levels each (scan index 1..15)
 *   I_PCM: [0,256) luma samples in raster order, [256,320) Cb, [320,384) Cr (values 0..255)
 * Intra16x16 and I_PCM macroblocks always have a record; the others only when cbp != 0. Levels of
 * blocks outside the coded block pattern are zero (every block of a record is transformed). */
#define MMR_H264_LEVELS_PER_MB 384u

#define MMR_H264_INTRA_4X4 0u
#define MMR_H264_INTRA_8X8 1u      /* also needs MMR_H264_MB_TRANSFORM_8X8 */
#define MMR_H264_INTRA_16X16 2u
#define MMR_H264_INTRA_PCM 3u

/*
 * Intra prediction of one MMR_H264_MB_INTRA macroblock (predict_intra4_block, predict_intra8_block,
 * predict_block, predict_chroma_block and their callers, crates/codecs/h264/src/decoder.rs:12224-12330,
 * :12377-12628, :13030-13564, :14499-14712; I_PCM place_pcm_macroblock :10982-11004). Neighbours are
 * available when they are inside the picture and in the same slice (slice_id); the top-right and
 * top-left samples follow the reference's rules. Residuals use the intra scaling lists (4x4 lists 0-2,
 * 8x8 list 0).
 */
typedef struct mmr_recon_h264_intra {
    uint8_t type;            /* MMR_H264_INTRA_* */
    uint8_t luma_mode;       /* Intra16x16: 0 vertical, 1 horizontal, 2 DC, 3 plane */
    uint8_t chroma_mode;     /* 0 DC, 1 horizontal, 2 vertical, 3 plane (not I_PCM) */
    uint8_t reserved;
    uint8_t modes[16];       /* Intra4x4: mode 0..8 of each 4x4 block, block-index order;
                                Intra8x8: modes[0..3] of the 8x8 blocks */
} mmr_recon_h264_intra;

#define MMR_H264_WEIGHTED_NONE 0u      /* default: average / copy */
#define MMR_H264_WEIGHTED_TABLE 1u     /* explicit or implicit: use the weight tables */

/*
 * One frame picture. Replaces H264 picture reconstruction after entropy decoding:
 * predict_inter_partition / predict_bi_inter_partition
 * (crates/codecs/h264/src/decoder.rs:11602-11656, :11901-12119) with luma_qpel,
 * chroma_epel and the weights (:13619-13926), the residual paths
 * add_luma_residual_blocks / _8x8_blocks / add_chroma_residual (:12630-12789), intra
 * prediction of intra macroblocks (mmr_recon_h264_intra), and deblock -> filter_picture
 * (:12791-12848, crates/codecs/h264/src/deblock.rs:84-587).
 */
typedef struct mmr_recon_h264_picture {
    size_t struct_size;
    uint32_t width_mbs;
    uint32_t height_mbs;
    uint32_t weighted_mode;                     /* MMR_H264_WEIGHTED_* */
    uint32_t deblock;                           /* 1: run the loop filter */
    int32_t chroma_qp_offset[2];                /* Cb, Cr */
    const mmr_recon_h264_macroblock *macroblocks;  /* raster order */
    const mmr_recon_h264_block *blocks;         /* [mbs][16] */
    const int16_t *levels;                      /* [coded mbs][384] */
    const mmr_recon_h264_weight *weights;       /* [2][MAX_REFS][3] single-list (Y, Cb, Cr) */
    const mmr_recon_h264_bi_weight *bi_weights; /* [MAX_REFS][MAX_REFS][3] */
    const uint8_t *scaling_lists;               /* NULL: Flat_16; else [6][16] 4x4 then [2][64] 8x8, raster */
    uint8_t *planes[3];                         /* output coded planes (stride 16*width_mbs, 8*width_mbs) */
    uint32_t ref_count[2];                      /* list 0, list 1; {0, 0} for an all-intra (I / IDR) picture */
    mmr_recon_h264_reference refs[2][MMR_H264_MAX_REFS];
    const mmr_recon_h264_intra *intra;          /* [mbs], 4-byte aligned, read for MMR_H264_MB_INTRA macroblocks; NULL:
                                                   intra macroblocks report UNSUPPORTED. A struct_size
                                                   that ends before this member reads as NULL. */
    uint32_t transform_bypass;                  /* qpprime_y_zero_transform_bypass_flag: the 4x4 luma blocks
                                                   (Intra4x4 and inter; not 8x8 or Intra16x16) of macroblocks
                                                   with QP 0, and the chroma blocks of macroblocks with chroma
                                                   QP 0, skip the transform (decoder.rs:12432, :12651,
                                                   :12739-12764). A struct_size that ends before this member
                                                   reads as 0. */
    uint32_t reserved;
} mmr_recon_h264_picture;

MMR_RECON_API mmr_status mmr_recon_h264_submit(mmr_recon_context *context,
                                               const mmr_recon_h264_picture *pictures,
                                               uint32_t picture_count, uint32_t flags,
                                               uint64_t *out_ticket);

/* Implicit bi-prediction weights (implicit_b_prediction_weights,
 * crates/codecs/h264/src/decoder.rs:3577-3620) for one reference pair, host-side:
 * writes the (w0, w1) used with log2_denom 5 and zero offsets. */
MMR_RECON_API mmr_status mmr_recon_h264_implicit_weights(int32_t poc0, int32_t poc1,
                                                         int32_t current_poc,
                                                         int32_t long_term0, int32_t long_term1,
                                                         int32_t *out_w0, int32_t *out_w1);

/* ------------------------------------------------------- colour / blend */

/*
 * YCbCr -> RGB8 (packed, 3 bytes per pixel) at luma resolution.
 * BT.601 replaces ycbcr_to_rgb + the chroma mapping of video_frame_image
 * (crates/cli/src/terminal_preview.rs:7159-7238, :7308-7331).
 * BT.709 has no counterpart in the reference; its integer form is defined here.
 */
typedef struct mmr_recon_yuv_to_rgb_job {
    size_t struct_size;
    uint32_t width;                 /* luma width = output width, 1..65535 */
    uint32_t height;                /* 1..65535 */
    uint32_t chroma_width;          /* 1..width */
    uint32_t chroma_height;         /* 1..height */
    uint32_t luma_stride;
    uint32_t chroma_stride;
    uint32_t range;                 /* MMR_RANGE_* */
    uint32_t matrix;                /* MMR_MATRIX_* */
    const uint8_t *planes[3];
    uint8_t *rgb;                   /* width*height*3 */
    uint8_t *rgba;                  /* optional width*height*4 (alpha 255) instead of / besides rgb */
} mmr_recon_yuv_to_rgb_job;

MMR_RECON_API mmr_status mmr_recon_yuv_to_rgb_submit(mmr_recon_context *context,
                                                     const mmr_recon_yuv_to_rgb_job *jobs,
                                                     uint32_t job_count, uint32_t flags,
                                                     uint64_t *out_ticket);

#define MMR_BLEND_SOURCE_OVER 1u        /* straight-alpha RGBA8 (blend_rgba_pixel) */
#define MMR_BLEND_PREMULTIPLIED_OVER 2u /* premultiplied RGBA8 (LinearPixel::over at 8 bits) */

typedef struct mmr_recon_layer {
    const uint8_t *rgba;   /* RGBA8 layer */
    uint32_t width;
    uint32_t height;
    uint32_t stride;       /* bytes per row */
    uint32_t x;
    uint32_t y;
    uint32_t op;           /* MMR_BLEND_* */
    uint32_t opacity;      /* premultiplied op only: 0..255 */
} mmr_recon_layer;

#define MMR_RECON_MAX_LAYERS 8u

/*
 * Composite layers in order over an RGBA8 target in place.
 * Replaces CpuCompositionBackend's RGBA path (crates/render/src/frame_graph.rs:307-355,
 * blend_rgba8 / blend_rgba_pixel :504-545, :638-657); the premultiplied operator follows
 * LinearPixel::over (crates/mmfx/src/render.rs:1826-1854) with 8-bit channels.
 */
typedef struct mmr_recon_composite_job {
    size_t struct_size;
    uint8_t *target;       /* RGBA8 */
    uint32_t width;
    uint32_t height;
    uint32_t stride;
    uint32_t layer_count;
    mmr_recon_layer layers[MMR_RECON_MAX_LAYERS];
} mmr_recon_composite_job;

MMR_RECON_API mmr_status mmr_recon_composite_submit(mmr_recon_context *context,
                                                    const mmr_recon_composite_job *jobs,
                                                    uint32_t job_count, uint32_t flags,
                                                    uint64_t *out_ticket);

/*
 * Conversion conversions[i] followed by composite composites[i], in one pass: each target pixel is
 * converted, composited and stored once (the converted frame is not written and read back).
 * composites[i].target must be conversions[i].rgba with the same width and height and stride
 * 4 x width, and conversions[i].rgb NULL; the result is that of mmr_recon_yuv_to_rgb_submit then
 * mmr_recon_composite_submit on the same jobs. No counterpart in the reference, which converts
 * a frame (ycbcr_to_rgb, crates/cli/src/terminal_preview.rs:7159-7238) and composites over it
 * (crates/render/src/frame_graph.rs:307-355) as two passes: for callers that do both per frame.
 */
MMR_RECON_API mmr_status mmr_recon_yuv_composite_submit(mmr_recon_context *context,
                                                        const mmr_recon_yuv_to_rgb_job *conversions,
                                                        const mmr_recon_composite_job *composites,
                                                        uint32_t job_count, uint32_t flags,
                                                        uint64_t *out_ticket);

/* -------------------------------------------- MMFX linear light (u16 per channel) */

/*
 * Surfaces of premultiplied linear-light RGBA with one uint16 per channel, in the order red,
 * green, blue, alpha (8 bytes per pixel): MMFX's Surface of LinearPixel
 * (crates/mmfx/src/render.rs:190-280, :1780-1880). Conversions use the reference's
 * srgb_to_linear_u16 / linear_u16_to_srgb, evaluated on the host once per context.
 */
#define MMR_LINEAR_SOURCE_SURFACE 1u  /* a linear u16 surface (uint16[4] per pixel) */
#define MMR_LINEAR_SOURCE_SRGBA8 2u   /* straight-alpha sRGBA8 (uint8[4] per pixel), LinearPixel::from_rgba */

typedef struct mmr_recon_linear_layer {
    const void *pixels;    /* MMR_LINEAR_SOURCE_SURFACE: uint16[4]; MMR_LINEAR_SOURCE_SRGBA8: uint8[4] */
    uint32_t format;       /* MMR_LINEAR_SOURCE_* */
    uint32_t stride;       /* pixels per row */
    int32_t x;             /* target position of the layer's pixel (0, 0) */
    int32_t y;
    uint32_t width;        /* layer size in pixels */
    uint32_t height;
    int32_t clip_left;     /* target-space clip (Clip), right and bottom exclusive */
    int32_t clip_top;
    int32_t clip_right;
    int32_t clip_bottom;
    uint32_t opacity;      /* 0..65535 */
} mmr_recon_linear_layer;

/*
 * Layers in order over a linear target in place: for every target pixel inside the layer's clip,
 * the target and the layer, target = source.over(target, opacity) (LinearPixel::over,
 * crates/mmfx/src/render.rs:1826-1854, with multiply_channel / add_saturating :1847-1857).
 * A surface layer at (0, 0) of the target's size is Surface::blend_surface (:269-277). A sRGBA8
 * layer is execute_draw_image (:1410-1444) for a destination at integer coordinates of the
 * image's own size (its nearest-sample lookup then reads the image 1:1), with the product of
 * the coverage masks (or 65535) as opacity. Sizes up to 65535; at most MMR_RECON_MAX_LAYERS.
 */
typedef struct mmr_recon_linear_composite_job {
    size_t struct_size;
    uint16_t *target;      /* uint16[4] per pixel */
    uint32_t width;
    uint32_t height;
    uint32_t stride;       /* pixels per row */
    uint32_t layer_count;
    mmr_recon_linear_layer layers[MMR_RECON_MAX_LAYERS];
} mmr_recon_linear_composite_job;

MMR_RECON_API mmr_status mmr_recon_linear_composite_submit(mmr_recon_context *context,
                                                           const mmr_recon_linear_composite_job *jobs,
                                                           uint32_t job_count, uint32_t flags,
                                                           uint64_t *out_ticket);

#define MMR_LINEAR_TO_SRGBA8 1u    /* Surface::to_rgba8 (:233-254): unpremultiply, linear_u16_to_srgb */
#define MMR_LINEAR_FROM_SRGBA8 2u  /* LinearPixel::from_rgba (:1796-1808) for every pixel */

typedef struct mmr_recon_linear_convert_job {
    size_t struct_size;
    uint32_t direction;    /* MMR_LINEAR_TO_SRGBA8 / MMR_LINEAR_FROM_SRGBA8 */
    uint32_t width;        /* 1..65535 */
    uint32_t height;       /* 1..65535 */
    uint16_t *linear;      /* uint16[4] per pixel */
    uint32_t linear_stride;  /* pixels per row */
    uint32_t rgba_stride;    /* bytes per row */
    uint8_t *rgba;         /* straight-alpha sRGBA8 */
} mmr_recon_linear_convert_job;

MMR_RECON_API mmr_status mmr_recon_linear_convert_submit(mmr_recon_context *context,
                                                         const mmr_recon_linear_convert_job *jobs,
                                                         uint32_t job_count, uint32_t flags,
                                                         uint64_t *out_ticket);

/*
 * Blend a preconverted YUV 4:2:0 overlay with alpha into a YUV 4:2:0 target.
 * Replaces blend_yuv420 / blend_channel (crates/render/src/frame_graph.rs:548-671).
 */
typedef struct mmr_recon_yuv_blend_job {
    size_t struct_size;
    uint8_t *target[3];
    uint32_t target_width;
    uint32_t target_height;
    uint32_t target_stride[3];
    const uint8_t *overlay[3];     /* Y (width x height), U, V (width/2 x height/2), range already chosen */
    const uint8_t *rgba;           /* straight alpha in byte 3 of each RGBA pixel (width x height) */
    const uint8_t *chroma_alpha;   /* 2x2-averaged alpha (width/2 x height/2) */
    uint32_t width;
    uint32_t height;
    uint32_t x;
    uint32_t y;
} mmr_recon_yuv_blend_job;

MMR_RECON_API mmr_status mmr_recon_yuv_blend_submit(mmr_recon_context *context,
                                                    const mmr_recon_yuv_blend_job *jobs,
                                                    uint32_t job_count, uint32_t flags,
                                                    uint64_t *out_ticket);

/*
 * Overlay preparation for the YUV 4:2:0 blend. Replaces PreparedOverlay::prepare_yuv +
 * rgb_to_yuv (crates/render/src/project_compositor.rs:1098-1185): from a straight-alpha RGBA8
 * overlay (the cropped region from_canvas keeps, :1046-1096), BT.709 limited and full-range luma
 * per pixel and, per 2x2 cell (the last row / column of an odd size on its own), the rounded mean
 * alpha and the alpha-weighted mean U / V of both ranges (128 where the cell's alpha sums to 0).
 * Outputs: y_* width x height; u_* / v_* / chroma_alpha ceil(width/2) x ceil(height/2), all
 * tightly packed; any output may be NULL (not written).
 */
typedef struct mmr_recon_overlay_prepare_job {
    size_t struct_size;
    const uint8_t *rgba;
    uint32_t width;
    uint32_t height;
    uint8_t *y_limited;
    uint8_t *y_full;
    uint8_t *u_limited;
    uint8_t *v_limited;
    uint8_t *u_full;
    uint8_t *v_full;
    uint8_t *chroma_alpha;
} mmr_recon_overlay_prepare_job;

MMR_RECON_API mmr_status mmr_recon_overlay_prepare_submit(mmr_recon_context *context,
                                                          const mmr_recon_overlay_prepare_job *jobs,
                                                          uint32_t job_count, uint32_t flags,
                                                          uint64_t *out_ticket);

/* Scale filters (FrameScaleFilter): image::imageops::FilterType::Lanczos3 / Triangle. */
#define MMR_SCALE_FILTER_LANCZOS3 0u
#define MMR_SCALE_FILTER_TRIANGLE 1u

/* Visual scale modes (VisualScaleMode). */
#define MMR_SCALE_FIT 0u       /* preserve aspect, pad */
#define MMR_SCALE_FILL 1u      /* preserve aspect, crop */
#define MMR_SCALE_STRETCH 2u   /* whole canvas */
#define MMR_SCALE_NATIVE 3u    /* no resize, centred crop / pad */

/*
 * A planar 4:2:0 frame mapped onto an even-sized canvas. Replaces scale_yuv420_to_canvas_cpu
 * (crates/render/src/compositor.rs:89-141) with scaled_dimensions / scaled_even (:211-254) and
 * center_crop_or_pad (:277-305): each plane is resized with the Lanczos3 filter of
 * image::imageops::resize (the image crate 0.25.6: a vertical pass into f32, then a horizontal pass
 * rounding to u8, weights per output coordinate normalised to sum 1) with the Lanczos3 or Triangle
 * filter (FrameScaleFilter, crates/render/src/frame_graph.rs:219-224) unless its scaled size equals
 * its size, then centre-cropped / padded (luma 16 for MMR_RANGE_LIMITED else 0, chroma 128).
 * Width, height and the canvas size are even and non-zero; chroma planes are half size.
 */
typedef struct mmr_recon_scale_job {
    size_t struct_size;
    const uint8_t *src[3];
    uint32_t src_stride[3];
    uint32_t width;
    uint32_t height;
    uint8_t *dst[3];
    uint32_t dst_stride[3];
    uint32_t canvas_width;
    uint32_t canvas_height;
    uint32_t mode;      /* MMR_SCALE_* */
    uint32_t range;     /* MMR_RANGE_* of the frame (padding black) */
    uint32_t filter;    /* MMR_SCALE_FILTER_* */
} mmr_recon_scale_job;

MMR_RECON_API mmr_status mmr_recon_scale_submit(mmr_recon_context *context, const mmr_recon_scale_job *jobs,
                                                uint32_t job_count, uint32_t flags, uint64_t *out_ticket);

/* Codec selectors of the prepared-batch and serving calls. */
#define MMR_CODEC_MJPEG 1u
#define MMR_CODEC_MPEG2 2u
#define MMR_CODEC_DV 3u
#define MMR_CODEC_H264 4u

/* ------------------------------------------------------ prepared batches */

/*
 * New (no counterpart in the reference, whose reconstruction loops run per frame on the CPU): a
 * batch prepared once and replayed. mmr_recon_prepare uploads the descriptors' job records once
 * into a persistent device block and captures the batch's kernel launches in a HIP graph;
 * mmr_recon_prepared_launch replays that graph on the context's stream and returns a ticket like a
 * submit call. The device buffers the descriptors name must stay allocated at the same addresses
 * while the prepared batch lives (their contents may change between launches: a replay reads the
 * inputs anew). Launches of one prepared batch run in order on the context's stream.
 * codec: MMR_CODEC_MJPEG / MPEG2 / DV / H264; descriptors: an array of that codec's frame /
 * picture descriptors; flags as for the codec's submit call.
 */
typedef struct mmr_recon_prepared mmr_recon_prepared;

MMR_RECON_API mmr_status mmr_recon_prepare(mmr_recon_context *context, uint32_t codec,
                                           const void *descriptors, uint32_t count, uint32_t flags,
                                           mmr_recon_prepared **out_prepared);
MMR_RECON_API mmr_status mmr_recon_prepared_launch(mmr_recon_context *context,
                                                   mmr_recon_prepared *prepared,
                                                   uint64_t *out_ticket);
/* Waits for the context's stream, then releases the graph and the device block. */
MMR_RECON_API void mmr_recon_prepared_destroy(mmr_recon_context *context,
                                              mmr_recon_prepared *prepared);

/* ---------------------------------------------------------------- tuning */

/*
 * New (no counterpart in the reference, whose reconstruction runs on the CPU): the launch shapes
 * of the device kernels. Every setting reconstructs the same bytes; settings differ in speed
 * only. A context starts from the defaults below; mmr_recon_set_tuning replaces them (for the
 * submissions, prepared batches and server batches made afterwards), and mmr_recon_autotune
 * times the candidates of one codec on a caller's batch and keeps the fastest.
 */
typedef struct mmr_recon_tuning {
    size_t struct_size;
    uint32_t mjpeg_group_blocks;         /* 8x8 blocks per workgroup: 16, 32 (default) or 64 */
    uint32_t mpeg2_group_blocks;         /* same for MPEG-2 (default 32) */
    uint32_t dv_group_blocks;            /* same for DV (default 64) */
    uint32_t h264_deblock_band_rows;     /* macroblock rows per deblocking band, one wave each:
                                            4, 8 (default) or 16 */
    uint32_t h264_deblock_groups_per_cu; /* persistent deblocking workgroups per compute unit: 1..4
                                            (default 2) */
    uint32_t h264_deblock_publish;       /* macroblocks between two progress publications of a
                                            band's last row to the next band: 1..32 (default 2) */
} mmr_recon_tuning;

MMR_RECON_API mmr_status mmr_recon_get_tuning(mmr_recon_context *context, mmr_recon_tuning *out_tuning);
/* MMR_STATUS_INVALID_ARGUMENT for a value outside the ranges above (the tuning is unchanged). */
MMR_RECON_API mmr_status mmr_recon_set_tuning(mmr_recon_context *context, const mmr_recon_tuning *tuning);

typedef struct mmr_recon_tuning_trial {
    mmr_recon_tuning tuning;
    double kernel_ms;                    /* median device time of the batch's kernels */
} mmr_recon_tuning_trial;

/*
 * Reconstructs the batch (descriptors and flags as for the codec's submit call) with every
 * candidate setting of that codec's fields (the other fields keep their current values): once
 * untimed, then `repeats` timed times (0: 3). Keeps the candidate with the smallest median kernel
 * time in the context and copies it to out_best (optional). trials (optional, trial_capacity
 * entries) receives every candidate with its time, out_trials (optional) their number. Blocks
 * until done; the outputs then hold the batch's reconstruction. A failing run returns its status
 * and leaves the tuning as it was.
 */
MMR_RECON_API mmr_status mmr_recon_autotune(mmr_recon_context *context, uint32_t codec,
                                            const void *descriptors, uint32_t count, uint32_t flags,
                                            uint32_t repeats, mmr_recon_tuning *out_best,
                                            mmr_recon_tuning_trial *trials, uint32_t trial_capacity,
                                            uint32_t *out_trials);

/* ------------------------------------------------- resident frame cache */

/*
 * Library-owned, byte-budgeted LRU cache of device frame storage with generation protection.
 * Replaces DeviceResourceCache (crates/render/src/device_cache.rs:53-287): a semantic key
 * (FrameResourceKey, crates/render/src/frame_graph.rs:28-43) maps to one device allocation
 * that the cache creates on the first retain and reuses afterwards. Entries retained or fetched
 * in the current generation are protected: a retain that would have to evict one of them fails
 * with MMR_STATUS_INVALID_STATE and leaves the cache within budget. Evicted and removed storage
 * is released stream-ordered on the context's stream, so submissions of the context enqueued
 * before the release finish with it; work on other streams must be complete before the call
 * that evicts. Calls on one cache are serialised by the context.
 */
typedef struct mmr_recon_frame_cache mmr_recon_frame_cache;

typedef struct mmr_recon_frame_key {
    uint32_t namespace_id;         /* FrameResourceNamespace */
    uint32_t variant;              /* backend-independent variant discriminator */
    uint64_t owner;                /* stable media identifier, 0 for caller-owned targets */
    uint64_t revision;             /* deterministic content revision */
    int64_t local_frame;           /* source-local frame, -1 for a time-invariant resource */
    uint32_t width;
    uint32_t height;
} mmr_recon_frame_key;

typedef struct mmr_recon_frame_descriptor {   /* FrameDescriptor: compared on every hit */
    uint32_t width;
    uint32_t height;
    uint32_t format;
    uint32_t color;
    uint32_t alpha_mode;
} mmr_recon_frame_descriptor;

#define MMR_CACHE_INSERTED 1u
#define MMR_CACHE_REUSED 2u

typedef struct mmr_recon_frame_cache_stats {
    size_t struct_size;
    uint64_t resources;            /* retained entries */
    uint64_t retained_bytes;
    uint64_t byte_budget;
    uint64_t reuses;
    uint64_t insertions;
    uint64_t evictions;
    uint64_t generation;
} mmr_recon_frame_cache_stats;

/* byte_budget > 0. */
MMR_RECON_API mmr_status mmr_recon_frame_cache_create(mmr_recon_context *context, uint64_t byte_budget,
                                                      mmr_recon_frame_cache **out_cache);
MMR_RECON_API void mmr_recon_frame_cache_destroy(mmr_recon_frame_cache *cache);
/* begin_frame: the next generation; what is retained or fetched after it is protected until the next. */
MMR_RECON_API mmr_status mmr_recon_frame_cache_begin_frame(mmr_recon_frame_cache *cache);
/* retain_with: the entry's storage (reused: *out_status MMR_CACHE_REUSED) or a new device allocation
 * of `bytes` (MMR_CACHE_INSERTED) after evicting least recently used entries of earlier
 * generations. Errors: a descriptor different from the retained one under the same key
 * (INVALID_STATE), bytes 0 or above the budget (INVALID_DATA), not enough evictable bytes
 * (INVALID_STATE, nothing inserted). */
MMR_RECON_API mmr_status mmr_recon_frame_cache_retain(mmr_recon_frame_cache *cache, const mmr_recon_frame_key *key,
                                                      const mmr_recon_frame_descriptor *descriptor, uint64_t bytes,
                                                      void **out_device_ptr, uint32_t *out_status);
/* get: the entry's storage marked used in this generation, or *out_device_ptr = NULL when absent. */
MMR_RECON_API mmr_status mmr_recon_frame_cache_get(mmr_recon_frame_cache *cache, const mmr_recon_frame_key *key,
                                                   const mmr_recon_frame_descriptor *descriptor, void **out_device_ptr);
/* remove: release one entry (*out_removed 1) or nothing (0). */
MMR_RECON_API mmr_status mmr_recon_frame_cache_remove(mmr_recon_frame_cache *cache, const mmr_recon_frame_key *key,
                                                      int32_t *out_removed);
/* release_idle: release entries idle for more than max_idle_generations completed generations
 * (never those of the current generation); *out_released receives their number. */
MMR_RECON_API mmr_status mmr_recon_frame_cache_release_idle(mmr_recon_frame_cache *cache, uint64_t max_idle_generations,
                                                            uint64_t *out_released);
MMR_RECON_API mmr_status mmr_recon_frame_cache_clear(mmr_recon_frame_cache *cache);
MMR_RECON_API mmr_status mmr_recon_frame_cache_stats_get(mmr_recon_frame_cache *cache,
                                                         mmr_recon_frame_cache_stats *out_stats);

/* --------------------------------------------------------------- serving */

/*
 * Continuous-batching request loop. Replaces the bounded worker pool + request / event pattern
 * of the playback sources (DecodeExecutor, crates/playback/src/executor.rs:40-293;
 * Mpeg2PlaybackSource / H264PlaybackSource request handling, crates/playback/src/mpeg2.rs:151-184,
 * crates/playback/src/h264.rs:273-317) for device reconstruction.
 *
 * A request is a window of frames of one codec on one stream, optionally followed per frame by
 * a YUV->RGB conversion and an RGBA composite. Requests are admitted while the sum of admitted
 * requests' hbm_bytes stays within the budget; a scheduler thread keeps up to `slots` batches in
 * flight on separate HIP streams and fills each new batch with the next frames of every admitted
 * request (continuous batching), so requests that arrive while others run join the next batch.
 * Frames of one request complete in order; batches that carry frames of the same request are
 * ordered on the device by events.
 */
typedef struct mmr_recon_server mmr_recon_server;


typedef struct mmr_recon_server_config {
    size_t struct_size;
    uint64_t hbm_budget_bytes;     /* admission budget (0: context's budget or unlimited) */
    uint32_t slots;                /* batches in flight (HIP streams), 0: 4 */
    uint32_t max_batch_frames;     /* frames per batch and codec, 0: 64 */
    uint32_t max_frames_per_request_per_batch; /* 0: 8 */
    uint32_t max_pending;          /* queued (not yet admitted) requests before BUSY, 0: 4096 */
} mmr_recon_server_config;

typedef struct mmr_recon_request {
    size_t struct_size;
    uint32_t codec;                 /* MMR_CODEC_* */
    uint32_t stream_id;             /* requests of one stream (nonzero id) run in enqueue order:
                                       their frames may use references left resident by earlier
                                       requests of the stream; 0 = independent request */
    uint32_t frame_count;
    const void *frames;             /* frame_count codec descriptors (mmr_recon_*_frame / _picture);
                                       MPEG-2 forward_source / backward_source index this array */
    const mmr_recon_yuv_to_rgb_job *convert;     /* NULL or frame_count jobs */
    const mmr_recon_composite_job *composite;    /* NULL or frame_count jobs (after convert) */
    uint64_t hbm_bytes;             /* device memory the request holds while admitted; admission
                                       counts at least the bytes of the planes, RGB / RGBA images and
                                       composite targets its descriptors name */
    uint64_t user_tag;
} mmr_recon_request;

typedef struct mmr_recon_frame_completion {
    uint64_t request_id;
    uint64_t user_tag;
    uint32_t frame_index;
    mmr_status status;
    double latency_ms;              /* enqueue -> observed device completion of this frame */
} mmr_recon_frame_completion;

typedef struct mmr_recon_server_stats {
    size_t struct_size;
    uint64_t requests_enqueued;
    uint64_t requests_completed;
    uint64_t frames_completed;
    uint64_t batches_launched;
    uint64_t admitted_bytes;        /* currently admitted */
    uint64_t peak_admitted_bytes;
    uint32_t pending;               /* waiting for admission */
    uint32_t active;                /* admitted, not complete */
} mmr_recon_server_stats;

/* New: no counterpart in the reference (its executor runs decode closures on CPU threads). */
MMR_RECON_API mmr_status mmr_recon_server_create(mmr_recon_context *context,
                                                 const mmr_recon_server_config *config,
                                                 mmr_recon_server **out_server);
/* Non-blocking; copies the descriptors. MMR_STATUS_BUSY when max_pending requests wait. */
MMR_RECON_API mmr_status mmr_recon_server_enqueue(mmr_recon_server *server,
                                                  const mmr_recon_request *request,
                                                  uint64_t *out_request_id);
/* Moves up to capacity frame completions (in completion order) into out. */
MMR_RECON_API mmr_status mmr_recon_server_poll(mmr_recon_server *server,
                                               mmr_recon_frame_completion *out, uint32_t capacity,
                                               uint32_t *out_count);
/* Blocks until every enqueued request has completed. */
MMR_RECON_API mmr_status mmr_recon_server_drain(mmr_recon_server *server);
MMR_RECON_API mmr_status mmr_recon_server_stats_get(mmr_recon_server *server,
                                                    mmr_recon_server_stats *out);
MMR_RECON_API void mmr_recon_server_destroy(mmr_recon_server *server);

#ifdef __cplusplus
}
#endif

#endif /* MMR_RECON_H */
