/*
 * hevcr_syntax.h — per-frame reconstruction syntax handed from the host
 * (CABAC / slice parsing) to the reconstruction engine.
 *
 * This is the payload the reference's reconstruction functions read from
 * its parser state, flattened into plain C structs:
 *   - CU records  <- CUInfo grid + intra mode grids
 *                    (src/decoding/coding_tree.h:21-30, 57-59)
 *   - PU records  <- PUMotionInfo per PU (src/decoding/inter_prediction.h:16-20),
 *                    stored per PU instead of per 4x4 (store_pu_motion,
 *                    src/decoding/inter_prediction.cpp:17-27 expands it)
 *   - TU records  <- the transform-tree leaves visited by decode_transform_unit
 *                    (src/decoding/coding_tree.cpp:1256-1419), with the int16
 *                    coefficient blocks written by decode_residual_coding
 *                    (src/decoding/residual_coding.cpp:438: row-major c[y*N+x])
 *   - SAO records <- DecodingContext::SaoParams (src/decoding/coding_tree.h:73-80)
 *   - scaling     <- the active ScalingListData (src/syntax/sps.h:19-25)
 *   - frame params <- the SPS/PPS/slice-header fields the hot path reads
 *                    (src/syntax/sps.h:69-161, pps.h:13-71, slice_header.h:16-108)
 *
 * All records are listed in decoding order: CUs CTB by CTB in the picture's
 * CTB scan (raster order, or tile by tile with tiles: CtbAddrRsToTs,
 * src/syntax/pps.cpp derive_tile_scan) and in z-order inside each CTB, each
 * CU's PUs in partIdx order, each CU's TUs in transform-tree z-order.
 * The CUs tile the picture without overlap (its width and height are
 * multiples of the minimum CB size, as H.265 requires), the TUs of an intra
 * CU tile that CU, and no two TUs overlap; the engine checks this from the
 * records' order and sizes and rejects other frames with HEVCR_EINVAL.
 * Pictures of several slice segments (hevcr_slice, one record per segment
 * and a slice-segment index per CTB), tiles (hevcr_tiles), PCM and
 * transquant-bypass CUs (hevcr_cu.flags) are representable; only 4:2:0 is.
 */
#ifndef HEVCR_SYNTAX_H
#define HEVCR_SYNTAX_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Slice types, src/common/types.h:104-108 */
#define HEVCR_SLICE_B 0
#define HEVCR_SLICE_P 1
#define HEVCR_SLICE_I 2

/* Prediction modes, src/common/types.h:111-115 */
#define HEVCR_MODE_INTER 0
#define HEVCR_MODE_INTRA 1
#define HEVCR_MODE_SKIP  2

/* Partition modes, src/common/types.h:118-127 */
#define HEVCR_PART_2Nx2N 0
#define HEVCR_PART_2NxN  1
#define HEVCR_PART_Nx2N  2
#define HEVCR_PART_NxN   3
#define HEVCR_PART_2NxnU 4
#define HEVCR_PART_2NxnD 5
#define HEVCR_PART_nLx2N 6
#define HEVCR_PART_nRx2N 7

/* TU cbf bits */
#define HEVCR_CBF_Y  1u
#define HEVCR_CBF_CB 2u
#define HEVCR_CBF_CR 4u

/* PU prediction-list flags */
#define HEVCR_PRED_L0 1u
#define HEVCR_PRED_L1 2u

#define HEVCR_MAX_REFS 16
#define HEVCR_MAX_SLICES 256     /* slice segments per picture (the reference keeps a uint8 index per CTB) */
#define HEVCR_MAX_TILE_COLS 20   /* num_tile_columns_minus1 + 1 (H.265 Table A.8, level 6.2) */
#define HEVCR_MAX_TILE_ROWS 22

/* CU flags */
#define HEVCR_CU_PCM    1u  /* pcm_flag: intra 2Nx2N CU whose samples are coded directly */
#define HEVCR_CU_BYPASS 2u  /* cu_transquant_bypass_flag: residual = coefficient levels */

/* One coding unit (16 bytes). */
typedef struct hevcr_cu {
    uint16_t x, y;          /* luma position of the CU */
    uint8_t  log2_size;     /* 3..6 */
    uint8_t  pred_mode;     /* HEVCR_MODE_* */
    uint8_t  part_mode;     /* HEVCR_PART_* */
    int8_t   qp_y;          /* QpY of the CU, without QpBdOffsetY */
    uint8_t  intra_luma[4]; /* IntraPredModeY per PU (NxN: 4 PUs in z-order, else [0]) */
    uint8_t  intra_chroma;  /* IntraPredModeC (derived mode 0..34) */
    uint8_t  num_pu;        /* inter: number of PU records of this CU */
    uint16_t flags;         /* HEVCR_CU_PCM | HEVCR_CU_BYPASS */
} hevcr_cu;

/* One inter prediction unit (20 bytes). */
typedef struct hevcr_pu {
    uint16_t x, y, w, h;    /* luma position and size */
    int16_t  mv[2][2];      /* [list][x,y], quarter-pel */
    int8_t   ref_idx[2];    /* index into ref list 0/1, -1 if unused */
    uint8_t  pred_flags;    /* HEVCR_PRED_L0 | HEVCR_PRED_L1 */
    uint8_t  reserved;
} hevcr_pu;

/* One transform-tree leaf (28 bytes). A PCM CU has exactly one record, of
 * the CU's size with cbf = 7, whose three blocks hold the CU's samples
 * instead of coefficient levels: the values the reference writes to the
 * picture (decode_pcm_samples, coding_tree.cpp:1425-1460 — the pcm_sample
 * values as read, without the left shift to the bit depth), dense NxN (luma)
 * and (N/2)x(N/2) (chroma) or as compact lists. */
typedef struct hevcr_tu {
    uint16_t x, y;          /* luma position (x0, y0) */
    uint16_t xbase, ybase;  /* parent position; chroma of a 4x4 luma quad sits there */
    uint8_t  log2_size;     /* luma transform size 2..5 */
    uint8_t  blk_idx;       /* 0..3 within the parent */
    uint8_t  cbf;           /* HEVCR_CBF_* */
    uint8_t  tskip;         /* transform_skip_flag per component (bit c); must be 0 */
    uint32_t cu_idx;        /* owning CU record */
    uint32_t coeff_off[3];  /* per component: offset (int16 elements) into the
                               coefficient pool of its dense NxN block, or
                               (sig set) index of its compact list in sig */
} hevcr_tu;

/* SAO parameters of one CTB (24 bytes). */
typedef struct hevcr_sao {
    uint8_t type[3];        /* 0 off, 1 band, 2 edge (SaoTypeIdx) */
    uint8_t band_pos[3];    /* sao_band_position */
    uint8_t eo_class[3];    /* SaoEoClass */
    int8_t  offset[3][5];   /* SaoOffsetVal per category, already signed */
} hevcr_sao;

/* Explicit weighted-prediction entry of one reference index. */
typedef struct hevcr_weight {
    int16_t w[3];           /* LumaWeightLX, ChromaWeightLX[Cb], [Cr] */
    int16_t o[3];           /* offsets in 8-bit units (scaled by << (bd-8) in use) */
} hevcr_weight;

/* Scaling factors in force when sps.scaling_list_enabled_flag is set
 * (ScalingListData, src/syntax/sps.h:19-25; the PPS's lists when it carries
 * them, else the SPS's). perform_dequant (src/decoding/transform.cpp:308-356)
 * takes the factor m of coefficient (x, y) of an NxN block of component c as
 *   matrixId = c + 3 * (CU not intra), sizeId = log2(N) - 2,
 *   m = list[0][matrixId][4y + x]                      for N = 4,
 *   m = list[sizeId][matrixId][8 (y / r) + x / r], r = N / 8, otherwise,
 *       except m = dc[sizeId - 2][matrixId] (0 reads as 16) at (0, 0) for N >= 16,
 * i.e. list[][][] is indexed by raster position in the 4x4 / 8x8 base matrix. */
typedef struct hevcr_scaling {
    uint8_t list[4][6][64];
    uint8_t dc[2][6];
} hevcr_scaling;

typedef struct hevcr_frame_params {
    int32_t width, height;          /* luma samples, multiples of 8 */
    int32_t bit_depth;              /* 8..10, luma == chroma */
    int32_t ctb_log2;               /* 4..6 */
    int32_t min_cb_log2;            /* 3..ctb_log2 (MinCbLog2SizeY; intra NxN only at this size) */
    int32_t slice_type;             /* HEVCR_SLICE_* */
    int32_t poc;
    int32_t strong_intra_smoothing; /* sps.strong_intra_smoothing_enabled_flag */
    int32_t intra_smoothing_disabled;
    int32_t sao_enabled;            /* sps.sample_adaptive_offset_enabled_flag */
    int32_t pps_cb_qp_offset, pps_cr_qp_offset;
    int32_t slice_cb_qp_offset, slice_cr_qp_offset;
    int32_t deblocking_disabled;    /* slice_deblocking_filter_disabled_flag */
    int32_t beta_offset_div2, tc_offset_div2;
    int32_t weighted_pred;          /* explicit WP in force for this slice */
    int32_t luma_log2_wd;           /* luma_log2_weight_denom */
    int32_t chroma_log2_wd;         /* luma denom + delta_chroma_log2_weight_denom */
    int32_t num_ref[2];             /* active entries of RefPicList0/1 */
    int32_t ref_slot[2][HEVCR_MAX_REFS]; /* picture slot each list entry points at */
    int32_t ref_poc[2][HEVCR_MAX_REFS];  /* POC of that picture */
    hevcr_weight weights[2][HEVCR_MAX_REFS];
} hevcr_frame_params;

/* Parameters of one slice segment of a picture of several
 * (src/syntax/slice_header.h:16-108). The reference builds the picture's
 * reference lists once, from its first slice segment (decoder.cpp:108-115),
 * and every segment predicts from those lists: they are hevcr_frame_params'
 * num_ref / ref_slot / ref_poc. A dependent slice segment carries its
 * independent segment's values (decoder.cpp:203-211); the reference still
 * counts it as a slice of its own wherever it compares slice indices
 * (intra and filter boundaries). */
typedef struct hevcr_slice {
    int32_t slice_type;             /* HEVCR_SLICE_*; PUs use the lists of this type */
    int32_t slice_cb_qp_offset, slice_cr_qp_offset;
    int32_t deblocking_disabled;    /* slice_deblocking_filter_disabled_flag */
    int32_t beta_offset_div2, tc_offset_div2;
    int32_t loop_filter_across_slices; /* slice_loop_filter_across_slices_enabled_flag */
    int32_t weighted_pred;          /* explicit WP in force for this segment */
    int32_t luma_log2_wd, chroma_log2_wd;
    hevcr_weight weights[2][HEVCR_MAX_REFS];
} hevcr_slice;

/* Tile partitioning of the picture (pps.h:13-71): column widths and row
 * heights in CTBs, left to right and top to bottom. */
typedef struct hevcr_tiles {
    int32_t  cols, rows;            /* num_tile_columns, num_tile_rows */
    uint16_t col_width[HEVCR_MAX_TILE_COLS];
    uint16_t row_height[HEVCR_MAX_TILE_ROWS];
    int32_t  loop_filter_across_tiles; /* pps loop_filter_across_tiles_enabled_flag */
} hevcr_tiles;

typedef struct hevcr_frame_syntax {
    hevcr_frame_params p;
    const hevcr_cu  *cus;    uint32_t num_cus;
    const hevcr_pu  *pus;    uint32_t num_pus;
    const hevcr_tu  *tus;    uint32_t num_tus;
    const int16_t   *coeffs; uint64_t num_coeffs;
    const hevcr_sao *sao;    /* PicWidthInCtbs * PicHeightInCtbs records, raster order */
    const hevcr_scaling *scaling; /* NULL: flat factors, m = 16 (scaling_list_enabled_flag 0) */
    /* Optional compact coefficients (NULL: the dense pool above is used).
     * When set, coeff_off[c] of a coded block is the index in sig of a header
     * word holding the number k (0..N*N) of the block's nonzero coefficients,
     * followed by k words (pos | (uint32_t)(uint16_t)level << 16) with
     * pos = y * N + x, in increasing order of pos; coeffs and num_coeffs are
     * not read. The reference's residual coding writes the
     * same coefficients into a dense block (residual_coding.cpp:438); a
     * parser emits them here as it decodes them, without the zeros. The
     * engine checks each list's count and extent, not its positions: a
     * position of N*N or more is taken modulo N*N, and of two words with one
     * position the later one holds (a producer that breaks the order gets a
     * wrong block, never an access outside it). */
    const uint32_t  *sig;    uint64_t num_sig;
    /* Slice segments (NULL: the picture is one slice segment, p's). With
     * segments, p's slice-level fields other than slice_type and the lists
     * are not read; ctb_slice[a] is the segment holding the CTB of raster
     * address a (the reference's slice_idx grid, coding_tree.cpp:609-611),
     * segments covering consecutive runs of the CTB scan, 0 first. */
    const hevcr_slice *slices; uint32_t num_slices;
    const uint8_t   *ctb_slice;
    const hevcr_tiles *tiles;        /* NULL: one tile */
    int32_t pcm_loop_filter_disabled; /* sps.pcm_loop_filter_disabled_flag */
} hevcr_frame_syntax;

#ifdef __cplusplus
}
#endif

#endif /* HEVCR_SYNTAX_H */
