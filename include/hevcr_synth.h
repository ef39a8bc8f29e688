/*
 * hevcr_synth.h — deterministic, seeded generator of per-frame
 * reconstruction syntax and reference pictures.
 *
 * The reference decoder has no syntax-level input other than a bitstream, so
 * the workloads (one intra 416x240 frame, 1080p P/B frames, 4K Main10 frames,
 * mixed serving streams) are produced here from an explicit 64-bit seed with
 * integer-only arithmetic (SplitMix64 -> xoshiro256**, integer Bernoulli
 * thresholds, geometric levels, Irwin-Hall Gaussians, a fixed-point sine
 * table): the same seed yields the same bytes on every host.
 */
#ifndef HEVCR_SYNTH_H
#define HEVCR_SYNTH_H

#include "hevcr_syntax.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Probabilities are fixed point with 16 fractional bits (65536 == 1.0). */
typedef struct hevcr_synth_cfg {
    int32_t  width, height, bit_depth;
    int32_t  ctb_log2;
    int32_t  slice_type;        /* HEVCR_SLICE_* */
    int32_t  poc;
    int32_t  qp_min, qp_max;    /* QpY range, uniform per CU */
    int32_t  level_mul;         /* coefficient level multiplier */
    uint32_t p_cu_split;        /* CU quadtree split probability per level */
    uint32_t p_tu_split;        /* TU quadtree split probability per level */
    uint32_t p_intra_nxn;       /* 8x8 intra CU uses PART_NxN */
    uint32_t p_cbf_luma;        /* coded luma block */
    uint32_t p_cbf_chroma;      /* coded chroma block */
    uint32_t p_intra_cu;        /* intra CU inside a P/B frame */
    uint32_t p_skip;            /* skip CU inside a P/B frame */
    uint32_t p_root_cbf_zero;   /* inter CU without residual (rqt_root_cbf = 0) */
    uint32_t p_bi;              /* bi-predicted PU (B slices) */
    uint32_t p_wp;              /* PU points at an explicitly weighted list entry */
    int32_t  mv_sigma_qpel;     /* Gaussian MV standard deviation, quarter-pel */
    int32_t  mv_clamp_qpel;     /* |mv| clamp, quarter-pel */
    int32_t  num_refs;          /* distinct reference pictures, 1..8 */
    int32_t  num_refs_before;   /* of which POC < current (rest POC > current) */
    int32_t  weighted;          /* explicit WP: list entries cycle twice, second half weighted */
    int32_t  amp;               /* allow asymmetric partitions */
    int32_t  sao_enabled;
    uint32_t p_sao_off;         /* a CTB component without SAO */
    int32_t  deblocking_disabled;
    int32_t  strong_intra_smoothing;
    int32_t  scaling_lists;     /* 1: seeded scaling factors (1..255, DC included), 0: flat */
    /* Slice segments and tiles (0 or 1: one). Segment boundaries fall at
     * seeded points of the CTB scan; a segment after the first is dependent
     * (the previous independent segment's parameters) with probability
     * p_dependent, else independent with seeded parameters: slice type (I
     * or P inside B pictures, I inside P pictures, with p_slice_type_change),
     * slice chroma QP offsets, deblocking on/off, beta/tc offsets,
     * slice_loop_filter_across_slices_enabled_flag and weights. Tile column
     * widths and row heights are seeded (each at least one CTB). */
    int32_t  num_slices;
    uint32_t p_dependent;
    uint32_t p_slice_type_change;
    int32_t  tile_cols, tile_rows;
    int32_t  loop_filter_across_tiles;
    /* PCM: an intra 2Nx2N CU of 8x8..32x32 is PCM with probability p_pcm,
     * its samples uniform over pcm_bits bits; pcm_loop_filter_disabled as
     * the SPS flag. Transquant bypass: a CU has it with probability p_bypass. */
    uint32_t p_pcm;
    int32_t  pcm_bits;
    int32_t  pcm_loop_filter_disabled;
    uint32_t p_bypass;
} hevcr_synth_cfg;

typedef struct hevcr_synth_frame hevcr_synth_frame;

/* Workload presets:
 *   0: 416x240 8-bit intra frame (CTB 64, CU split 0.5/level to 8x8, TU 4..32, QP 22..37,
 *      deblocking and random per-CTB SAO on)
 *   1: 1920x1080 8-bit B frame (PUs 8..64 with AMP, 60% uni / 40% bi, MV sigma 8 px clamp
 *      128 px, 20% explicitly weighted, 30% of residual-bearing CUs CBF-zero, QP 27..37)
 *   2: 1920x1080 8-bit intra frame with config-0 syntax
 *   3: 3840x2160 10-bit B frame (levels x4, QP 17..37, MV sigma 24 px clamp 256 px)
 *   4: 3840x2160 10-bit intra frame (levels x4, QP 17..37)
 *   5: 1920x1080 8-bit P frame (serving GOP member, uni-pred only)
 * Returns 0, or -1 for an unknown preset. */
int hevcr_synth_preset(int preset, hevcr_synth_cfg *cfg);

/* Generates one frame of syntax. Returns NULL on invalid configuration. */
hevcr_synth_frame *hevcr_synth_generate(const hevcr_synth_cfg *cfg, uint64_t seed);
const hevcr_frame_syntax *hevcr_synth_syntax(const hevcr_synth_frame *f);
void hevcr_synth_free(hevcr_synth_frame *f);

/* Number of distinct reference pictures the frame's lists point at; list entry
 * ref_slot[l][i] is an index in [0, count). */
int hevcr_synth_num_ref_pictures(const hevcr_synth_frame *f);

/* Seeded reference picture: uniform noise low-passed by a 9x9 box filter plus a
 * fixed-point sinusoidal gradient, clipped to the bit depth. Planes are
 * row-major with stride == width (luma w*h, chroma (w/2)*(h/2)). */
int hevcr_synth_ref_picture(int width, int height, int bit_depth, uint64_t seed,
                            uint16_t *y, uint16_t *cb, uint16_t *cr);

#ifdef __cplusplus
}
#endif

#endif /* HEVCR_SYNTH_H */
