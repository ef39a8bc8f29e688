/*
 * hevc_oracle.h — CPU restatement of the reference decoder's reconstruction
 * path, used ONLY as test infrastructure (the checker for the HIP engine and
 * the CPU baseline leg of the benchmark). Nothing in the product links it.
 */
#ifndef HEVC_ORACLE_H
#define HEVC_ORACLE_H

#include "hevcr_syntax.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Reconstructs one frame: PU prediction, TU dequant + inverse transform +
 * intra prediction + reconstruction in decoding order, then deblocking and
 * SAO. ref_planes holds 3 uint16 planes (stride == width) per reference
 * picture slot named by syn->p.ref_slot; out_* are 3 plane buffers or NULL.
 * Returns 0, or a negative code for syntax the restatement does not cover. */
int oracle_recon_frame(const hevcr_frame_syntax *syn, const uint16_t *const *ref_planes,
                       int num_ref_pics, uint16_t *const *out_recon,
                       uint16_t *const *out_dbk, uint16_t *const *out_sao);

/* Dequantisation (flat scaling) + inverse transform of one NxN block. */
int oracle_residual_block(int log2, int cidx, int is_intra, int qp_prime, int bit_depth,
                          int tskip, const int16_t *coeffs, int16_t *out);

/* The two in-loop filters alone, in place on planes that already hold the
 * pre-deblock (resp. post-deblock) reconstruction. They need the frame's
 * syntax for the side information (grids, QPs, motion, SAO parameters). */
int oracle_filters(const hevcr_frame_syntax *syn, uint16_t *const *planes, int do_deblock, int do_sao);

#ifdef __cplusplus
}
#endif

#endif
