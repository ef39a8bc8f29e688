/*
 * hevcr.h — C-ABI of the MI355X HEVC reconstruction engine (libhevcr.so).
 *
 * The engine replaces the reference decoder's internal reconstruction
 * operators, which its CABAC loop calls one block at a time
 * (src/decoding/coding_tree.cpp:14-26 declares them for that caller):
 *   perform_dequant            src/decoding/transform.cpp:293
 *   perform_transform_inverse  src/decoding/transform.cpp:374
 *   perform_intra_prediction   src/decoding/intra_prediction.cpp:428
 *   perform_inter_prediction   src/decoding/interpolation.cpp:306
 *   apply_deblocking           src/filters/deblocking.cpp:357 (called at decoder.cpp:270)
 *   apply_sao                  src/filters/sao.cpp:20         (called at decoder.cpp:271)
 * with one per-frame call that takes the whole frame's syntax
 * (hevcr_syntax.h) and reconstructs it on the GPU into a decoded-picture
 * buffer kept in HBM. Pictures leave the device only on request.
 *
 * Conventions (mirroring src/wasm/hevc_api.h:43-45): functions return
 * HEVCR_OK (0) or a negative code; hevcr_last_error() describes the last
 * failure of the calling thread. Invalid syntax is rejected on the host
 * before any kernel is launched. One engine may be used from one thread at a
 * time; engines are independent.
 */
#ifndef HEVCR_H
#define HEVCR_H

#include "hevcr_syntax.h"

#ifdef __cplusplus
extern "C" {
#endif

#define HEVCR_OK            0
#define HEVCR_ERROR        -1  /* HEVC_ERROR of hevc_api.h */
#define HEVCR_NEED_MORE    -2  /* HEVC_NEED_MORE of hevc_api.h */
#define HEVCR_EINVAL       -3  /* invalid argument or syntax */
#define HEVCR_ENOMEM       -4  /* device or host allocation failed / budget exceeded */
#define HEVCR_EDEVICE      -5  /* no usable GPU or a HIP call failed */
#define HEVCR_EUNSUPPORTED -6  /* syntax outside the engine's scope */

/* Stop points for hevcr_frame_run (stage outputs for parity checks). */
#define HEVCR_RUN_FULL        0  /* reconstruction, deblocking, SAO */
#define HEVCR_RUN_STOP_RECON  1  /* prediction + residual only (pre-deblock picture) */
#define HEVCR_RUN_STOP_DEBLOCK 2 /* + deblocking (pre-SAO picture) */

/* Timed stages (hevcr_stage_times). */
#define HEVCR_STAGE_EXPAND   0
#define HEVCR_STAGE_MC       1
#define HEVCR_STAGE_RESIDUAL 2
#define HEVCR_STAGE_INTRA    3
#define HEVCR_STAGE_DEBLOCK  4
#define HEVCR_STAGE_SAO      5
#define HEVCR_NUM_STAGES     6

/* Timed kernels (hevcr_kernel_times), named by hevcr_kernel_name. */
#define HEVCR_KERNEL_EXPAND     0
#define HEVCR_KERNEL_MC         1  /* tiles other than 8x8 */
#define HEVCR_KERNEL_MC8        2  /* 8x8 tiles, four per wavefront */
#define HEVCR_KERNEL_RESIDUAL4  3  /* + 1, 2, 3 for 8x8, 16x16, 32x32 blocks */
#define HEVCR_KERNEL_INTRA      7
#define HEVCR_KERNEL_DEBLOCK    8
#define HEVCR_KERNEL_SAO        9
#define HEVCR_KERNEL_LOOPFILTER 10 /* deblocking + SAO in one pass (64x64 CTBs, SAO on) */
#define HEVCR_NUM_KERNELS      11

typedef struct hevcr_engine hevcr_engine;
typedef struct hevcr_frame hevcr_frame;
typedef struct hevcr_batch hevcr_batch;

/* Last error message of the calling thread ("" if none). */
const char *hevcr_last_error(void);

/* Engine on one GPU. Fails with HEVCR_EDEVICE when no GPU is usable. */
int hevcr_engine_create(int device, hevcr_engine **out);
void hevcr_engine_destroy(hevcr_engine *e);
/* Waits for the engine's queued work. HEVCR_EDEVICE if a persistent intra
 * launch since the last call gave up waiting for a neighbour (its pictures
 * are then not to be trusted; the flag is cleared). */
int hevcr_engine_sync(hevcr_engine *e);

/* A decoding stream: its pictures (the DPB, dpb_slots pictures of
 * width x height at bit_depth 8..10, 4:2:0; 1..128 slots) live in HBM: 1 byte per sample at
 * 8 bits, 2 bytes above. */
int hevcr_stream_create(hevcr_engine *e, int width, int height, int bit_depth, int dpb_slots, int *stream_id);
int hevcr_stream_destroy(hevcr_engine *e, int stream_id);

/* Copies a picture between host uint16 planes (stride == plane width) and a
 * DPB slot. Download serialises exactly the decoded samples (converted on the
 * device, as HEVCR_EGRESS_U16 below, then one copy per plane). */
int hevcr_picture_upload(hevcr_engine *e, int stream_id, int slot, const uint16_t *y, const uint16_t *cb,
                         const uint16_t *cr);
int hevcr_picture_download(hevcr_engine *e, int stream_id, int slot, uint16_t *y, uint16_t *cb, uint16_t *cr);

/* ---- Egress ----------------------------------------------------------------
 * Pictures out of the DPB in the reference's two output layouts:
 *  HEVCR_EGRESS_YUV — Picture::write_yuv (picture.cpp:42-99): for Y, Cb, Cr
 *    the conformance window (crop = left, right, top, bottom in luma samples,
 *    even; NULL = none), rows without padding, samples as bytes at 8 bits and
 *    as 16-bit little-endian words above: the bytes the picture adds to a
 *    .yuv file;
 *  HEVCR_EGRESS_U16 — the planes of hevc_api.h's HEVCFrame before its crop
 *    offsets (hevc_api.cpp:49-77): Y, Cb, Cr whole, uint16 samples, stride ==
 *    plane width (crop ignored). */
#define HEVCR_EGRESS_YUV 0
#define HEVCR_EGRESS_U16 1

/* Bytes of one picture of this size and bit depth in `format`. */
int hevcr_egress_size(int width, int height, int bit_depth, int format, const int *crop, uint64_t *bytes);
/* Writes n pictures one after another into dst (each at the sum of the sizes
 * of those before it) with one kernel launch per sample size; crops: 4 ints
 * per picture, or NULL. dst is device memory, pinned host memory from
 * hevcr_host_alloc (written by the GPU directly, no staging copy), or any
 * other host memory (written to device memory first, then copied). Runs after
 * the work queued on the engine's stream; returns when dst holds the pictures. */
int hevcr_picture_egress(hevcr_engine *e, int n, const int *stream_ids, const int *slots, const int *crops,
                         int format, void *dst, uint64_t dst_bytes);
/* Pinned host memory the GPU can write directly (zero-copy egress). */
int hevcr_host_alloc(uint64_t bytes, void **out);
void hevcr_host_free(void *p);

/* The compact coefficient form of a frame whose coefficients are dense
 * (hevcr_frame_syntax.sig): with sig NULL, *num_sig receives the number of
 * words it needs; otherwise sig (cap words) receives the lists and
 * tu_coeff_off (3 per TU record) each coded block's index in sig, to be put
 * into a copy of the TU records. Host only; a parser that knows its nonzero
 * coefficients can emit the form directly instead. */
int hevcr_syntax_compact(const hevcr_frame_syntax *syn, uint32_t *sig, uint64_t cap, uint64_t *num_sig,
                         uint32_t *tu_coeff_off);

/* Plain form: validates and uploads the frame's syntax, reconstructs it into
 * DPB slot out_slot (list entries reference slots through
 * syn->p.ref_slot), one kernel per stage, in order; returns when queued. */
int hevcr_decode_frame(hevcr_engine *e, int stream_id, const hevcr_frame_syntax *syn, int out_slot);

/* The same split in two, so that syntax can be made resident once and the
 * reconstruction launched many times (benchmarks, replay). Preparing copies
 * the syntax and its derived work lists to the device on a stream of its own,
 * so it does not wait for queued reconstruction work; unlike the other calls,
 * hevcr_frame_prepare may run on several threads at once (not concurrently
 * with creating or destroying streams). */
int hevcr_frame_prepare(hevcr_engine *e, int stream_id, const hevcr_frame_syntax *syn, hevcr_frame **out);
int hevcr_frame_run(hevcr_engine *e, hevcr_frame *f, int out_slot, int stop);
void hevcr_frame_release(hevcr_engine *e, hevcr_frame *f);

/* Batched form: n prepared frames (any streams, resolutions and bit depths)
 * reconstructed together, each stage one launch over all of them (intra: one
 * launch per intra wave of CTUs), optionally captured once into a hipGraph and
 * replayed. Frame i writes DPB slot out_slots[i] of its stream; no two frames
 * may write the same slot, and no frame may reference a slot the batch
 * writes (frames of a batch are independent). The frames must stay alive
 * while the batch exists. The reference reconstructs one picture at a time
 * (decoder.cpp:52-304); this has no single counterpart there. */
int hevcr_batch_create(hevcr_engine *e, hevcr_frame *const *frames, const int *out_slots, int n, int stop,
                       hevcr_batch **out);
int hevcr_batch_run(hevcr_engine *e, hevcr_batch *b, int use_graph);
void hevcr_batch_release(hevcr_engine *e, hevcr_batch *b);
int hevcr_batch_bytes(const hevcr_batch *b, double *bytes);

/* ---- Tuning ------------------------------------------------------------------
 * How frames of one picture size and bit depth are cut into work, fixed when a
 * frame is prepared (every choice gives the same pictures, bit for bit):
 *  mc8          — 8x8 PUs four to a wavefront (k_mc8), else as 16x16-class
 *                 tiles (k_mc); PUs predicted as tiles only;
 *  lf_fused     — deblocking + SAO in one kernel over 64x64 areas where
 *                 eligible (SAO on, 32x32 or 64x64 CTBs), else k_deblock then
 *                 k_sao;
 *  intra_group4 — independent 4x4 intra blocks of a dependency level four to
 *                 a step (16 lanes each), else one block per step;
 *  intra_persistent — the intra stage as one launch in which resident
 *                 workgroups take CTUs in dependency order and wait for their
 *                 neighbours' flags, else one launch per dependency wave;
 *  mc_strips    — motion compensation by strips of 4 luma columns, a lane
 *                 each (k_mcs), for every band of a PU whose reference windows
 *                 lie inside the picture: 1 = up to 16 rows (uni-predicted
 *                 8-bit pictures; 8 otherwise), 2 = up to 8 rows; 0 = tiles
 *                 only (k_mc, k_mc8). Bands that need the edge clamping are
 *                 always tiles.
 * All fields are 1 until set. No counterpart in the reference (it has one
 * CPU code path). */
typedef struct hevcr_tuning {
    int32_t mc8, lf_fused, intra_group4, intra_persistent, mc_strips;
} hevcr_tuning;
#define HEVCR_TUNE_CANDIDATES 32 /* candidate k: mc8 = k & 1, lf_fused = k >> 1 & 1, intra_group4 = k >> 2 & 1,
                                    intra_persistent = k >> 3 & 1, mc_strips = k >> 4 & 1 */
int hevcr_tuning_set(hevcr_engine *e, int width, int height, int bit_depth, const hevcr_tuning *t);
int hevcr_tuning_get(hevcr_engine *e, int width, int height, int bit_depth, hevcr_tuning *t);
/* The autotuner: for n frames of one picture size and bit depth (their
 * syntax, streams and output slots as for hevcr_batch_create), prepares and
 * runs the batched form (graph replays) under every candidate, checks that
 * every candidate's pictures equal the first's (device digests; a mismatch is
 * HEVCR_ERROR), times `reps` replays each (HIP events), stores the fastest
 * candidate as that size and bit depth's tuning and returns it, with the
 * average ms per replay of each candidate in ms[HEVCR_TUNE_CANDIDATES]. */
int hevcr_autotune(hevcr_engine *e, int n, const int *stream_ids, const hevcr_frame_syntax *const *syn,
                   const int *out_slots, int reps, hevcr_tuning *best, double *ms);

/* Algorithmic HBM bytes of one run of a prepared frame, per stage
 * (HEVCR_NUM_STAGES entries) and per kernel (HEVCR_NUM_KERNELS entries). */
int hevcr_frame_bytes(const hevcr_frame *f, double *bytes);
int hevcr_frame_kernel_bytes(const hevcr_frame *f, double *bytes);
int hevcr_batch_kernel_bytes(const hevcr_batch *b, double *bytes);

/* Per-stage GPU time (HIP events on the engine's stream) accumulated over the
 * runs since the last reset: ms[HEVCR_NUM_STAGES], launches[HEVCR_NUM_STAGES]. */
int hevcr_set_profiling(hevcr_engine *e, int on);
int hevcr_stage_times(hevcr_engine *e, double *ms, uint64_t *launches);
int hevcr_reset_stage_times(hevcr_engine *e);
/* The same per kernel (every launch of it): ms[HEVCR_NUM_KERNELS], launches[HEVCR_NUM_KERNELS]. */
int hevcr_kernel_times(hevcr_engine *e, double *ms, uint64_t *launches);
const char *hevcr_kernel_name(int kernel);

/* Batched dequantisation + inverse transform of independent blocks
 * (perform_dequant + perform_transform_inverse over a block list): n blocks,
 * block i of size 1 << log2[i] with component cidx[i], qP' qp[i], flags[i]
 * (bit 0 intra, bit 1 transform skip), coefficients at coeffs + off[i]
 * (dense row-major int16); residuals written at out + off[i]. Host buffers. */
int hevcr_residual_blocks(hevcr_engine *e, int bit_depth, uint32_t n, const uint8_t *log2, const uint8_t *cidx,
                          const uint8_t *qp, const uint8_t *flags, const uint32_t *off, const int16_t *coeffs,
                          uint64_t num_coeffs, int16_t *out);

/* hevcr_residual_blocks without allocation: the same arguments, results and
 * errors, for callers that cannot allocate device memory (a harness that
 * forbids hipMalloc, a tight serving loop). No counterpart in the reference.
 * Staged coefficients, block lists, the residual plane and the outputs live
 * in one static device array of the library (about 6.3 MiB per device, held
 * until the process ends); at most 8192 blocks, 32768 rows of 32 residual
 * samples (the blocks' sizes summed) and 2^20 coefficients per call, else
 * HEVCR_EINVAL. Calls share that array on a GPU: calls that overlap on one
 * GPU (engines on different host threads) race. Waits for its result. */
int hevcr_residual_blocks_noalloc(hevcr_engine *e, int bit_depth, uint32_t n, const uint8_t *log2,
                                  const uint8_t *cidx, const uint8_t *qp, const uint8_t *flags,
                                  const uint32_t *off, const int16_t *coeffs, uint64_t num_coeffs, int16_t *out);

/* Verification digests of n DPB pictures, three per picture (Y, Cb, Cr) in
 * digests[3 * i + c]: the sum over the plane's samples, in raster order i, of
 * (v + 1) * mix64(i ^ key_c) modulo 2^64 (mix64 = SplitMix64 finaliser,
 * key_c = 0x5851F42D4C957F2D * (c + 1)). Waits for queued work. Equal
 * pictures give equal digests; the reference has no counterpart (its checks
 * hash the written YUV file, main.cpp:281-325), this one lets pictures be
 * compared without copying them to the host. */
int hevcr_picture_digest(hevcr_engine *e, int n, const int *stream_ids, const int *slots, uint64_t *digests);

/* SHA-256 (FIPS 180-4) of each plane of n DPB pictures, computed on the GPU
 * over the bytes Picture::write_yuv writes for the plane (picture.cpp:42-99:
 * the conformance window crop = left, right, top, bottom in luma samples,
 * even, 4 ints per picture or NULL for none; rows without padding; samples
 * as bytes at 8 bits, 16-bit little-endian words above). digests receives
 * 96 bytes per picture: the 32-byte digests of Y, Cb, Cr. Runs after the work
 * queued on the engine's stream and waits for the result. The reference's
 * checks hash the same bytes on the host (tools/oracle_test.sh hashes the
 * written .yuv file; main.cpp:281-325 writes it); this keeps the pictures on
 * the device. A plane is one message, hashed by one GPU lane, so a call is
 * fast in aggregate (many planes at once), not per plane. */
int hevcr_picture_sha256(hevcr_engine *e, int n, const int *stream_ids, const int *slots, const int *crops,
                         uint8_t *digests);

/* Measurement helper with no counterpart in the reference: copies `bytes`
 * bytes between two device buffers `reps` times with a plain vector-copy
 * kernel (k_copy) and waits. Used to calibrate hardware byte counters
 * against a kernel of known traffic. */
int hevcr_bench_copy(hevcr_engine *e, uint64_t bytes, int reps);

/* ---- Serving --------------------------------------------------------------
 * A request loop over one engine with continuous batching. The reference
 * decodes one stream per decoder instance (hevc_api.h:14-86, one
 * HEVCDecoder per stream, hevc_decoder_feed per chunk); a server instead
 * serves many streams on one GPU:
 *  - a stream is admitted when its DPB (dpb_slots pictures) fits in the
 *    remaining budget, first come first served; a stream that does not fit
 *    yet waits until closed streams free enough;
 *  - frames are prepared with hevcr_frame_prepare on the stream id the server
 *    returns (also before admission) and submitted; a stream's frames run in
 *    submission order;
 *  - each hevcr_server_step launches one batch holding the next frame of up
 *    to max_batch admitted streams (oldest arrival first), if fewer than
 *    max_inflight steps are queued on the device. Frames that arrive while
 *    steps run join a later step;
 *  - steps run on `lanes` device streams, concurrently: a step goes to the
 *    lane with the fewest queued steps, and takes only streams that have no
 *    frame in an unfinished step of another lane. Steps complete (and are
 *    polled) in any order across lanes.
 * Submitted frames must stay alive until their completion has been polled.
 * Single-threaded, like the engine. */
typedef struct hevcr_server hevcr_server;

typedef struct hevcr_server_cfg {
    uint64_t dpb_budget_bytes; /* device memory the admitted streams' DPBs may hold */
    int32_t max_batch;         /* frames per step, 1..1024 */
    int32_t max_inflight;      /* steps queued on the device at once, 1..8 */
    int32_t lanes;             /* device streams steps run on concurrently, 1..4; a step takes at
                                  most ceil(admitted streams / lanes) frames (lanes > 1) */
    int32_t ingest_threads;    /* host threads preparing frames submitted as syntax, 0..64 (0: only
                                  hevcr_server_submit of prepared frames) */
} hevcr_server_cfg;

typedef struct hevcr_done {
    uint64_t tag;              /* as submitted */
    int32_t stream_id;
    int32_t out_slot;
    int32_t status;            /* HEVCR_OK */
    int32_t step_frames;       /* frames of the step that reconstructed it */
    double t_submit;           /* arrival (hevcr_server_clock seconds) */
    double t_launch;           /* its step was launched */
    double t_done;             /* the host saw its step complete */
} hevcr_done;

typedef struct hevcr_server_state {
    int32_t streams, admitted, queued_frames, inflight_steps;
    uint64_t budget_used, frames_done, steps_done;
    uint64_t ingested_frames, ingest_bytes; /* frames prepared from syntax, bytes they uploaded */
    double ingest_busy_s;                   /* seconds the ingest threads spent preparing, summed */
    /* that time by phase: validation, work lists, copy into pinned staging,
     * the upload (launch and wait) */
    double ingest_validate_s, ingest_lists_s, ingest_stage_s, ingest_upload_s;
} hevcr_server_state;

int hevcr_server_create(hevcr_engine *e, const hevcr_server_cfg *cfg, hevcr_server **out);
void hevcr_server_destroy(hevcr_server *s); /* waits for queued steps, frees the streams */
double hevcr_server_clock(void);            /* monotonic seconds */
int hevcr_server_open(hevcr_server *s, int width, int height, int bit_depth, int dpb_slots, int *stream_id);
/* The stream's DPB is freed once its submitted frames are done. */
int hevcr_server_close(hevcr_server *s, int stream_id);
int hevcr_server_admitted(hevcr_server *s, int stream_id, int *admitted);
/* t_arrival < 0: now. */
int hevcr_server_submit(hevcr_server *s, hevcr_frame *f, int out_slot, uint64_t tag, double t_arrival);
/* Ingest: a frame submitted as host syntax (the parser's output for stream
 * stream_id) is validated and prepared by one of the server's ingest threads
 * (its work lists built on the host, its device block from the engine's
 * pool, its upload on the thread's own device stream from pinned staging)
 * while steps run, and is run by the first step after it is ready, in
 * submission order within its stream. The syntax must stay valid until the
 * frame's completion has been polled; the server releases the prepared
 * frame itself when its step retires. A frame that fails validation
 * completes with that error as status. Tuning must not change meanwhile.
 * While a step is queued and submitted frames are still being prepared, a
 * step that would hold fewer frames than a full one is not launched yet. */
int hevcr_server_submit_syntax(hevcr_server *s, int stream_id, const hevcr_frame_syntax *syn, int out_slot,
                               uint64_t tag, double t_arrival);
/* *launched = frames in the launched step, 0 if none was launched. */
int hevcr_server_step(hevcr_server *s, int *launched);
/* Retires finished steps, oldest first (wait != 0: blocks until at least one
 * completion is available, if any step is queued), and returns up to cap
 * completions. */
int hevcr_server_poll(hevcr_server *s, int wait, hevcr_done *out, int cap, int *n);
int hevcr_server_info(hevcr_server *s, hevcr_server_state *info);

#ifdef __cplusplus
}
#endif

#endif /* HEVCR_H */
