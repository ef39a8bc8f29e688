/*
 * hevcr_parse.h — the engine's host front end: Annex-B HEVC bitstream in,
 * reconstruction syntax (hevcr_frame_syntax) and output order out.
 *
 * This is the CPU half the reference decoder runs before its reconstruction
 * operators (NAL parsing src/bitstream/nal_parser.cpp, parameter sets and
 * slice headers src/syntax/, CABAC and the coding tree
 * src/decoding/cabac.cpp + coding_tree.cpp + residual_coding.cpp, merge /
 * AMVP / TMVP src/decoding/inter_prediction.cpp, POC / RPS / reference lists
 * and output bumping src/decoding/dpb.cpp), emitting per picture exactly the
 * records the reconstruction engine consumes.
 *
 * Supported: Main and Main 10 (4:2:0, 8..10 bits), one slice per picture,
 * wavefront parallel processing (entry points), no tiles, no PCM, no
 * transquant bypass; anything else is reported as HEVCR_EUNSUPPORTED.
 *
 * Usage: feed Annex-B bytes in any split, then pull events in order:
 *   HEVCR_EVENT_DECODE  reconstruct `syntax` into DPB slot `slot` of the
 *                       stream of sequence `sequence` (hevcr_decode_frame);
 *   HEVCR_EVENT_OUTPUT  picture `slot` is next in output order (crop it with
 *                       `crop`); its slot is not reused before this event.
 * A new `sequence` number (new SPS geometry) comes after the previous
 * sequence's last OUTPUT event: create a stream of `width` x `height`,
 * `bit_depth`, `dpb_slots` slots for it.
 */
#ifndef HEVCR_PARSE_H
#define HEVCR_PARSE_H

#include <stddef.h>
#include <stdint.h>

#include "hevcr_syntax.h"

#ifdef __cplusplus
extern "C" {
#endif

#define HEVCR_EVENT_DECODE 1
#define HEVCR_EVENT_OUTPUT 2

typedef struct hevcr_parser hevcr_parser;

typedef struct hevcr_parser_event {
    int32_t kind;            /* HEVCR_EVENT_* */
    int32_t sequence;        /* increments when the picture geometry changes */
    int32_t slot;            /* DPB slot */
    int32_t poc;
    int32_t width, height, bit_depth, dpb_slots;
    int32_t crop[4];         /* conformance window in luma samples: left, right, top, bottom */
    int32_t decode_index;    /* decoding order of the picture */
    int32_t profile_idc, level_idc;
    const hevcr_frame_syntax *syntax;  /* DECODE only; valid until the next hevcr_parser_next call */
} hevcr_parser_event;

int hevcr_parser_create(hevcr_parser **out);
void hevcr_parser_destroy(hevcr_parser *p);
/* Feeds Annex-B bytes (start codes included). Complete NAL units are parsed
 * at once; the last one waits for the next start code or the flush. */
int hevcr_parser_feed(hevcr_parser *p, const uint8_t *data, size_t size);
/* End of stream: parses what is buffered and queues the remaining outputs. */
int hevcr_parser_flush(hevcr_parser *p);
/* Next event: 1 and *ev filled, 0 when none is pending, < 0 on error
 * (the feed or flush that failed already returned it). */
int hevcr_parser_next(hevcr_parser *p, hevcr_parser_event *ev);
/* Message of the last error of this parser. */
const char *hevcr_parser_error(const hevcr_parser *p);

#ifdef __cplusplus
}
#endif

#endif /* HEVCR_PARSE_H */
