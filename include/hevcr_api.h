/*
 * hevcr_api.h — the reference decoder's C API (src/wasm/hevc_api.h:14-86),
 * kept name for name, type for type and return code for return code, served
 * by the host front end (hevcr_parse.h) and the GPU reconstruction engine
 * (hevcr.h) instead of the reference's CPU decoder. A program written
 * against hevc_api.h links libhevcr_decoder.so instead and is unchanged.
 *
 * Differences a caller can see: streams outside the engine's scope (more
 * than one slice per picture, tiles, PCM, transquant bypass, 4:2:2/4:4:4,
 * more than 10 bits) return HEVC_ERROR, with the reason in
 * hevcr_decoder_error(). Pictures are reconstructed on the GPU while the
 * host parses the next part of the stream and are downloaded in batches
 * before each call returns. Every decoder of a process shares one engine
 * on HIP device 0, or on the one HEVCR_DEVICE names; as in the reference, one
 * decoder is used from one thread at a time, and decoders on different
 * threads are independent (their engine calls are serialised internally).
 */
#ifndef HEVCR_API_H
#define HEVCR_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct HEVCDecoder HEVCDecoder;

/* hevc_api.h:18-31 */
typedef struct {
    const uint16_t *y;
    const uint16_t *cb;
    const uint16_t *cr;
    int width;          /* luma width after the conformance window crop */
    int height;
    int stride_y;       /* in samples */
    int stride_c;
    int chroma_width;
    int chroma_height;
    int bit_depth;
    int poc;
} HEVCFrame;

/* hevc_api.h:34-41 */
typedef struct {
    int width;
    int height;
    int bit_depth;
    int chroma_format;
    int profile;
    int level;
} HEVCStreamInfo;

#define HEVC_OK          0
#define HEVC_ERROR      -1
#define HEVC_NEED_MORE  -2

HEVCDecoder *hevc_decoder_create(void);
void hevc_decoder_destroy(HEVCDecoder *dec);
/* Whole bitstream: decodes every picture; frames are then read by index in display order. */
int hevc_decoder_decode(HEVCDecoder *dec, const uint8_t *data, size_t size);
int hevc_decoder_get_frame_count(HEVCDecoder *dec);
int hevc_decoder_get_frame(HEVCDecoder *dec, int index, HEVCFrame *frame);
int hevc_decoder_get_info(HEVCDecoder *dec, HEVCStreamInfo *info);
/* Incremental: feed chunks, drain what the bumping process has output, flush at the end. */
int hevc_decoder_feed(HEVCDecoder *dec, const uint8_t *data, size_t size);
int hevc_decoder_drain(HEVCDecoder *dec, int *count);
int hevc_decoder_get_drained_frame(HEVCDecoder *dec, int index, HEVCFrame *frame);
int hevc_decoder_flush(HEVCDecoder *dec);

/* No counterpart in hevc_api.h: the reason of the last HEVC_ERROR. */
const char *hevcr_decoder_error(const HEVCDecoder *dec);

#ifdef __cplusplus
}
#endif

#endif /* HEVCR_API_H */
