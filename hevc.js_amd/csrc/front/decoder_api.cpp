// The reference decoder's C API (src/wasm/hevc_api.h, hevc_api.cpp) served
// by the host front end and the GPU engine: parser events drive
// hevcr_decode_frame (DECODE) and hevcr_picture_download (OUTPUT).
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "hevcr.h"
#include "hevcr_api.h"
#include "hevcr_parse.h"

namespace {

// A picture as output: full planes (stride == plane width) and its crop.
struct OutPicture {
    std::vector<uint16_t> plane[3];
    int width = 0, height = 0, bit_depth = 8, poc = 0;
    int crop[4] = {0, 0, 0, 0};
};

}  // namespace

struct HEVCDecoder {
    hevcr_engine* engine = nullptr;
    hevcr_parser* parser = nullptr;
    int stream = -1, sequence = -1;
    std::vector<std::unique_ptr<OutPicture>> output;   // hevc_decoder_decode: every picture, display order
    std::vector<std::unique_ptr<OutPicture>> pending;  // incremental: output since the last drain
    std::vector<std::unique_ptr<OutPicture>> drained;
    HEVCStreamInfo info{};
    bool have_info = false;
    std::string error;

    ~HEVCDecoder() {
        if (engine && stream >= 0) hevcr_stream_destroy(engine, stream);
        if (engine) hevcr_engine_destroy(engine);
        if (parser) hevcr_parser_destroy(parser);
    }
};

namespace {

int fail(HEVCDecoder* d, const std::string& what) {
    d->error = what;
    return HEVC_ERROR;
}

// Runs every queued parser event; OUTPUT pictures go to `sink`.
int run_events(HEVCDecoder* d, std::vector<std::unique_ptr<OutPicture>>& sink) {
    hevcr_parser_event ev;
    while (true) {
        const int rc = hevcr_parser_next(d->parser, &ev);
        if (rc < 0) return fail(d, std::string("parse: ") + hevcr_parser_error(d->parser));
        if (rc == 0) return HEVC_OK;
        if (ev.sequence != d->sequence) {  // new picture geometry: a new engine stream
            if (d->stream >= 0) hevcr_stream_destroy(d->engine, d->stream);
            d->stream = -1;
            if (hevcr_stream_create(d->engine, ev.width, ev.height, ev.bit_depth, ev.dpb_slots, &d->stream) != HEVCR_OK)
                return fail(d, std::string("engine stream: ") + hevcr_last_error());
            d->sequence = ev.sequence;
        }
        if (ev.kind == HEVCR_EVENT_DECODE) {
            if (hevcr_decode_frame(d->engine, d->stream, ev.syntax, ev.slot) != HEVCR_OK)
                return fail(d, std::string("engine: ") + hevcr_last_error());
            continue;
        }
        auto pic = std::make_unique<OutPicture>();
        pic->width = ev.width;
        pic->height = ev.height;
        pic->bit_depth = ev.bit_depth;
        pic->poc = ev.poc;
        std::memcpy(pic->crop, ev.crop, sizeof(pic->crop));
        pic->plane[0].resize(static_cast<size_t>(ev.width) * ev.height);
        pic->plane[1].resize(static_cast<size_t>(ev.width / 2) * (ev.height / 2));
        pic->plane[2].resize(pic->plane[1].size());
        if (hevcr_picture_download(d->engine, d->stream, ev.slot, pic->plane[0].data(), pic->plane[1].data(),
                                   pic->plane[2].data()) != HEVCR_OK)
            return fail(d, std::string("download: ") + hevcr_last_error());
        if (!d->have_info) {
            d->info.width = ev.width - ev.crop[0] - ev.crop[1];
            d->info.height = ev.height - ev.crop[2] - ev.crop[3];
            d->info.bit_depth = ev.bit_depth;
            d->info.chroma_format = 1;
            d->info.profile = ev.profile_idc;
            d->info.level = ev.level_idc;
            d->have_info = true;
        }
        sink.push_back(std::move(pic));
    }
}

void fill_frame(const OutPicture& p, HEVCFrame* f) {
    // hevc_api.cpp:49-77: plane pointers offset by the conformance window
    const int l = p.crop[0], r = p.crop[1], t = p.crop[2], b = p.crop[3];
    f->width = p.width - l - r;
    f->height = p.height - t - b;
    f->stride_y = p.width;
    f->stride_c = p.width / 2;
    f->y = p.plane[0].data() + static_cast<size_t>(t) * p.width + l;
    f->cb = p.plane[1].data() + static_cast<size_t>(t / 2) * (p.width / 2) + l / 2;
    f->cr = p.plane[2].data() + static_cast<size_t>(t / 2) * (p.width / 2) + l / 2;
    f->chroma_width = f->width / 2;
    f->chroma_height = f->height / 2;
    f->bit_depth = p.bit_depth;
    f->poc = p.poc;
}

int ensure(HEVCDecoder* d) {
    if (!d->engine) {
        const char* dev = std::getenv("HEVCR_DEVICE");
        if (hevcr_engine_create(dev ? std::atoi(dev) : 0, &d->engine) != HEVCR_OK)
            return fail(d, std::string("engine: ") + hevcr_last_error());
    }
    if (!d->parser && hevcr_parser_create(&d->parser) != HEVCR_OK) return fail(d, "parser allocation");
    return HEVC_OK;
}

}  // namespace

extern "C" {

HEVCDecoder* hevc_decoder_create(void) { return new (std::nothrow) HEVCDecoder(); }

void hevc_decoder_destroy(HEVCDecoder* dec) { delete dec; }

int hevc_decoder_decode(HEVCDecoder* dec, const uint8_t* data, size_t size) {
    if (!dec || !data || size == 0) return HEVC_ERROR;
    if (ensure(dec)) return HEVC_ERROR;
    dec->output.clear();
    if (hevcr_parser_feed(dec->parser, data, size) < 0 || hevcr_parser_flush(dec->parser) < 0) {
        run_events(dec, dec->output);  // pictures before the error stay readable, as in hevc_api.cpp:31-36
        dec->error = std::string("parse: ") + hevcr_parser_error(dec->parser);
        return dec->output.empty() ? HEVC_ERROR : HEVC_OK;
    }
    run_events(dec, dec->output);  // on an engine error the pictures before it stay readable
    return dec->output.empty() ? HEVC_ERROR : HEVC_OK;
}

int hevc_decoder_get_frame_count(HEVCDecoder* dec) { return dec ? static_cast<int>(dec->output.size()) : 0; }

int hevc_decoder_get_frame(HEVCDecoder* dec, int index, HEVCFrame* frame) {
    if (!dec || !frame || index < 0 || index >= static_cast<int>(dec->output.size())) return HEVC_ERROR;
    fill_frame(*dec->output[static_cast<size_t>(index)], frame);
    return HEVC_OK;
}

int hevc_decoder_get_info(HEVCDecoder* dec, HEVCStreamInfo* info) {
    if (!dec || !info || !dec->have_info) return HEVC_ERROR;
    *info = dec->info;
    return HEVC_OK;
}

int hevc_decoder_feed(HEVCDecoder* dec, const uint8_t* data, size_t size) {
    if (!dec || !data || size == 0) return HEVC_ERROR;
    if (ensure(dec)) return HEVC_ERROR;
    if (hevcr_parser_feed(dec->parser, data, size) < 0)
        return fail(dec, std::string("parse: ") + hevcr_parser_error(dec->parser));
    return run_events(dec, dec->pending);
}

int hevc_decoder_drain(HEVCDecoder* dec, int* count) {
    if (!dec || !count) return HEVC_ERROR;
    dec->drained = std::move(dec->pending);
    dec->pending.clear();
    *count = static_cast<int>(dec->drained.size());
    return HEVC_OK;
}

int hevc_decoder_get_drained_frame(HEVCDecoder* dec, int index, HEVCFrame* frame) {
    if (!dec || !frame || index < 0 || index >= static_cast<int>(dec->drained.size())) return HEVC_ERROR;
    fill_frame(*dec->drained[static_cast<size_t>(index)], frame);
    return HEVC_OK;
}

int hevc_decoder_flush(HEVCDecoder* dec) {
    if (!dec) return HEVC_ERROR;
    if (ensure(dec)) return HEVC_ERROR;
    if (hevcr_parser_flush(dec->parser) < 0) return fail(dec, std::string("parse: ") + hevcr_parser_error(dec->parser));
    const int rc = run_events(dec, dec->pending);
    // as in hevc_api.cpp:143-151, the flushed pictures are readable through
    // hevc_decoder_get_drained_frame at once (with any not drained before)
    dec->drained = std::move(dec->pending);
    dec->pending.clear();
    return rc;
}

const char* hevcr_decoder_error(const HEVCDecoder* dec) { return dec ? dec->error.c_str() : "null decoder"; }

}  // extern "C"
