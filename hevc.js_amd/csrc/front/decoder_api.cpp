// The reference decoder's C API (src/wasm/hevc_api.h, hevc_api.cpp) served
// by the host front end and the GPU engine: parser events drive
// hevcr_decode_frame (DECODE) and batched hevcr_picture_egress (OUTPUT).
//
// Work flow of one call:
//  - the bitstream is fed to the parser in chunks; after each chunk the
//    queued events run, so the GPU reconstructs the pictures parsed so far
//    while the host parses the next chunk;
//  - an OUTPUT event only records the picture; its download is deferred and
//    batched with the other recorded pictures into one egress call, made when
//    a DECODE is about to rewrite a recorded picture's slot, before the
//    stream changes, and before the call returns (the only waits on the GPU);
//  - every HEVCDecoder of a process shares one engine per device (a mutex
//    serialises their engine calls; parsing runs outside it), as one
//    serving engine carries many streams.
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "hevcr.h"
#include "hevcr_api.h"
#include "hevcr_parse.h"

namespace {

// Bytes handed to the parser at a time: small enough that reconstruction of
// the pictures parsed so far overlaps the parsing of the next chunk, large
// enough that the per-chunk event handling stays negligible.
constexpr size_t kFeedChunk = 32 * 1024;

// A picture as output: full uint16 planes (stride == plane width), Y then
// Cb then Cr, inside a buffer shared with the pictures downloaded by the
// same egress call.
struct OutPicture {
    std::shared_ptr<std::vector<uint16_t>> buf;
    size_t off = 0;  // first sample of Y in buf
    int width = 0, height = 0, bit_depth = 8, poc = 0;
    int crop[4] = {0, 0, 0, 0};
    const uint16_t* plane(int c) const {
        const size_t luma = static_cast<size_t>(width) * height;
        const size_t chroma = static_cast<size_t>(width / 2) * (height / 2);
        return buf->data() + off + (c == 0 ? 0 : luma + (c == 2 ? chroma : 0));
    }
};

// One engine per device for the whole process.
struct SharedEngine {
    std::mutex m;  // held for every engine call of any decoder
    hevcr_engine* e = nullptr;
    int users = 0;
};
std::mutex g_engines_mu;
std::map<int, SharedEngine*> g_engines;

}  // namespace

struct HEVCDecoder {
    SharedEngine* shared = nullptr;
    int device = 0;
    hevcr_parser* parser = nullptr;
    int stream = -1, sequence = -1;
    std::vector<std::unique_ptr<OutPicture>> output;   // hevc_decoder_decode: every picture, display order
    std::vector<std::unique_ptr<OutPicture>> pending;  // incremental: output since the last drain
    std::vector<std::unique_ptr<OutPicture>> drained;
    // recorded OUTPUT events whose samples are not downloaded yet
    struct Deferred {
        OutPicture* pic;
        int slot;
    };
    std::vector<Deferred> deferred;
    std::string error;

    ~HEVCDecoder();
};

namespace {

int fail(HEVCDecoder* d, const std::string& what) {
    d->error = what;
    return HEVC_ERROR;
}

int acquire_engine(HEVCDecoder* d) {
    const char* dev = std::getenv("HEVCR_DEVICE");
    d->device = dev ? std::atoi(dev) : 0;
    std::lock_guard<std::mutex> g(g_engines_mu);
    SharedEngine*& se = g_engines[d->device];
    if (!se) se = new SharedEngine();
    if (!se->e && hevcr_engine_create(d->device, &se->e) != HEVCR_OK)
        return fail(d, std::string("engine: ") + hevcr_last_error());
    se->users++;
    d->shared = se;
    return HEVC_OK;
}

void release_engine(HEVCDecoder* d) {
    if (!d->shared) return;
    std::lock_guard<std::mutex> g(g_engines_mu);
    SharedEngine* se = d->shared;
    d->shared = nullptr;
    if (--se->users == 0 && se->e) {
        hevcr_engine_destroy(se->e);  // the last decoder of this device: free the GPU resources
        se->e = nullptr;
    }
}

// Downloads every deferred picture with one egress call (HEVCFrame's uint16
// layout); the engine mutex is held by the caller.
int download_deferred(HEVCDecoder* d) {
    if (d->deferred.empty()) return HEVC_OK;
    std::vector<int> ids, slots;
    size_t total = 0;
    for (const auto& q : d->deferred) {
        ids.push_back(d->stream);
        slots.push_back(q.slot);
        q.pic->off = total;
        total += static_cast<size_t>(q.pic->width) * q.pic->height * 3 / 2;
    }
    auto buf = std::make_shared<std::vector<uint16_t>>(total);
    for (const auto& q : d->deferred) q.pic->buf = buf;
    const int n = static_cast<int>(d->deferred.size());
    d->deferred.clear();
    if (hevcr_picture_egress(d->shared->e, n, ids.data(), slots.data(), nullptr, HEVCR_EGRESS_U16, buf->data(),
                             total * sizeof(uint16_t)) != HEVCR_OK)
        return fail(d, std::string("download: ") + hevcr_last_error());
    return HEVC_OK;
}

// Runs every queued parser event; OUTPUT pictures go to `sink` (their
// samples follow with the next download_deferred).
int run_events(HEVCDecoder* d, std::vector<std::unique_ptr<OutPicture>>& sink) {
    std::lock_guard<std::mutex> g(d->shared->m);
    hevcr_engine* e = d->shared->e;
    hevcr_parser_event ev;
    while (true) {
        const int rc = hevcr_parser_next(d->parser, &ev);
        if (rc < 0) return fail(d, std::string("parse: ") + hevcr_parser_error(d->parser));
        if (rc == 0) return HEVC_OK;
        if (ev.sequence != d->sequence) {  // new picture geometry: a new engine stream
            if (download_deferred(d) != HEVC_OK) return HEVC_ERROR;
            if (d->stream >= 0) hevcr_stream_destroy(e, d->stream);
            d->stream = -1;
            if (hevcr_stream_create(e, ev.width, ev.height, ev.bit_depth, ev.dpb_slots, &d->stream) != HEVCR_OK)
                return fail(d, std::string("engine stream: ") + hevcr_last_error());
            d->sequence = ev.sequence;
        }
        if (ev.kind == HEVCR_EVENT_DECODE) {
            for (const auto& q : d->deferred)
                if (q.slot == ev.slot) {  // the slot is about to be rewritten: take its picture out first
                    if (download_deferred(d) != HEVC_OK) return HEVC_ERROR;
                    break;
                }
            if (hevcr_decode_frame(e, d->stream, ev.syntax, ev.slot) != HEVCR_OK)
                return fail(d, std::string("engine: ") + hevcr_last_error());
            continue;
        }
        auto pic = std::make_unique<OutPicture>();
        pic->width = ev.width;
        pic->height = ev.height;
        pic->bit_depth = ev.bit_depth;
        pic->poc = ev.poc;
        std::memcpy(pic->crop, ev.crop, sizeof(pic->crop));
        d->deferred.push_back({pic.get(), ev.slot});
        sink.push_back(std::move(pic));
    }
}

// Events, then the downloads of what they output: the pictures in `sink` are
// complete when this returns HEVC_OK.
int settle(HEVCDecoder* d, std::vector<std::unique_ptr<OutPicture>>& sink) {
    int rc = run_events(d, sink);
    std::lock_guard<std::mutex> g(d->shared->m);
    if (download_deferred(d) != HEVC_OK) rc = HEVC_ERROR;
    if (rc != HEVC_OK) {  // pictures whose samples never arrived are not handed out
        while (!sink.empty() && !sink.back()->buf) sink.pop_back();
    }
    return rc;
}

// Feeds `data` in chunks, running the events of each chunk before parsing the next.
int feed_chunks(HEVCDecoder* d, const uint8_t* data, size_t size, std::vector<std::unique_ptr<OutPicture>>& sink) {
    for (size_t at = 0; at < size; at += kFeedChunk) {
        const size_t n = size - at < kFeedChunk ? size - at : kFeedChunk;
        if (hevcr_parser_feed(d->parser, data + at, n) < 0) {
            settle(d, sink);  // pictures before the error stay readable, as in hevc_api.cpp:31-36
            return fail(d, std::string("parse: ") + hevcr_parser_error(d->parser));
        }
        if (run_events(d, sink) != HEVC_OK) {
            settle(d, sink);
            return HEVC_ERROR;
        }
    }
    return HEVC_OK;
}

void fill_frame(const OutPicture& p, HEVCFrame* f) {
    // hevc_api.cpp:49-77: plane pointers offset by the conformance window
    const int l = p.crop[0], r = p.crop[1], t = p.crop[2], b = p.crop[3];
    f->width = p.width - l - r;
    f->height = p.height - t - b;
    f->stride_y = p.width;
    f->stride_c = p.width / 2;
    f->y = p.plane(0) + static_cast<size_t>(t) * p.width + l;
    f->cb = p.plane(1) + static_cast<size_t>(t / 2) * (p.width / 2) + l / 2;
    f->cr = p.plane(2) + static_cast<size_t>(t / 2) * (p.width / 2) + l / 2;
    f->chroma_width = f->width / 2;
    f->chroma_height = f->height / 2;
    f->bit_depth = p.bit_depth;
    f->poc = p.poc;
}

int ensure(HEVCDecoder* d) {
    if (!d->shared && acquire_engine(d) != HEVC_OK) return HEVC_ERROR;
    if (!d->parser && hevcr_parser_create(&d->parser) != HEVCR_OK) return fail(d, "parser allocation");
    return HEVC_OK;
}

}  // namespace

HEVCDecoder::~HEVCDecoder() {
    if (shared) {
        {
            std::lock_guard<std::mutex> g(shared->m);
            if (stream >= 0) hevcr_stream_destroy(shared->e, stream);
        }
        release_engine(this);
    }
    if (parser) hevcr_parser_destroy(parser);
}

// Every entry point catches what the C++ side may throw (std::bad_alloc from
// the containers above) at the C boundary, as hevc_api.cpp:30-41 does.
extern "C" {

HEVCDecoder* hevc_decoder_create(void) { return new (std::nothrow) HEVCDecoder(); }

void hevc_decoder_destroy(HEVCDecoder* dec) {
    try {
        delete dec;
    } catch (...) {
    }
}

int hevc_decoder_decode(HEVCDecoder* dec, const uint8_t* data, size_t size) {
    if (!dec || !data || size == 0) return HEVC_ERROR;
    try {
        if (ensure(dec)) return HEVC_ERROR;
        dec->output.clear();
        if (feed_chunks(dec, data, size, dec->output) != HEVC_OK)
            return dec->output.empty() ? HEVC_ERROR : HEVC_OK;  // partial output, as hevc_api.cpp:36-40
        if (hevcr_parser_flush(dec->parser) < 0) {
            settle(dec, dec->output);
            dec->error = std::string("parse: ") + hevcr_parser_error(dec->parser);
            return dec->output.empty() ? HEVC_ERROR : HEVC_OK;
        }
        settle(dec, dec->output);  // on an engine error the pictures before it stay readable
        return dec->output.empty() ? HEVC_ERROR : HEVC_OK;
    } catch (...) {
        dec->error = "exception in hevc_decoder_decode";
        while (!dec->output.empty() && !dec->output.back()->buf) dec->output.pop_back();
        return dec->output.empty() ? HEVC_ERROR : HEVC_OK;
    }
}

int hevc_decoder_get_frame_count(HEVCDecoder* dec) { return dec ? static_cast<int>(dec->output.size()) : 0; }

int hevc_decoder_get_frame(HEVCDecoder* dec, int index, HEVCFrame* frame) {
    if (!dec || !frame || index < 0 || index >= static_cast<int>(dec->output.size())) return HEVC_ERROR;
    fill_frame(*dec->output[static_cast<size_t>(index)], frame);
    return HEVC_OK;
}

// hevc_api.cpp:153-166: the first picture of the last hevc_decoder_decode
// call, HEVC_ERROR when it output none; profile and level are reported as 0.
int hevc_decoder_get_info(HEVCDecoder* dec, HEVCStreamInfo* info) {
    if (!dec || !info || dec->output.empty()) return HEVC_ERROR;
    const OutPicture& p = *dec->output[0];
    info->width = p.width - p.crop[0] - p.crop[1];
    info->height = p.height - p.crop[2] - p.crop[3];
    info->bit_depth = p.bit_depth;
    info->chroma_format = 1;
    info->profile = 0;
    info->level = 0;
    return HEVC_OK;
}

int hevc_decoder_feed(HEVCDecoder* dec, const uint8_t* data, size_t size) {
    if (!dec || !data || size == 0) return HEVC_ERROR;
    try {
        if (ensure(dec)) return HEVC_ERROR;
        if (feed_chunks(dec, data, size, dec->pending) != HEVC_OK) return HEVC_ERROR;
        return settle(dec, dec->pending);
    } catch (...) {
        dec->error = "exception in hevc_decoder_feed";
        return HEVC_ERROR;
    }
}

int hevc_decoder_drain(HEVCDecoder* dec, int* count) {
    if (!dec || !count) return HEVC_ERROR;
    try {
        dec->drained = std::move(dec->pending);
        dec->pending.clear();
        *count = static_cast<int>(dec->drained.size());
        return HEVC_OK;
    } catch (...) {
        *count = 0;
        return HEVC_ERROR;
    }
}

int hevc_decoder_get_drained_frame(HEVCDecoder* dec, int index, HEVCFrame* frame) {
    if (!dec || !frame || index < 0 || index >= static_cast<int>(dec->drained.size())) return HEVC_ERROR;
    fill_frame(*dec->drained[static_cast<size_t>(index)], frame);
    return HEVC_OK;
}

int hevc_decoder_flush(HEVCDecoder* dec) {
    if (!dec) return HEVC_ERROR;
    try {
        if (ensure(dec)) return HEVC_ERROR;
        if (hevcr_parser_flush(dec->parser) < 0) {
            settle(dec, dec->pending);
            return fail(dec, std::string("parse: ") + hevcr_parser_error(dec->parser));
        }
        const int rc = settle(dec, dec->pending);
        // as in hevc_api.cpp:143-151, the flushed pictures are readable through
        // hevc_decoder_get_drained_frame at once (with any not drained before)
        dec->drained = std::move(dec->pending);
        dec->pending.clear();
        return rc;
    } catch (...) {
        dec->error = "exception in hevc_decoder_flush";
        return HEVC_ERROR;
    }
}

const char* hevcr_decoder_error(const HEVCDecoder* dec) { return dec ? dec->error.c_str() : "null decoder"; }

}  // extern "C"
