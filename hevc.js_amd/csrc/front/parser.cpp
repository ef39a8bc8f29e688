// Host front end: NAL units to pictures. POC (8.3.1), reference picture set
// and marking (8.3.2), reference lists (8.3.4), DPB slots and output order
// (C.5.2 bumping), and the C-ABI of include/hevcr_parse.h. The reference's
// counterparts are src/decoding/decoder.cpp and src/decoding/dpb.cpp.
#include <algorithm>
#include <cstring>
#include <deque>
#include <memory>
#include <string>
#include <vector>

#include "hevcr.h"
#include "hevcr_parse.h"
#include "params.hpp"
#include "slice_data.hpp"

using namespace hevcf;

// A picture whose slice segments are still arriving: the state its slice
// parser refers to (kept at stable addresses until the picture is finished).
struct PendingPicture {
    std::shared_ptr<Sps> sps;
    std::shared_ptr<Pps> pps;
    std::deque<SliceHeader> headers;            // every segment's header, in order
    const SliceHeader* independent = nullptr;   // the last independent segment's
    RefLists refs;
    std::unique_ptr<DpbPic> cur;
    std::unique_ptr<PictureRecords> rec;
    std::unique_ptr<PictureDecoder> dec;
    int nal_type = 0;
    int poc = 0;
};

struct hevcr_parser {
    SpsTable sps{};
    PpsTable pps{};
    std::unique_ptr<PendingPicture> pending;
    std::vector<uint8_t> buf;              // bytes not yet parsed (from the last start code on)
    std::vector<std::unique_ptr<DpbPic>> dpb;
    std::deque<hevcr_parser_event> events;
    std::deque<std::unique_ptr<PictureRecords>> records;  // kept alive for queued DECODE events
    std::deque<hevcr_frame_syntax> syntaxes;
    std::unique_ptr<PictureRecords> current_records;      // of the last DECODE event handed out
    hevcr_frame_syntax current_syntax{};
    std::string error;
    int status = 0;
    // POC state
    int prev_tid0_lsb = 0, prev_tid0_msb = 0;
    bool first_picture = true, after_eos = false, skip_rasl = false;
    int cvs = 0, decode_index = 0, sequence = 0;
    int seq_w = -1, seq_h = -1, seq_bd = -1, seq_slots = -1;
    const Sps* active_sps = nullptr;
};

namespace {

int fail(hevcr_parser* p, int code, const char* msg) {
    p->status = code;
    p->error = msg;
    return code;
}

bool is_irap(int t) { return t >= NAL_BLA_W_LP && t <= 23; }
bool is_idr(int t) { return t == NAL_IDR_W_RADL || t == NAL_IDR_N_LP; }
bool sub_layer_nonref(int t) { return t <= 14 && (t % 2) == 0; }

void queue_output(hevcr_parser* P, DpbPic* d) {
    hevcr_parser_event ev{};
    ev.kind = HEVCR_EVENT_OUTPUT;
    ev.sequence = P->sequence;
    ev.slot = d->slot;
    ev.poc = d->poc;
    ev.width = P->seq_w;
    ev.height = P->seq_h;
    ev.bit_depth = P->seq_bd;
    ev.dpb_slots = P->seq_slots;
    const Sps& s = *P->active_sps;
    for (int k = 0; k < 4; k++) ev.crop[k] = s.conf[k];
    ev.decode_index = d->id;
    ev.profile_idc = s.profile_idc;
    ev.level_idc = s.level_idc;
    P->events.push_back(ev);
    d->output_needed = false;
}

// Outputs the picture with the smallest POC among those waiting (C.5.2.4).
bool bump(hevcr_parser* P) {
    DpbPic* best = nullptr;
    for (auto& d : P->dpb)
        if (d->output_needed && (!best || d->poc < best->poc)) best = d.get();
    if (!best) return false;
    queue_output(P, best);
    return true;
}

void drop_unused(hevcr_parser* P) {
    P->dpb.erase(std::remove_if(P->dpb.begin(), P->dpb.end(),
                                [](const std::unique_ptr<DpbPic>& d) {
                                    return !d->output_needed && !d->short_term && !d->long_term;
                                }),
                 P->dpb.end());
}

int count_output(hevcr_parser* P) {
    int n = 0;
    for (auto& d : P->dpb) n += d->output_needed;
    return n;
}

// The first slice segment of a picture: POC, reference picture set and
// marking, bumping before the picture, reference lists and the picture's
// DPB slot; then its slice parser, which the segments feed.
void start_picture(hevcr_parser* P, const Nal& nal, const SliceHeader& first) {
    const SliceHeader& sh = first;
    auto pps_ptr = P->pps[static_cast<size_t>(sh.pps_id)];
    auto sps_ptr = P->sps[static_cast<size_t>(pps_ptr->sps_id)];
    const Pps& pps = *pps_ptr;
    const Sps& sps = *sps_ptr;
    if (sps.chroma_format_idc != 1) throw ParseError(-6, "chroma format other than 4:2:0");
    if (sps.bit_depth != sps.bit_depth_c || sps.bit_depth > 10) throw ParseError(-6, "bit depth");
    if ((sps.width & 7) || (sps.height & 7)) throw ParseError(-6, "picture size not a multiple of 8");
    const int t = nal.type;
    const bool irap = is_irap(t);
    const bool no_rasl_output = irap && (is_idr(t) || t <= NAL_BLA_N_LP || P->first_picture || P->after_eos);
    if (irap) P->skip_rasl = no_rasl_output;
    if ((t == NAL_RASL_N || t == NAL_RASL_R) && P->skip_rasl) {  // not decodable (8.1.3): its segments are skipped
        P->pending.reset();
        return;
    }
    // POC (8.3.1)
    const int max_lsb = 1 << sps.log2_max_poc_lsb;
    int msb;
    if (irap && no_rasl_output) {
        msb = 0;
    } else {
        const int lsb = static_cast<int>(sh.poc_lsb), pl = P->prev_tid0_lsb, pm = P->prev_tid0_msb;
        if (lsb < pl && pl - lsb >= max_lsb / 2)
            msb = pm + max_lsb;
        else if (lsb > pl && lsb - pl > max_lsb / 2)
            msb = pm - max_lsb;
        else
            msb = pm;
    }
    const int poc = msb + static_cast<int>(sh.poc_lsb);
    if (nal.tid == 0 && !(t >= NAL_RADL_N && t <= NAL_RASL_R) && !sub_layer_nonref(t)) {
        P->prev_tid0_lsb = static_cast<int>(sh.poc_lsb);
        P->prev_tid0_msb = msb;
    }
    // new coded video sequence: the previous one's pictures all leave
    if (irap && no_rasl_output) {
        for (auto& d : P->dpb) d->short_term = d->long_term = false;
        while (bump(P)) {
        }
        drop_unused(P);
        P->cvs++;
    }
    // sequence geometry and the engine's DPB size
    const int slots = std::min(sps.max_dec_pic_buffering + 2, 17);
    if (sps.width != P->seq_w || sps.height != P->seq_h || sps.bit_depth != P->seq_bd || slots > P->seq_slots) {
        if (!P->dpb.empty()) {
            for (auto& d : P->dpb) d->short_term = d->long_term = false;
            while (bump(P)) {
            }
            drop_unused(P);
        }
        P->sequence++;
        P->seq_w = sps.width;
        P->seq_h = sps.height;
        P->seq_bd = sps.bit_depth;
        P->seq_slots = slots;
    }
    P->active_sps = P->sps[static_cast<size_t>(pps.sps_id)].get();
    // reference picture set (8.3.2)
    std::vector<DpbPic*> st_before, st_after, lt_curr;
    std::vector<DpbPic*> keep_st, keep_lt;
    if (!is_idr(t)) {
        auto find_st = [&](int p) -> DpbPic* {
            for (auto& d : P->dpb)
                if (d->short_term && d->poc == p && d->cvs == P->cvs) return d.get();
            return nullptr;
        };
        const StRps& r = sh.st_rps;
        for (int i = 0; i < r.num_negative; i++) {
            DpbPic* d = find_st(poc + r.delta_s0[i]);
            if (r.used_s0[i] && !d) throw ParseError(-1, "missing short-term reference picture");
            if (d) (r.used_s0[i] ? st_before : keep_st).push_back(d);
        }
        for (int i = 0; i < r.num_positive; i++) {
            DpbPic* d = find_st(poc + r.delta_s1[i]);
            if (r.used_s1[i] && !d) throw ParseError(-1, "missing short-term reference picture");
            if (d) (r.used_s1[i] ? st_after : keep_st).push_back(d);
        }
        uint32_t msb_cycle = 0;
        for (int i = 0; i < sh.num_lt; i++) {
            // DeltaPocMsbCycleLt (7-52): accumulated within the SPS and the slice entries separately
            msb_cycle = (i == 0 || i == sh.num_lt_sps) ? sh.lt_msb_cycle[i] : msb_cycle + sh.lt_msb_cycle[i];
            int p = static_cast<int>(sh.lt_poc_lsb[i]);
            bool full = sh.lt_msb_present[i];
            if (full) p += poc - static_cast<int>(msb_cycle) * max_lsb - (poc & (max_lsb - 1));
            DpbPic* d = nullptr;
            for (auto& q : P->dpb)
                if ((q->short_term || q->long_term) && q->cvs == P->cvs &&
                    (full ? q->poc == p : (q->poc & (max_lsb - 1)) == p))
                    d = q.get();
            if (sh.lt_used[i] && !d) throw ParseError(-1, "missing long-term reference picture");
            if (d) (sh.lt_used[i] ? lt_curr : keep_lt).push_back(d);
        }
    }
    for (auto& d : P->dpb) {
        const bool in_st = std::count(st_before.begin(), st_before.end(), d.get()) ||
                           std::count(st_after.begin(), st_after.end(), d.get()) ||
                           std::count(keep_st.begin(), keep_st.end(), d.get());
        const bool in_lt = std::count(lt_curr.begin(), lt_curr.end(), d.get()) ||
                           std::count(keep_lt.begin(), keep_lt.end(), d.get());
        d->long_term = in_lt;
        d->short_term = in_st && !in_lt;
    }
    drop_unused(P);
    // bumping before the current picture (C.5.2.2)
    while (true) {
        const int waiting = count_output(P);
        const bool reorder = waiting > sps.max_num_reorder;
        const bool full = static_cast<int>(P->dpb.size()) >= sps.max_dec_pic_buffering;
        if (!(reorder || full) || !bump(P)) break;
        drop_unused(P);
    }
    // reference lists (8.3.4), built once from the first segment (the reference
    // does the same, decoder.cpp:108-115)
    auto pend = std::make_unique<PendingPicture>();
    pend->sps = sps_ptr;
    pend->pps = pps_ptr;
    pend->nal_type = t;
    pend->poc = poc;
    RefLists& refs = pend->refs;
    if (sh.slice_type != HEVCR_SLICE_I) {
        const int total = static_cast<int>(st_before.size() + st_after.size() + lt_curr.size());
        if (total == 0) throw ParseError(-1, "P/B slice without reference pictures");
        for (int l = 0; l < (sh.slice_type == HEVCR_SLICE_B ? 2 : 1); l++) {
            std::vector<std::pair<DpbPic*, bool>> temp;
            const int n = std::max(sh.num_ref_idx[l], total);
            const auto& first = l ? st_after : st_before;
            const auto& second = l ? st_before : st_after;
            while (static_cast<int>(temp.size()) < n) {
                for (DpbPic* d : first)
                    if (static_cast<int>(temp.size()) < n) temp.emplace_back(d, false);
                for (DpbPic* d : second)
                    if (static_cast<int>(temp.size()) < n) temp.emplace_back(d, false);
                for (DpbPic* d : lt_curr)
                    if (static_cast<int>(temp.size()) < n) temp.emplace_back(d, true);
            }
            refs.num[l] = sh.num_ref_idx[l];
            for (int i = 0; i < sh.num_ref_idx[l]; i++) {
                const int k = sh.list_mod[l] ? static_cast<int>(sh.list_entry[l][i]) : i;
                if (k >= static_cast<int>(temp.size())) throw ParseError(-1, "list_entry");
                refs.pic[l][i] = temp[static_cast<size_t>(k)].first;
                refs.poc[l][i] = temp[static_cast<size_t>(k)].first->poc;
                refs.lt[l][i] = temp[static_cast<size_t>(k)].second;
            }
        }
    }
    // slot of the current picture
    auto cur = std::make_unique<DpbPic>();
    cur->id = P->decode_index++;
    cur->poc = poc;
    cur->cvs = P->cvs;
    cur->width = sps.width;
    cur->height = sps.height;
    {
        std::vector<bool> used(static_cast<size_t>(P->seq_slots), false);
        for (auto& d : P->dpb) used[static_cast<size_t>(d->slot)] = true;
        int s = 0;
        while (s < P->seq_slots && used[static_cast<size_t>(s)]) s++;
        if (s == P->seq_slots) throw ParseError(-1, "decoded picture buffer overflow");
        cur->slot = s;
    }
    pend->cur = std::move(cur);
    pend->rec = std::make_unique<PictureRecords>();
    pend->headers.push_back(first);
    pend->independent = &pend->headers.back();
    pend->dec = std::make_unique<PictureDecoder>(sps, pps, pend->headers.front(), pend->refs, *pend->cur, *pend->rec);
    P->pending = std::move(pend);
}

// The picture after its last slice segment: records, its DECODE event,
// the picture into the DPB and bumping after it (C.5.2.3).
void finish_picture(hevcr_parser* P) {
    std::unique_ptr<PendingPicture> pend = std::move(P->pending);
    const Sps& sps = *pend->sps;
    const Pps& pps = *pend->pps;
    const SliceHeader& sh = pend->headers.front();
    const RefLists& refs = pend->refs;
    const int t = pend->nal_type, poc = pend->poc;
    std::unique_ptr<DpbPic> cur = std::move(pend->cur);
    std::unique_ptr<PictureRecords> rec = std::move(pend->rec);
    pend->dec->finish();
    if (sps.scaling_list_enabled) {  // the PPS's lists when it carries them, else the SPS's
        rec->has_scaling = true;
        rec->scaling = pps.scaling_list_present ? pps.scaling : sps.scaling;
    }
    // frame parameters
    hevcr_frame_syntax syn{};
    hevcr_frame_params& fp = syn.p;
    fp.width = sps.width;
    fp.height = sps.height;
    fp.bit_depth = sps.bit_depth;
    fp.ctb_log2 = sps.log2_ctb;
    fp.min_cb_log2 = sps.log2_min_cb;
    fp.slice_type = sh.slice_type;
    fp.poc = poc;
    fp.strong_intra_smoothing = sps.strong_intra_smoothing;
    fp.intra_smoothing_disabled = 0;
    fp.sao_enabled = sps.sao;
    fp.pps_cb_qp_offset = pps.cb_qp_offset;
    fp.pps_cr_qp_offset = pps.cr_qp_offset;
    fp.slice_cb_qp_offset = sh.cb_qp_offset;
    fp.slice_cr_qp_offset = sh.cr_qp_offset;
    fp.deblocking_disabled = sh.deblocking_disabled;
    fp.beta_offset_div2 = sh.beta_offset_div2;
    fp.tc_offset_div2 = sh.tc_offset_div2;
    fp.weighted_pred = (sh.slice_type == HEVCR_SLICE_P && pps.weighted_pred) ||
                       (sh.slice_type == HEVCR_SLICE_B && pps.weighted_bipred);
    fp.luma_log2_wd = sh.luma_log2_wd;
    fp.chroma_log2_wd = sh.chroma_log2_wd;
    for (int l = 0; l < 2; l++) {
        fp.num_ref[l] = refs.num[l];
        for (int i = 0; i < refs.num[l]; i++) {
            fp.ref_slot[l][i] = refs.pic[l][i]->slot;
            fp.ref_poc[l][i] = refs.poc[l][i];
            for (int c = 0; c < 3; c++) {
                fp.weights[l][i].w[c] = sh.weights[l][i].w[c];
                fp.weights[l][i].o[c] = sh.weights[l][i].o[c];
            }
        }
    }
    hevcr_parser_event ev{};
    ev.kind = HEVCR_EVENT_DECODE;
    ev.sequence = P->sequence;
    ev.slot = cur->slot;
    ev.poc = poc;
    ev.width = sps.width;
    ev.height = sps.height;
    ev.bit_depth = sps.bit_depth;
    ev.dpb_slots = P->seq_slots;
    for (int k = 0; k < 4; k++) ev.crop[k] = sps.conf[k];
    ev.decode_index = cur->id;
    ev.profile_idc = sps.profile_idc;
    ev.level_idc = sps.level_idc;
    ev.syntax = nullptr;  // bound when handed out
    P->events.push_back(ev);
    P->syntaxes.push_back(syn);
    P->records.push_back(std::move(rec));
    // the current picture joins the DPB; bumping after it (C.5.2.3)
    cur->short_term = true;
    cur->output_needed = sh.pic_output && !(t == NAL_RASL_N || t == NAL_RASL_R ? P->skip_rasl : false);
    P->dpb.push_back(std::move(cur));
    while (count_output(P) > sps.max_num_reorder && bump(P)) {
    }
    drop_unused(P);
    P->first_picture = false;
    P->after_eos = false;
}

// A slice segment NAL unit: the first of a picture starts it (finishing a
// picture still open), the others feed the open picture; a picture is
// finished as soon as its segments cover it.
void slice_segment(hevcr_parser* P, const Nal& nal, Bits& b) {
    SliceHeader sh;
    const SliceHeader* independent = P->pending ? P->pending->independent : nullptr;
    parse_slice_header(b, nal.type, P->sps, P->pps, sh, independent);
    if (sh.first_slice) {
        if (P->pending) throw ParseError(-1, "picture with CTBs in no slice segment");
        start_picture(P, nal, sh);
        if (!P->pending) return;  // skipped (RASL)
    } else {
        if (!P->pending) throw ParseError(-1, "slice segment without the first segment of its picture");
        if (sh.pps_id != P->pending->headers.front().pps_id) throw ParseError(-1, "slice segments of one picture with different PPSs");
        P->pending->headers.push_back(sh);
        if (!sh.dependent) P->pending->independent = &P->pending->headers.back();
    }
    PendingPicture& pp = *P->pending;
    pp.dec->segment(pp.headers.back(), nal.rbsp, nal.epb);
    if (pp.dec->complete()) finish_picture(P);
}


void handle_nal(hevcr_parser* P, const uint8_t* p, size_t n) {
    Nal nal;
    if (!make_nal(p, n, nal)) return;
    if (nal.layer != 0) return;
    Bits b(nal.rbsp.data(), nal.rbsp.size());
    switch (nal.type) {
        case NAL_SPS: {
            auto s = std::make_shared<Sps>();
            parse_sps(b, *s);
            P->sps[static_cast<size_t>(s->id)] = s;
            return;
        }
        case NAL_PPS: {
            auto q = std::make_shared<Pps>();
            parse_pps(b, *q, P->sps);
            P->pps[static_cast<size_t>(q->id)] = q;
            return;
        }
        case NAL_EOS:
            if (P->pending) throw ParseError(-1, "picture with CTBs in no slice segment");
            P->after_eos = true;
            return;
        default:
            break;
    }
    if (nal.type > 21 && nal.type != 22 && nal.type != 23) return;  // non-VCL
    if (nal.type >= 10 && nal.type <= 15) return;                  // reserved non-IRAP VCL
    if (nal.type == 22 || nal.type == 23) return;                  // reserved IRAP
    slice_segment(P, nal, b);
}

// Parses every complete NAL unit in the buffer (all of it when `final`).
int drain(hevcr_parser* P, bool final) {
    if (P->status) return P->status;
    std::vector<uint8_t>& B = P->buf;
    size_t pos = next_start(B.data(), 0, B.size());
    size_t consumed = 0;
    try {
        while (pos < B.size()) {
            size_t nxt = next_start(B.data(), pos, B.size());
            if (nxt >= B.size() && !final) break;  // the last unit may still grow
            size_t end = nxt >= B.size() ? B.size() : nxt - 3;
            while (end > pos && B[end - 1] == 0) end--;  // trailing_zero_8bits / zero_byte
            handle_nal(P, B.data() + pos, end - pos);
            consumed = nxt >= B.size() ? B.size() : nxt - 3;
            pos = nxt;
        }
    } catch (const ParseError& e) {
        return fail(P, e.code, e.what());
    } catch (const std::exception& e) {
        return fail(P, HEVCR_ERROR, e.what());
    }
    if (final) consumed = B.size();
    B.erase(B.begin(), B.begin() + static_cast<std::ptrdiff_t>(consumed));
    return HEVCR_OK;
}

}  // namespace

extern "C" {

int hevcr_parser_create(hevcr_parser** out) {
    if (!out) return HEVCR_EINVAL;
    *out = new (std::nothrow) hevcr_parser();
    return *out ? HEVCR_OK : HEVCR_ENOMEM;
}

void hevcr_parser_destroy(hevcr_parser* p) { delete p; }

int hevcr_parser_feed(hevcr_parser* p, const uint8_t* data, size_t size) {
    if (!p || (size && !data)) return HEVCR_EINVAL;
    if (p->status) return p->status;
    p->buf.insert(p->buf.end(), data, data + size);
    return drain(p, false);
}

int hevcr_parser_flush(hevcr_parser* p) {
    if (!p) return HEVCR_EINVAL;
    int rc = drain(p, true);
    if (rc) return rc;
    for (auto& d : p->dpb) d->short_term = d->long_term = false;
    while (bump(p)) {
    }
    drop_unused(p);
    return HEVCR_OK;
}

int hevcr_parser_next(hevcr_parser* p, hevcr_parser_event* ev) {
    if (!p || !ev) return HEVCR_EINVAL;
    if (p->events.empty()) return p->status ? p->status : 0;
    *ev = p->events.front();
    p->events.pop_front();
    if (ev->kind == HEVCR_EVENT_DECODE) {
        p->current_records = std::move(p->records.front());
        p->records.pop_front();
        p->current_syntax = p->syntaxes.front();
        p->syntaxes.pop_front();
        PictureRecords& r = *p->current_records;
        hevcr_frame_syntax& s = p->current_syntax;
        s.cus = r.cus.data();
        s.num_cus = static_cast<uint32_t>(r.cus.size());
        s.pus = r.pus.data();
        s.num_pus = static_cast<uint32_t>(r.pus.size());
        s.tus = r.tus.data();
        s.num_tus = static_cast<uint32_t>(r.tus.size());
        s.coeffs = r.coeffs.data();
        s.num_coeffs = r.coeffs.size();
        s.sao = r.sao.data();
        s.scaling = r.has_scaling ? &r.scaling : nullptr;
        if (r.slices.size() > 1) {
            s.slices = r.slices.data();
            s.num_slices = static_cast<uint32_t>(r.slices.size());
            s.ctb_slice = r.ctb_slice.data();
        }
        s.tiles = r.has_tiles ? &r.tiles : nullptr;
        s.pcm_loop_filter_disabled = r.pcm_loop_filter_disabled;
        ev->syntax = &s;
    }
    return 1;
}

const char* hevcr_parser_error(const hevcr_parser* p) { return p ? p->error.c_str() : "null parser"; }

}  // extern "C"
