// Test stream writer: writes HEVC Annex-B streams of random content that
// exercise slice segments (dependent ones included), tiles, wavefront entry
// points, PCM, transquant bypass, weighted prediction per segment and the
// other picture/slice-level options, for the coded-stream golden cases.
//
// The slice data is written by running the front end's own slice parser
// (hevc.js_amd/csrc/front/slice_decoder.hpp) over an engine that, instead of
// decoding a bin, chooses it (at random, with a probability per context) and
// arithmetic-codes it (H.264 9.3.4.2-9.3.4.5, the encoder counterpart of
// H.265 9.3.4.3): every bin the parser asks for is written in the order it
// asks, so the parser reads the stream back exactly as it was written.
// Parameter sets and slice headers are written from H.265 7.3.2-7.3.6 and
// parsed back with the front end's header parser before use. The goldens are
// what the reference decoder (oracle/_ref/hevc-decode) makes of these streams;
// whether the front end and the engine reproduce them is what the tests check.
//
//   stream-writer OUT.265 key=value ...     (keys: see Config)
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "bits.hpp"
#include "params.hpp"
#include "slice_data.hpp"
#include "slice_decoder.hpp"

using namespace hevcf;

namespace {

// ---------------------------------------------------------------------------
// configuration

struct Config {
    std::map<std::string, std::string> kv;
    std::map<std::string, bool> used;
    double get(const char* k, double d) {
        used[k] = true;
        auto it = kv.find(k);
        return it == kv.end() ? d : std::atof(it->second.c_str());
    }
    int geti(const char* k, int d) { return static_cast<int>(get(k, d)); }
    std::string gets(const char* k, const char* d) {
        used[k] = true;
        auto it = kv.find(k);
        return it == kv.end() ? std::string(d) : it->second;
    }
};

// xorshift64*: fixed across compilers and libraries
struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed * 0x9E3779B97F4A7C15ull + 0x1234567ull) {
        if (!s) s = 1;
        for (int i = 0; i < 4; i++) next();
    }
    uint64_t next() {
        s ^= s >> 12;
        s ^= s << 25;
        s ^= s >> 27;
        return s * 2685821657736338717ull;
    }
    int below(int n) { return n <= 1 ? 0 : static_cast<int>(next() % static_cast<uint64_t>(n)); }
    int range(int lo, int hi) { return lo + below(hi - lo + 1); }  // inclusive
    bool coin(double p) { return static_cast<double>(next() >> 11) * (1.0 / 9007199254740992.0) < p; }
};

int ceil_log2(int v) {
    int n = 0;
    while ((1 << n) < v) n++;
    return n;
}

// ---------------------------------------------------------------------------
// bit writer

struct BitWriter {
    std::vector<uint8_t> buf;
    size_t pos = 0;  // bits written
    void bit(int b) {
        if ((pos & 7) == 0) buf.push_back(0);
        if (b) buf.back() = static_cast<uint8_t>(buf.back() | (0x80u >> (pos & 7)));
        pos++;
    }
    void u(uint32_t v, int n) {
        for (int i = n - 1; i >= 0; i--) bit(static_cast<int>((v >> i) & 1u));
    }
    void flag(bool b) { bit(b ? 1 : 0); }
    void ue(uint32_t v) {
        const uint64_t x = static_cast<uint64_t>(v) + 1;
        int len = 0;
        while ((x >> len) > 1) len++;
        for (int i = 0; i < len; i++) bit(0);
        for (int i = len; i >= 0; i--) bit(static_cast<int>((x >> i) & 1u));
    }
    void se(int32_t v) { ue(v > 0 ? static_cast<uint32_t>(2 * v - 1) : static_cast<uint32_t>(-2 * v)); }
    void align_zero() {
        while (pos & 7) bit(0);
    }
    void trailing() {  // rbsp_trailing_bits() / byte_alignment(): a one, then zeros
        bit(1);
        align_zero();
    }
};

// ---------------------------------------------------------------------------
// arithmetic encoder (H.264 9.3.4.2-9.3.4.5; the same range and context
// state machine as the decoder of H.265 9.3.4.3)

class ArithEncoder {
public:
    void start(BitWriter* w) {
        w_ = w;
        low_ = 0;
        range_ = 510;
        first_ = true;
        outstanding_ = 0;
    }
    void decision(uint8_t& s, int bin) {
        const int state = s >> 1, mps = s & 1;
        const uint32_t lps = Cabac::kLps[state][(range_ >> 6) & 3];
        range_ -= lps;
        if (bin != mps) {
            low_ += range_;
            range_ = lps;
            s = static_cast<uint8_t>((Cabac::kNextLps[state] << 1) | (state == 0 ? 1 - mps : mps));
        } else {
            s = static_cast<uint8_t>((Cabac::kNextMps[state] << 1) | mps);
        }
        renorm();
    }
    void bypass(int bin) {
        low_ <<= 1;
        if (bin) low_ += range_;
        if (low_ >= 1024) {
            put(1);
            low_ -= 1024;
        } else if (low_ < 512) {
            put(0);
        } else {
            low_ -= 512;
            outstanding_++;
        }
    }
    // A terminating bin; a one flushes the coder, whose last bit written is
    // the one that ends the codeword (stop / alignment bit).
    void terminate(int bin) {
        range_ -= 2;
        if (bin) {
            low_ += range_;
            range_ = 2;
            renorm();
            put(static_cast<int>((low_ >> 9) & 1));
            w_->u(((low_ >> 7) & 3) | 1, 2);
        } else {
            renorm();
        }
    }

private:
    void renorm() {
        while (range_ < 256) {
            if (low_ < 256) {
                put(0);
            } else if (low_ >= 512) {
                low_ -= 512;
                put(1);
            } else {
                low_ -= 256;
                outstanding_++;
            }
            range_ <<= 1;
            low_ <<= 1;
        }
    }
    void put(int b) {
        if (first_)
            first_ = false;
        else
            w_->bit(b);
        for (; outstanding_ > 0; outstanding_--) w_->bit(1 - b);
    }
    BitWriter* w_ = nullptr;
    uint32_t low_ = 0, range_ = 510;
    bool first_ = true;
    int outstanding_ = 0;
};

// ---------------------------------------------------------------------------
// the engine the slice parser runs on: chooses each bin, codes it

struct BinOdds {
    double p[C_NUM];
    double bypass = 0.5, prefix = 0.35, pcm = 0.3;
    int prefix_cap = 4;
};

class BinChooser {
public:
    BinChooser(const BinOdds& odds, Rng& rng, bool tiles) : odds_(odds), rng_(rng), tiles_(tiles) {}

    void begin_segment(int ctbs) {
        seg_ctbs_ = ctbs;
        ctbs_ = 0;
        subs_.clear();
    }
    std::vector<BitWriter>& substreams() { return subs_; }

    // --- the engine interface the slice parser uses
    void init_contexts(int init_type, int slice_qp) { model_.init_contexts(init_type, slice_qp); }
    void save(uint8_t* out) const { model_.save(out); }
    void load(const uint8_t* in) { model_.load(in); }
    void start(const uint8_t*, const uint8_t*) { new_substream(); }
    int decode(int ci) {
        const int b = rng_.coin(odds_.p[ci]) ? 1 : 0;
        enc_.decision(model_.ctx[ci], b);
        run_ = 0;
        return b;
    }
    int bypass() {
        const int b = rng_.coin(odds_.bypass) ? 1 : 0;
        enc_.bypass(b);
        run_ = 0;
        return b;
    }
    int bypass_prefix() {
        const int b = (run_ < odds_.prefix_cap && rng_.coin(odds_.prefix)) ? 1 : 0;
        enc_.bypass(b);
        run_ = b ? run_ + 1 : 0;
        return b;
    }
    uint32_t bypass_bits(int n) {
        uint32_t v = 0;
        for (int i = 0; i < n; i++) v = (v << 1) | static_cast<uint32_t>(bypass());
        return v;
    }
    int terminate(int role) {
        int b = 0;
        if (role == Cabac::kEndOfSlice)
            b = ++ctbs_ == seg_ctbs_ ? 1 : 0;
        else if (role == Cabac::kEndOfSubset)
            b = 1;
        else
            b = rng_.coin(odds_.pcm) ? 1 : 0;
        enc_.terminate(b);
        if (b) {
            // The decoders this stream is for read the alignment that
            // follows the codeword as one bit and then zeros up to a byte
            // boundary (bitstream_reader.cpp:117-124 after
            // coding_tree.cpp:630 and :1432), one bit past the codeword's
            // last: where that codeword ends on a byte boundary they take a
            // whole byte more, which is written for them.
            BitWriter& w = subs_.back();
            const bool realigned = role == Cabac::kPcmFlag || (role == Cabac::kEndOfSubset && tiles_);
            if (realigned && (w.pos & 7) == 0) w.u(0x80, 8);
            w.align_zero();
        }
        run_ = 0;
        return b;
    }
    void restart_aligned() { new_substream(); }
    void raw_begin() {}
    uint32_t raw_bits(int n) {
        const uint32_t v = static_cast<uint32_t>(rng_.next() >> 20) & ((1u << n) - 1u);
        subs_.back().u(v, n);
        return v;
    }
    void raw_end() { enc_.start(&subs_.back()); }

private:
    void new_substream() {
        subs_.emplace_back();
        enc_.start(&subs_.back());
    }
    const BinOdds& odds_;
    Rng& rng_;
    Cabac model_;
    ArithEncoder enc_;
    std::vector<BitWriter> subs_;
    bool tiles_ = false;
    int seg_ctbs_ = 0, ctbs_ = 0, run_ = 0;
};

// ---------------------------------------------------------------------------
// NAL units

std::vector<uint8_t> nal_unit(int type, const std::vector<uint8_t>& rbsp, std::vector<size_t>* coded_pos = nullptr) {
    std::vector<uint8_t> out{0, 0, 0, 1, static_cast<uint8_t>(type << 1), 1};
    if (coded_pos) coded_pos->assign(rbsp.size() + 1, 0);
    int zeros = 0;
    size_t epb = 0;
    for (size_t i = 0; i < rbsp.size(); i++) {
        const uint8_t b = rbsp[i];
        if (zeros >= 2 && b <= 3) {
            out.push_back(3);
            epb++;
            zeros = 0;
        }
        if (coded_pos) (*coded_pos)[i] = i + epb;
        out.push_back(b);
        zeros = b == 0 ? zeros + 1 : 0;
    }
    if (coded_pos) (*coded_pos)[rbsp.size()] = rbsp.size() + epb;
    if (!rbsp.empty() && rbsp.back() == 0) out.push_back(3);
    return out;
}

void profile_tier_level(BitWriter& w, int bit_depth) {
    const int profile = bit_depth > 8 ? 2 : 1;  // Main 10 / Main
    w.u(0, 2);                                  // general_profile_space
    w.u(0, 1);                                  // general_tier_flag
    w.u(static_cast<uint32_t>(profile), 5);
    for (int j = 0; j < 32; j++) w.flag(j == profile || (profile == 1 && j == 2));
    w.flag(true);   // progressive_source
    w.flag(false);  // interlaced_source
    w.flag(false);  // non_packed_constraint
    w.flag(true);   // frame_only_constraint
    w.u(0, 32);     // 43 reserved zero bits and general_inbld_flag
    w.u(0, 12);
    w.u(120, 8);    // general_level_idc (4.0)
}

// ---------------------------------------------------------------------------
// the stream

struct Seq {
    int w, h, bd, ctb, mincb, mintb, maxtb, dinter, dintra;
    bool amp, sao, pcm, pcm_lf, tmvp, sis, scaling;
    int pcm_bits_l, pcm_bits_c, pcm_min, pcm_max;
    int crop_r, crop_b;  // conformance window, in chroma samples
    int refs, max_dec, reorder;
};

struct PicSet {
    bool dep, sdh, cip, tskip, cuqpd, slice_cqp, wp, wbp, tqb, tiles, wpp, lf_tiles, lf_slices, dbctl, dbover, dbdis,
        listmod;
    int nref_default[2], init_qp, qpd_depth, cb, cr, tcols, trows, beta, tc, pml;
    bool tuniform;
    std::vector<int> tcw, trh;
};

// Per-segment header choices (the fields a dependent segment does not code
// are ignored for it).
struct SegHeader {
    bool first = false, dependent = false;
    uint32_t address = 0;
    int slice_type = HEVCR_SLICE_I;
    uint32_t poc_lsb = 0;
    StRps rps;
    bool tmvp = false, sao_l = false, sao_c = false;
    bool override_nref = false;
    int nref[2] = {0, 0};
    bool listmod[2] = {false, false};
    uint32_t entry[2][16] = {};
    bool mvd_l1_zero = false, cabac_init = false, col_l0 = true;
    int col_idx = 0;
    int luma_denom = 0, chroma_denom_delta = 0;
    bool lw[2][16] = {}, cw[2][16] = {};
    int dlw[2][16] = {}, lo[2][16] = {}, dcw[2][16][2] = {}, dco[2][16][2] = {};
    int max_merge = 5, qp_delta = 0, cb = 0, cr = 0;
    bool db_override = false, db_disabled = false;
    int beta = 0, tc = 0;
    bool lf_slices = false;
    std::vector<uint32_t> entry_sizes;
};

void write_sps(BitWriter& w, const Seq& s) {
    w.u(0, 4);
    w.u(0, 3);
    w.u(1, 1);
    profile_tier_level(w, s.bd);
    w.ue(0);
    w.ue(1);
    w.ue(static_cast<uint32_t>(s.w));
    w.ue(static_cast<uint32_t>(s.h));
    w.flag(s.crop_r || s.crop_b);
    if (s.crop_r || s.crop_b) {
        w.ue(0);
        w.ue(static_cast<uint32_t>(s.crop_r));
        w.ue(0);
        w.ue(static_cast<uint32_t>(s.crop_b));
    }
    w.ue(static_cast<uint32_t>(s.bd - 8));
    w.ue(static_cast<uint32_t>(s.bd - 8));
    w.ue(8 - 4);  // log2_max_pic_order_cnt_lsb = 8
    w.flag(true);
    w.ue(static_cast<uint32_t>(s.max_dec - 1));
    w.ue(static_cast<uint32_t>(s.reorder));
    w.ue(0);
    w.ue(static_cast<uint32_t>(s.mincb - 3));
    w.ue(static_cast<uint32_t>(s.ctb - s.mincb));
    w.ue(static_cast<uint32_t>(s.mintb - 2));
    w.ue(static_cast<uint32_t>(s.maxtb - s.mintb));
    w.ue(static_cast<uint32_t>(s.dinter));
    w.ue(static_cast<uint32_t>(s.dintra));
    w.flag(s.scaling);
    if (s.scaling) w.flag(false);  // the default lists
    w.flag(s.amp);
    w.flag(s.sao);
    w.flag(s.pcm);
    if (s.pcm) {
        w.u(static_cast<uint32_t>(s.pcm_bits_l - 1), 4);
        w.u(static_cast<uint32_t>(s.pcm_bits_c - 1), 4);
        w.ue(static_cast<uint32_t>(s.pcm_min - 3));
        w.ue(static_cast<uint32_t>(s.pcm_max - s.pcm_min));
        w.flag(s.pcm_lf);
    }
    w.ue(0);  // num_short_term_ref_pic_sets: every slice codes its own
    w.flag(false);
    w.flag(s.tmvp);
    w.flag(s.sis);
    w.flag(false);  // vui
    w.flag(false);  // extensions
    w.trailing();
}

void write_vps(BitWriter& w, const Seq& s) {
    w.u(0, 4);
    w.flag(true);
    w.flag(true);
    w.u(0, 6);
    w.u(0, 3);
    w.flag(true);
    w.u(0xffff, 16);
    profile_tier_level(w, s.bd);
    w.flag(true);
    w.ue(static_cast<uint32_t>(s.max_dec - 1));
    w.ue(static_cast<uint32_t>(s.reorder));
    w.ue(0);
    w.u(0, 6);
    w.ue(0);
    w.flag(false);
    w.flag(false);
    w.trailing();
}

void write_pps(BitWriter& w, const PicSet& p) {
    w.ue(0);
    w.ue(0);
    w.flag(p.dep);
    w.flag(false);
    w.u(0, 3);
    w.flag(p.sdh);
    w.flag(p.cip);
    w.ue(static_cast<uint32_t>(p.nref_default[0] - 1));
    w.ue(static_cast<uint32_t>(p.nref_default[1] - 1));
    w.se(p.init_qp - 26);
    w.flag(false);  // constrained_intra_pred
    w.flag(p.tskip);
    w.flag(p.cuqpd);
    if (p.cuqpd) w.ue(static_cast<uint32_t>(p.qpd_depth));
    w.se(p.cb);
    w.se(p.cr);
    w.flag(p.slice_cqp);
    w.flag(p.wp);
    w.flag(p.wbp);
    w.flag(p.tqb);
    w.flag(p.tiles);
    w.flag(p.wpp);
    if (p.tiles) {
        w.ue(static_cast<uint32_t>(p.tcols - 1));
        w.ue(static_cast<uint32_t>(p.trows - 1));
        w.flag(p.tuniform);
        if (!p.tuniform) {
            for (int i = 0; i + 1 < p.tcols; i++) w.ue(static_cast<uint32_t>(p.tcw[static_cast<size_t>(i)] - 1));
            for (int i = 0; i + 1 < p.trows; i++) w.ue(static_cast<uint32_t>(p.trh[static_cast<size_t>(i)] - 1));
        }
        w.flag(p.lf_tiles);
    }
    w.flag(p.lf_slices);
    w.flag(p.dbctl);
    if (p.dbctl) {
        w.flag(p.dbover);
        w.flag(p.dbdis);
        if (!p.dbdis) {
            w.se(p.beta);
            w.se(p.tc);
        }
    }
    w.flag(false);  // pps scaling lists
    w.flag(p.listmod);
    w.ue(static_cast<uint32_t>(p.pml - 2));
    w.flag(false);
    w.flag(false);
    w.trailing();
}

void write_slice_header(BitWriter& w, const SegHeader& h, int nal_type, const Seq& s, const PicSet& p, int pic_ctbs) {
    const bool idr = nal_type == NAL_IDR_W_RADL || nal_type == NAL_IDR_N_LP;
    w.flag(h.first);
    if (nal_type >= 16 && nal_type <= 23) w.flag(false);
    w.ue(0);
    if (!h.first) {
        if (p.dep) w.flag(h.dependent);
        w.u(h.address, ceil_log2(pic_ctbs));
    }
    if (h.dependent) {
        // what the decoders these streams are for read of a dependent
        // segment's header (slice_header.cpp:141-144, decoder.cpp:243-244):
        // nothing after the address, the slice data from the next byte
        // boundary on (7.3.6.1 would code the entry points and a
        // byte_alignment() here)
        if (w.pos & 7) w.trailing();
        return;
    }
    if (!h.dependent) {
        w.ue(static_cast<uint32_t>(h.slice_type));
        int total = 0;
        if (!idr) {
            w.u(h.poc_lsb, 8);
            w.flag(false);  // st_ref_pic_set coded here (index 0 of 0: no prediction)
            const StRps& r = h.rps;
            w.ue(static_cast<uint32_t>(r.num_negative));
            w.ue(static_cast<uint32_t>(r.num_positive));
            for (int i = 0; i < r.num_negative; i++) {
                w.ue(static_cast<uint32_t>((i ? r.delta_s0[i - 1] : 0) - r.delta_s0[i] - 1));
                w.flag(r.used_s0[i]);
                total += r.used_s0[i];
            }
            for (int i = 0; i < r.num_positive; i++) {
                w.ue(static_cast<uint32_t>(r.delta_s1[i] - (i ? r.delta_s1[i - 1] : 0) - 1));
                w.flag(r.used_s1[i]);
                total += r.used_s1[i];
            }
            if (s.tmvp) w.flag(h.tmvp);
        }
        if (s.sao) {
            w.flag(h.sao_l);
            w.flag(h.sao_c);
        }
        const bool pb = h.slice_type != HEVCR_SLICE_I, b = h.slice_type == HEVCR_SLICE_B;
        if (pb) {
            w.flag(h.override_nref);
            if (h.override_nref) {
                w.ue(static_cast<uint32_t>(h.nref[0] - 1));
                if (b) w.ue(static_cast<uint32_t>(h.nref[1] - 1));
            }
            if (p.listmod && total > 1) {
                for (int l = 0; l < (b ? 2 : 1); l++) {
                    w.flag(h.listmod[l]);
                    if (h.listmod[l])
                        for (int i = 0; i < h.nref[l]; i++) w.u(h.entry[l][i], ceil_log2(total));
                }
            }
            if (b) w.flag(h.mvd_l1_zero);
            if (p.cip) w.flag(h.cabac_init);
            if (h.tmvp) {
                if (b) w.flag(h.col_l0);
                if ((h.col_l0 && h.nref[0] > 1) || (!h.col_l0 && h.nref[1] > 1)) w.ue(static_cast<uint32_t>(h.col_idx));
            }
            if ((p.wp && !b) || (p.wbp && b)) {
                w.ue(static_cast<uint32_t>(h.luma_denom));
                w.se(h.chroma_denom_delta);
                for (int l = 0; l < (b ? 2 : 1); l++) {
                    for (int i = 0; i < h.nref[l]; i++) w.flag(h.lw[l][i]);
                    for (int i = 0; i < h.nref[l]; i++) w.flag(h.cw[l][i]);
                    for (int i = 0; i < h.nref[l]; i++) {
                        if (h.lw[l][i]) {
                            w.se(h.dlw[l][i]);
                            w.se(h.lo[l][i]);
                        }
                        if (h.cw[l][i])
                            for (int j = 0; j < 2; j++) {
                                w.se(h.dcw[l][i][j]);
                                w.se(h.dco[l][i][j]);
                            }
                    }
                }
            }
            w.ue(static_cast<uint32_t>(5 - h.max_merge));
        }
        w.se(h.qp_delta);
        if (p.slice_cqp) {
            w.se(h.cb);
            w.se(h.cr);
        }
        if (p.dbctl && p.dbover) w.flag(h.db_override);
        bool dis = p.dbctl && p.dbdis;
        if (h.db_override) {
            w.flag(h.db_disabled);
            dis = h.db_disabled;
            if (!h.db_disabled) {
                w.se(h.beta);
                w.se(h.tc);
            }
        }
        if (p.lf_slices && (h.sao_l || h.sao_c || !dis)) w.flag(h.lf_slices);
    }
    if (p.tiles || p.wpp) {
        w.ue(static_cast<uint32_t>(h.entry_sizes.size()));
        if (!h.entry_sizes.empty()) {
            uint32_t mx = 1;
            for (uint32_t e : h.entry_sizes) mx = std::max(mx, e);
            int len = 1;
            while (len < 32 && ((mx - 1) >> len)) len++;
            w.ue(static_cast<uint32_t>(len - 1));
            for (uint32_t e : h.entry_sizes) w.u(e - 1, len);
        }
    }
    w.trailing();  // byte_alignment()
}

// ---------------------------------------------------------------------------

struct Segment {
    int first_ts = 0, ctbs = 0;
    bool dependent = false;
};

class Writer {
public:
    explicit Writer(Config& c) : cfg_(c) {}

    int run(const char* path) {
        read_config();
        std::vector<uint8_t> out;
        {
            BitWriter w;
            write_vps(w, seq_);
            append(out, nal_unit(NAL_VPS, w.buf));
        }
        {
            BitWriter w;
            write_sps(w, seq_);
            Bits b(w.buf.data(), w.buf.size());
            auto s = std::make_shared<Sps>();
            parse_sps(b, *s);
            spst_[0] = s;
            append(out, nal_unit(NAL_SPS, w.buf));
        }
        {
            BitWriter w;
            write_pps(w, pps_);
            Bits b(w.buf.data(), w.buf.size());
            auto q = std::make_shared<Pps>();
            parse_pps(b, *q, spst_);
            ppst_[0] = q;
            append(out, nal_unit(NAL_PPS, w.buf));
        }
        for (auto& kv : cfg_.kv)
            if (!cfg_.used.count(kv.first)) {
                std::fprintf(stderr, "unknown key %s\n", kv.first.c_str());
                return 2;
            }
        const std::vector<int> order = decode_order();
        std::vector<int> decoded;  // POCs in decoding order
        for (size_t k = 0; k < order.size(); k++) {
            const int poc = order[k];
            bool done = false;
            for (int attempt = 0; attempt < 40 && !done; attempt++) {
                try {
                    std::vector<uint8_t> pic = picture(static_cast<int>(k), poc, decoded, attempt);
                    append(out, pic);
                    done = true;
                } catch (const ParseError& e) {
                    if (verbose_) std::fprintf(stderr, "picture %d attempt %d: %s\n", poc, attempt, e.what());
                }
            }
            if (!done) {
                std::fprintf(stderr, "picture %d: no attempt parsed\n", poc);
                return 1;
            }
            decoded.push_back(poc);
        }
        FILE* f = std::fopen(path, "wb");
        if (!f) return 1;
        std::fwrite(out.data(), 1, out.size(), f);
        std::fclose(f);
        std::printf("%s: %zu bytes, %zu pictures\n", path, out.size(), order.size());
        return 0;
    }

private:
    static void append(std::vector<uint8_t>& a, const std::vector<uint8_t>& b) { a.insert(a.end(), b.begin(), b.end()); }

    void read_config() {
        Config& c = cfg_;
        seed_ = static_cast<uint64_t>(c.geti("seed", 1));
        verbose_ = c.geti("verbose", 0) != 0;
        npics_ = c.geti("pics", 3);
        gop_ = c.gets("gop", "ld");
        types_ = c.gets("types", "pb");
        segs_ = c.geti("segs", 3);
        pdep_ = c.get("pdep", 0.5);
        pseg_i_ = c.get("pseg_i", 0.2);
        seq_.w = c.geti("w", 416);
        seq_.h = c.geti("h", 240);
        seq_.bd = c.geti("bd", 8);
        seq_.ctb = c.geti("ctb", 5);
        seq_.mincb = c.geti("mincb", 3);
        seq_.mintb = c.geti("mintb", 2);
        seq_.maxtb = c.geti("maxtb", std::min(5, seq_.ctb));
        seq_.dinter = c.geti("dinter", 1);
        seq_.dintra = c.geti("dintra", 1);
        seq_.amp = c.geti("amp", 1);
        seq_.sao = c.geti("sao", 1);
        seq_.pcm = c.geti("pcm", 0);
        seq_.pcm_bits_l = c.geti("pcm_bits", seq_.bd - 1);
        seq_.pcm_bits_c = c.geti("pcm_bits_c", seq_.pcm_bits_l);
        seq_.pcm_min = c.geti("pcm_min", 3);
        seq_.pcm_max = c.geti("pcm_max", std::min(5, seq_.ctb));
        seq_.pcm_lf = c.geti("pcm_lf_disabled", 0);
        seq_.tmvp = c.geti("tmvp", 1);
        seq_.sis = c.geti("sis", 1);
        seq_.scaling = c.geti("scaling", 0);
        seq_.crop_r = c.geti("crop_r", 0);
        seq_.crop_b = c.geti("crop_b", 0);
        seq_.refs = c.geti("refs", 2);
        seq_.reorder = gop_ == "ra" ? 3 : 0;
        seq_.max_dec = std::min(16, seq_.refs + 2 + seq_.reorder);
        pps_.dep = c.geti("dep", 1);
        pps_.sdh = c.geti("sdh", 1);
        pps_.cip = c.geti("cip", 1);
        pps_.nref_default[0] = c.geti("nref0", 1);
        pps_.nref_default[1] = c.geti("nref1", 1);
        pps_.init_qp = c.geti("qp", 30);
        pps_.tskip = c.geti("tskip", 1);
        pps_.cuqpd = c.geti("cuqpd", 1);
        pps_.qpd_depth = c.geti("qpd_depth", 1);
        pps_.cb = c.geti("cb", 0);
        pps_.cr = c.geti("cr", 0);
        pps_.slice_cqp = c.geti("slice_cqp", 1);
        pps_.wp = c.geti("wp", 0);
        pps_.wbp = c.geti("wbp", 0);
        pps_.tqb = c.geti("tqb", 0);
        pps_.tcols = c.geti("tcols", 1);
        pps_.trows = c.geti("trows", 1);
        pps_.tiles = pps_.tcols > 1 || pps_.trows > 1;
        pps_.tuniform = c.geti("tuniform", 1);
        pps_.wpp = c.geti("wpp", 0);
        pps_.lf_tiles = c.geti("lf_tiles", 1);
        pps_.lf_slices = c.geti("lf_slices", 1);
        pps_.dbctl = c.geti("dbctl", 1);
        pps_.dbover = c.geti("dbover", 1);
        pps_.dbdis = c.geti("dbdis", 0);
        pps_.beta = c.geti("beta", 0);
        pps_.tc = c.geti("tc", 0);
        pps_.listmod = c.geti("listmod", 0);
        pps_.pml = c.geti("pml", 2);
        const int wc = (seq_.w + (1 << seq_.ctb) - 1) >> seq_.ctb, hc = (seq_.h + (1 << seq_.ctb) - 1) >> seq_.ctb;
        if (pps_.tiles && !pps_.tuniform) {  // random explicit widths / heights
            Rng r(seed_ * 7 + 3);
            split_sizes(r, wc, pps_.tcols, pps_.tcw);
            split_sizes(r, hc, pps_.trows, pps_.trh);
        }
        // bin odds per context
        BinOdds& o = odds_;
        for (double& p : o.p) p = 0.5;
        auto set = [&](int a, int n, double v) {
            for (int i = 0; i < n; i++) o.p[a + i] = v;
        };
        set(C_SAO_MERGE, 1, c.get("p_sao_merge", 0.3));
        set(C_SAO_TYPE, 1, 0.6);
        set(C_SPLIT_CU, 3, c.get("p_split", 0.55));
        set(C_TQ_BYPASS, 1, c.get("p_bypass", 0.15));
        set(C_SKIP, 3, c.get("p_skip", 0.3));
        set(C_PRED_MODE, 1, c.get("p_intra", 0.2));
        set(C_PART_MODE, 1, c.get("p_2nx2n", 0.6));
        set(C_PREV_INTRA, 1, 0.6);
        set(C_RQT_ROOT_CBF, 1, 0.7);
        set(C_MERGE_FLAG, 1, c.get("p_merge", 0.4));
        set(C_INTER_PRED, 4, c.get("p_bi", 0.4));
        set(C_REF_IDX, 2, 0.4);
        set(C_SPLIT_TF, 3, 0.4);
        set(C_CBF_LUMA, 2, c.get("p_cbf", 0.6));
        set(C_CBF_CHROMA, 4, 0.4);
        set(C_MVD_G0, 1, 0.6);
        set(C_QP_DELTA, 1, c.get("p_qpd", 0.25));
        set(C_QP_DELTA + 1, 1, 0.4);
        set(C_TSKIP, 2, 0.3);
        set(C_SIG, C_GT1 - C_SIG, c.get("p_sig", 0.35));
        set(C_GT1, C_GT2 - C_GT1, 0.3);
        set(C_GT2, C_NUM - C_GT2, 0.3);
        o.pcm = c.get("p_pcm", 0.3);
    }

    static void split_sizes(Rng& r, int total, int parts, std::vector<int>& out) {
        out.assign(static_cast<size_t>(parts), 1);
        for (int left = total - parts; left > 0; left--) out[static_cast<size_t>(r.below(parts))]++;
    }

    std::vector<int> decode_order() const {
        std::vector<int> o{0};
        if (gop_ == "ra") {
            for (int b = 0; b < npics_ - 1; b += 4) {
                if (b + 4 <= npics_ - 1) {
                    for (int d : {4, 2, 1, 3}) o.push_back(b + d);
                } else {
                    for (int d = b + 1; d < npics_; d++) o.push_back(d);
                }
            }
        } else {
            for (int p = 1; p < npics_; p++) o.push_back(p);
        }
        return o;
    }

    // CTB scan (6.5.1): tile-scan order to raster, and tile of each CTB
    void scan(std::vector<int>& ts2rs, std::vector<int>& tile_of_ts) const {
        const Sps& s = *spst_[0];
        const Pps& p = *ppst_[0];
        std::vector<int> cw, rh;
        tile_sizes(s, p, cw, rh);
        const int wc = s.width_ctbs();
        ts2rs.clear();
        tile_of_ts.clear();
        int y0 = 0;
        for (size_t r = 0; r < rh.size(); r++) {
            int x0 = 0;
            for (size_t c = 0; c < cw.size(); c++) {
                for (int y = y0; y < y0 + rh[r]; y++)
                    for (int x = x0; x < x0 + cw[c]; x++) {
                        ts2rs.push_back(y * wc + x);
                        tile_of_ts.push_back(static_cast<int>(r * cw.size() + c));
                    }
                x0 += cw[c];
            }
            y0 += rh[r];
        }
    }

    // Segment boundaries: each segment lies in one tile or is made of whole
    // tiles; with wavefronts a segment that starts inside a CTB row ends in
    // that row; a segment that starts a tile is independent.
    std::vector<Segment> plan(Rng& r) const {
        std::vector<int> ts2rs, tile;
        scan(ts2rs, tile);
        const int n = static_cast<int>(ts2rs.size());
        const int wc = spst_[0]->width_ctbs();
        std::vector<Segment> segs;
        auto add = [&](int first, int end, bool tile_start) {
            Segment s;
            s.first_ts = first;
            s.ctbs = end - first;
            s.dependent = !segs.empty() && !tile_start && pps_.dep && r.coin(pdep_);
            segs.push_back(s);
        };
        if (!pps_.tiles) {
            const int target = std::max(1, n / std::max(1, segs_));
            int pos = 0;
            while (pos < n) {
                int end;
                if (segs_ <= 1) {
                    end = n;
                } else {
                    end = std::min(n, pos + std::max(1, r.range(target / 2, target + target / 2)));
                    if (pps_.wpp && (ts2rs[static_cast<size_t>(pos)] % wc) != 0)
                        end = std::min(end, pos + (wc - ts2rs[static_cast<size_t>(pos)] % wc));
                    if (n - end < 2 && r.coin(0.5)) end = n;
                }
                add(pos, end, false);
                pos = end;
            }
            return segs;
        }
        int t = 0;
        const int ntiles = tile.back() + 1;
        std::vector<int> tile_first(static_cast<size_t>(ntiles) + 1, n);
        for (int i = n - 1; i >= 0; i--) tile_first[static_cast<size_t>(tile[static_cast<size_t>(i)])] = i;
        while (t < ntiles) {
            const int a = tile_first[static_cast<size_t>(t)];
            const int tn = tile_first[static_cast<size_t>(t) + 1] - a;
            if (tn > 1 && segs_ > 1 && r.coin(0.5)) {  // split this tile
                const int parts = std::min(tn, r.range(2, 3));
                std::vector<int> cuts;
                while (static_cast<int>(cuts.size()) < parts - 1) {
                    const int c = r.range(1, tn - 1);
                    if (std::find(cuts.begin(), cuts.end(), c) == cuts.end()) cuts.push_back(c);
                }
                std::sort(cuts.begin(), cuts.end());
                int pos = 0;
                for (size_t i = 0; i <= cuts.size(); i++) {
                    const int end = i < cuts.size() ? cuts[i] : tn;
                    add(a + pos, a + end, pos == 0);
                    pos = end;
                }
                t++;
            } else {  // one to three whole tiles
                const int g = std::min(ntiles - t, segs_ > 1 ? r.range(1, 3) : ntiles);
                add(a, tile_first[static_cast<size_t>(t + g)], true);
                t += g;
            }
        }
        return segs;
    }

    // entry points of a segment: substreams after its first (tiles / rows)
    int entry_count(const Segment& s) const {
        std::vector<int> ts2rs, tile;
        scan(ts2rs, tile);
        const int wc = spst_[0]->width_ctbs();
        int n = 0;
        for (int i = s.first_ts + 1; i < s.first_ts + s.ctbs; i++) {
            if (pps_.tiles && tile[static_cast<size_t>(i)] != tile[static_cast<size_t>(i) - 1]) n++;
            else if (pps_.wpp && ts2rs[static_cast<size_t>(i)] % wc == 0) n++;
        }
        return n;
    }

    // header choices of an independent segment
    SegHeader choose_header(Rng& r, int slice_type, const StRps& rps, int total_curr, const SegHeader* first) const {
        SegHeader h;
        h.slice_type = slice_type;
        h.rps = rps;
        h.tmvp = first ? first->tmvp : (seq_.tmvp && r.coin(0.8));
        h.sao_l = seq_.sao && r.coin(0.7);
        h.sao_c = seq_.sao && r.coin(0.7);
        if (slice_type != HEVCR_SLICE_I) {
            const int maxn = std::min(4, std::max(1, total_curr + 1));
            for (int l = 0; l < 2; l++) {
                h.nref[l] = first && first->slice_type != HEVCR_SLICE_I ? first->nref[l] : r.range(1, maxn);
                h.listmod[l] = first ? first->listmod[l] : (pps_.listmod && total_curr > 1 && r.coin(0.5));
                for (int i = 0; i < 16; i++)
                    h.entry[l][i] = first ? first->entry[l][i] : static_cast<uint32_t>(r.below(total_curr));
            }
            if (slice_type == HEVCR_SLICE_P) h.nref[1] = 0;
            h.override_nref = h.nref[0] != pps_.nref_default[0] ||
                              (slice_type == HEVCR_SLICE_B && h.nref[1] != pps_.nref_default[1]) || r.coin(0.3);
            h.mvd_l1_zero = r.coin(0.3);
            h.cabac_init = pps_.cip && r.coin(0.5);
            // the collocated picture of every segment is the first one's
            h.col_l0 = first ? first->col_l0 : (slice_type != HEVCR_SLICE_B || r.coin(0.5));
            if (slice_type == HEVCR_SLICE_P) h.col_l0 = true;
            const int cl = h.col_l0 ? 0 : 1;
            h.col_idx = first ? first->col_idx : r.below(h.nref[cl]);
            if (h.col_idx >= h.nref[cl]) h.col_idx = 0;
            h.luma_denom = r.range(0, 7);
            const int cd = r.range(0, 7);
            h.chroma_denom_delta = cd - h.luma_denom;
            for (int l = 0; l < 2; l++)
                for (int i = 0; i < h.nref[l]; i++) {
                    h.lw[l][i] = r.coin(0.6);
                    h.cw[l][i] = r.coin(0.5);
                    h.dlw[l][i] = r.range(-20, 20);
                    h.lo[l][i] = r.range(-30, 30);
                    for (int j = 0; j < 2; j++) {
                        h.dcw[l][i][j] = r.range(-20, 20);
                        h.dco[l][i][j] = r.range(-60, 60);
                    }
                }
            h.max_merge = r.range(1, 5);
        }
        h.qp_delta = r.range(-4, 4);
        if (pps_.slice_cqp) {
            h.cb = r.range(-4, 4);
            h.cr = r.range(-4, 4);
        }
        h.db_override = pps_.dbctl && pps_.dbover && r.coin(0.5);
        if (h.db_override) {
            h.db_disabled = r.coin(0.3);
            h.beta = r.range(-6, 6);
            h.tc = r.range(-6, 6);
        }
        h.lf_slices = r.coin(0.5);
        return h;
    }

    // One picture: its slice segment NAL units.
    std::vector<uint8_t> picture(int k, int poc, const std::vector<int>& decoded, int attempt) {
        Rng r(seed_ * 1000003ull + static_cast<uint64_t>(k) * 7919ull + static_cast<uint64_t>(attempt) * 104729ull);
        const int nal_type = k == 0 ? NAL_IDR_W_RADL : NAL_TRAIL_R;
        // RPS: the last `refs` decoded pictures, some of them used
        StRps rps;
        std::vector<int> before, after;  // used, closest first
        if (k > 0) {
            std::vector<int> set(decoded.end() - std::min<int>(static_cast<int>(decoded.size()), seq_.refs), decoded.end());
            std::vector<int> neg, pos;
            for (int q : set) (q < poc ? neg : pos).push_back(q);
            std::sort(neg.begin(), neg.end(), std::greater<int>());
            std::sort(pos.begin(), pos.end());
            bool any = false;
            for (int q : neg) {
                rps.delta_s0[rps.num_negative] = q - poc;
                const bool u = r.coin(0.8);
                rps.used_s0[rps.num_negative++] = u;
                if (u) before.push_back(q);
                any |= u;
            }
            for (int q : pos) {
                rps.delta_s1[rps.num_positive] = q - poc;
                const bool u = r.coin(0.8);
                rps.used_s1[rps.num_positive++] = u;
                if (u) after.push_back(q);
                any |= u;
            }
            if (!any) {  // at least one reference
                rps.used_s0[0] = rps.num_negative > 0;
                if (rps.num_negative)
                    before.push_back(neg[0]);
                else {
                    rps.used_s1[0] = true;
                    after.push_back(pos[0]);
                }
            }
        }
        const int total = static_cast<int>(before.size() + after.size());
        int pic_type = HEVCR_SLICE_I;
        if (k > 0) {
            const char t = types_[static_cast<size_t>(r.below(static_cast<int>(types_.size())))];
            pic_type = t == 'b' ? HEVCR_SLICE_B : t == 'p' ? HEVCR_SLICE_P : HEVCR_SLICE_I;
        }
        const Sps& sps = *spst_[0];
        const Pps& pps = *ppst_[0];
        const int pic_ctbs = sps.width_ctbs() * sps.height_ctbs();
        std::vector<Segment> segs = plan(r);
        std::vector<int> ts2rs, tile;
        scan(ts2rs, tile);

        auto cur = std::make_unique<DpbPic>();
        cur->poc = poc;
        cur->id = k;
        cur->width = sps.width;
        cur->height = sps.height;
        PictureRecords rec;
        RefLists refs;
        std::deque<SliceHeader> parsed;
        std::unique_ptr<sd::SliceDecoder<BinChooser>> dec;
        BinChooser eng(odds_, r, pps.tiles);
        std::vector<uint8_t> out;
        SegHeader first_h;
        const SliceHeader* independent = nullptr;
        for (size_t si = 0; si < segs.size(); si++) {
            const Segment& sg = segs[si];
            SegHeader h;
            if (si == 0) {
                h = choose_header(r, pic_type, rps, total, nullptr);
                first_h = h;
            } else if (sg.dependent) {
                h = first_h;
            } else {
                int t = pic_type;
                if (r.coin(pseg_i_)) t = HEVCR_SLICE_I;
                else if (pic_type == HEVCR_SLICE_B && r.coin(0.3)) t = HEVCR_SLICE_P;
                h = choose_header(r, t, rps, total, &first_h);
            }
            h.first = si == 0;
            h.dependent = sg.dependent;
            h.address = static_cast<uint32_t>(ts2rs[static_cast<size_t>(sg.first_ts)]);
            h.poc_lsb = static_cast<uint32_t>(poc & 255);
            // a dependent segment codes no entry points (see write_slice_header)
            h.entry_sizes.assign(sg.dependent ? 0 : static_cast<size_t>(entry_count(sg)), 1u);
            // header with placeholder entry points, parsed back
            BitWriter hw;
            write_slice_header(hw, h, nal_type, seq_, pps_, pic_ctbs);
            SliceHeader sh;
            {
                Bits b(hw.buf.data(), hw.buf.size());
                parse_slice_header(b, nal_type, spst_, ppst_, sh, independent);
            }
            parsed.push_back(sh);
            const SliceHeader& shr = parsed.back();
            if (!shr.dependent) independent = &shr;
            if (si == 0) {
                build_lists(shr, before, after, refs);
                dec = std::make_unique<sd::SliceDecoder<BinChooser>>(sps, pps, parsed.front(), refs, *cur, rec, eng);
            }
            eng.begin_segment(sg.ctbs);
            std::vector<uint8_t> dummy(shr.data_offset + h.entry_sizes.size() + 16, 0);
            dec->segment(shr, dummy.data(), dummy.size(), {});
            // the coded segment, entry points settled (offsets count the
            // emulation prevention bytes of their substreams)
            std::vector<BitWriter>& subs = eng.substreams();
            if (!h.dependent && subs.size() != h.entry_sizes.size() + 1) throw ParseError(-1, "substream count");
            std::vector<uint8_t> nal;
            for (int it = 0; it < 16; it++) {
                BitWriter w;
                write_slice_header(w, h, nal_type, seq_, pps_, pic_ctbs);
                std::vector<uint8_t> rbsp = w.buf;
                std::vector<size_t> starts;
                for (auto& s : subs) {
                    starts.push_back(rbsp.size());
                    rbsp.insert(rbsp.end(), s.buf.begin(), s.buf.end());
                }
                starts.push_back(rbsp.size());
                std::vector<size_t> cpos;
                nal = nal_unit(nal_type, rbsp, &cpos);
                bool same = true;
                for (size_t e = 0; e < h.entry_sizes.size(); e++) {
                    const uint32_t sz = static_cast<uint32_t>(cpos[starts[e + 1]] - cpos[starts[e]]);
                    if (h.entry_sizes[e] != sz) same = false;
                    h.entry_sizes[e] = sz;
                }
                if (same) break;
                if (it == 15) throw ParseError(-1, "entry points did not settle");
            }
            append(out, nal);
        }
        if (!dec->complete()) throw ParseError(-1, "segments do not cover the picture");
        dec->finish();
        dpb_[poc] = std::move(cur);
        return out;
    }

    // Reference lists (8.3.4) from the first segment, as the front end builds them.
    void build_lists(const SliceHeader& sh, const std::vector<int>& before, const std::vector<int>& after,
                     RefLists& refs) {
        if (sh.slice_type == HEVCR_SLICE_I) return;
        const int total = static_cast<int>(before.size() + after.size());
        for (int l = 0; l < (sh.slice_type == HEVCR_SLICE_B ? 2 : 1); l++) {
            std::vector<int> temp;
            const int n = std::max(sh.num_ref_idx[l], total);
            const auto& a = l ? after : before;
            const auto& b = l ? before : after;
            while (static_cast<int>(temp.size()) < n) {
                for (int q : a)
                    if (static_cast<int>(temp.size()) < n) temp.push_back(q);
                for (int q : b)
                    if (static_cast<int>(temp.size()) < n) temp.push_back(q);
            }
            refs.num[l] = sh.num_ref_idx[l];
            for (int i = 0; i < sh.num_ref_idx[l]; i++) {
                const int q = temp[static_cast<size_t>(sh.list_mod[l] ? static_cast<int>(sh.list_entry[l][i]) : i)];
                refs.pic[l][i] = dpb_.at(q).get();
                refs.poc[l][i] = q;
                refs.lt[l][i] = false;
            }
        }
    }

    Config& cfg_;
    uint64_t seed_ = 1;
    bool verbose_ = false;
    int npics_ = 3, segs_ = 3;
    double pdep_ = 0.5, pseg_i_ = 0.2;
    std::string gop_, types_;
    Seq seq_{};
    PicSet pps_{};
    BinOdds odds_{};
    SpsTable spst_{};
    PpsTable ppst_{};
    std::map<int, std::unique_ptr<DpbPic>> dpb_;
};

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s OUT.265 [key=value ...]\n", argv[0]);
        return 2;
    }
    Config c;
    for (int i = 2; i < argc; i++) {
        const char* eq = std::strchr(argv[i], '=');
        if (!eq) {
            std::fprintf(stderr, "bad argument %s\n", argv[i]);
            return 2;
        }
        c.kv[std::string(argv[i], static_cast<size_t>(eq - argv[i]))] = eq + 1;
    }
    try {
        Writer w(c);
        return w.run(argv[1]);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
