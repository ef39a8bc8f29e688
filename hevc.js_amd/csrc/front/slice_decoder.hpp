// Host front end: the slice parser (H.265 7.3.8 slice segment data parsed
// with CABAC, 9.3) and the decoding-process derivations the engine's
// reconstruction records need: intra mode derivation (8.4.2, 8.4.3), luma
// QP prediction (8.6.1), and merge / AMVP / temporal motion vector
// prediction (8.5.3.2). The reference's counterparts are
// src/decoding/coding_tree.cpp, residual_coding.cpp, syntax_elements.cpp and
// inter_prediction.cpp.
//
// The parser is a template over its arithmetic decoder: slice_data.cpp
// instantiates it with Cabac; an engine that chooses the bins it is asked
// for (and arithmetic-codes them) turns the same syntax into a writer.
#pragma once

#include <algorithm>
#include <cstdlib>
#include <cstring>

#include "slice_data.hpp"

namespace hevcf {
namespace sd {
constexpr int kUnsupported = -6;

// ScanOrder[log2 size][scanIdx][pos] -> (x, y) packed as x | y << 4 (6.5.3-6.5.5).
struct ScanTables {
    uint8_t pos[4][3][64];
    ScanTables() {
        for (int l = 0; l < 4; l++) {
            const int n = 1 << l;
            int i = 0, x = 0, y = 0;
            while (i < n * n) {  // up-right diagonal
                while (y >= 0) {
                    if (x < n && y < n) pos[l][0][i++] = static_cast<uint8_t>(x | y << 4);
                    y--;
                    x++;
                }
                y = x;
                x = 0;
            }
            i = 0;
            for (int yy = 0; yy < n; yy++)
                for (int xx = 0; xx < n; xx++) pos[l][1][i++] = static_cast<uint8_t>(xx | yy << 4);  // horizontal
            i = 0;
            for (int xx = 0; xx < n; xx++)
                for (int yy = 0; yy < n; yy++) pos[l][2][i++] = static_cast<uint8_t>(xx | yy << 4);  // vertical
        }
    }
};
inline const ScanTables kScan;

// Merge pruning compares the whole stored motion of a unit, both lists'
// vectors and indices whether used or not, as the reference does
// (inter_prediction.cpp:297-302); a list-1 vector survives in the store where
// an 8x4 / 4x8 unit lost its list-1 prediction (:468-472).
inline bool same_motion(const Mv4& a, const Mv4& b) {
    return (a.flags & 3) == (b.flags & 3) && a.ref[0] == b.ref[0] && a.ref[1] == b.ref[1] &&
           a.mv[0][0] == b.mv[0][0] && a.mv[0][1] == b.mv[0][1] && a.mv[1][0] == b.mv[1][0] &&
           a.mv[1][1] == b.mv[1][1];
}

inline int clip3(int lo, int hi, int v) { return v < lo ? lo : (v > hi ? hi : v); }

// Motion vector scaling by POC distances (8-179..8-183).
inline int16_t scale_mv(int mv, int td, int tb) {
    td = clip3(-128, 127, td);
    tb = clip3(-128, 127, tb);
    const int tx = (16384 + (std::abs(td) >> 1)) / td;
    const int dsf = clip3(-4096, 4095, (tb * tx + 32) >> 6);
    const int p = dsf * mv;
    return static_cast<int16_t>(clip3(-32768, 32767, (p >= 0 ? 1 : -1) * ((std::abs(p) + 127) >> 8)));
}

// The slice parser of one picture, fed the picture's slice segments in
// decoding order. Engine is the arithmetic decoder (Cabac); the template
// lets a bin-choosing engine drive the same syntax.
template <class Engine>
class SliceDecoder {
public:
    // `first`: the picture's first slice segment, whose lists (refs) and
    // collocated picture every segment uses, as the reference does
    // (decoder.cpp:108-115).
    SliceDecoder(const Sps& sps, const Pps& pps, const SliceHeader& first, const RefLists& refs, DpbPic& cur,
                 PictureRecords& out, Engine& cab)
        : sps_(sps), pps_(pps), sh_(&first), refs_(refs), cur_(cur), out_(out), cab_(cab) {
        W_ = sps.width;
        H_ = sps.height;
        w4_ = (W_ + 3) >> 2;
        h4_ = (H_ + 3) >> 2;
        const size_t n4 = static_cast<size_t>(w4_) * h4_;
        mode4_.assign(n4, 0);
        ipm4_.assign(n4, 1);
        mvok4_.assign(n4, 0);
        mot4_.assign(n4, Mv4{});
        depth4_.assign(n4, 0);
        qp4_.assign(n4, 0);
        qbd_ = 6 * (sps.bit_depth - 8);
        log2_qg_ = sps.log2_ctb - (pps.cu_qp_delta ? pps.diff_cu_qp_delta_depth : 0);
        par_mrg_ = pps.log2_parallel_merge_level;
        // NoBackwardPredFlag (8.5.3.2.9): no reference picture after the current one
        no_backward_ = true;
        for (int l = 0; l < 2; l++)
            for (int i = 0; i < refs.num[l]; i++)
                if (refs.poc[l][i] > cur.poc) no_backward_ = false;
        col_ = nullptr;
        if (first.temporal_mvp && first.slice_type != HEVCR_SLICE_I) {
            const int l = (first.slice_type == HEVCR_SLICE_B && !first.collocated_from_l0) ? 1 : 0;
            if (first.collocated_ref_idx >= refs.num[l]) throw ParseError(-1, "collocated_ref_idx");
            col_ = refs.pic[l][first.collocated_ref_idx];
        }
        // CTB scan and tiles (6.5.1)
        const int wc = sps.width_ctbs(), hc = sps.height_ctbs(), n = wc * hc;
        std::vector<int> cw, rh;
        tile_sizes(sps, pps, cw, rh);
        if (pps.entropy_sync && (cw.size() > 1 || rh.size() > 1))
            throw ParseError(kUnsupported, "tiles together with wavefront parallel processing");
        tile_.assign(static_cast<size_t>(n), 0);
        rs2ts_.assign(static_cast<size_t>(n), 0);
        ts2rs_.assign(static_cast<size_t>(n), 0);
        int ts = 0, y0 = 0;
        for (size_t r = 0; r < rh.size(); r++) {
            int x0 = 0;
            for (size_t c = 0; c < cw.size(); c++) {
                for (int y = y0; y < y0 + rh[r]; y++)
                    for (int x = x0; x < x0 + cw[c]; x++) {
                        const int a = y * wc + x;
                        tile_[static_cast<size_t>(a)] = static_cast<uint16_t>(r * cw.size() + c);
                        rs2ts_[static_cast<size_t>(a)] = ts;
                        ts2rs_[static_cast<size_t>(ts++)] = a;
                    }
                x0 += cw[c];
            }
            y0 += rh[r];
        }
        ctb_seg_.assign(static_cast<size_t>(n), 0);
        out_.sao.assign(static_cast<size_t>(n), hevcr_sao{});
        if (pps.tiles && (cw.size() > 1 || rh.size() > 1)) {
            out_.has_tiles = true;
            out_.tiles = hevcr_tiles{};
            out_.tiles.cols = static_cast<int32_t>(cw.size());
            out_.tiles.rows = static_cast<int32_t>(rh.size());
            for (size_t c = 0; c < cw.size(); c++) out_.tiles.col_width[c] = static_cast<uint16_t>(cw[c]);
            for (size_t r = 0; r < rh.size(); r++) out_.tiles.row_height[r] = static_cast<uint16_t>(rh[r]);
            out_.tiles.loop_filter_across_tiles = pps.loop_filter_across_tiles;
        }
        out_.pcm_loop_filter_disabled = sps.pcm && sps.pcm_loop_filter_disabled;
    }

    // All CTBs of the picture parsed.
    bool complete() const { return next_ts_ == static_cast<int>(ts2rs_.size()); }

    // One slice segment (7.3.8.1): its CTBs from slice_segment_address on, in
    // the picture's CTB scan, until end_of_slice_segment_flag.
    void segment(const SliceHeader& sh, const uint8_t* data, size_t size, const std::vector<uint32_t>& epb) {
        const int wc = sps_.width_ctbs(), hc = sps_.height_ctbs();
        if (sh.segment_address >= static_cast<uint32_t>(wc * hc)) throw ParseError(-1, "slice_segment_address");
        const int first_ts = rs2ts_[sh.segment_address];
        if (first_ts != next_ts_) throw ParseError(-1, "slice segment does not continue the picture");
        if (nseg_ >= HEVCR_MAX_SLICES) throw ParseError(kUnsupported, "more than 256 slice segments in a picture");
        // every segment uses the lists of the picture's first segment
        // (decoder.cpp:108-115); a segment that would index past them has no
        // defined result there
        if (sh.slice_type != HEVCR_SLICE_I)
            for (int l = 0; l < (sh.slice_type == HEVCR_SLICE_B ? 2 : 1); l++)
                if (sh.num_ref_idx[l] > refs_.num[l])
                    throw ParseError(kUnsupported, "slice segment with more references than the first segment's lists");
        sh_ = &sh;
        seg_ = nseg_++;
        slice_qp_ = sh.slice_qp(pps_);
        if (slice_qp_ < -qbd_ || slice_qp_ > 51) throw ParseError(-1, "SliceQpY");
        // substream starts (RBSP offsets) from the entry points of wavefront
        // rows; tiles are read one after the other from the aligned end of the
        // one before, their entry points unused, as the reference reads them
        // (coding_tree.cpp:626-639)
        std::vector<size_t> starts{sh.data_offset};
        if (!sh.entry_points.empty() && pps_.entropy_sync) {
            size_t coded = rbsp_to_coded(sh.data_offset, epb);
            for (uint32_t e : sh.entry_points) {
                coded += e;
                starts.push_back(coded_to_rbsp(coded, epb));
            }
        }
        const int init_type = sh.slice_type == HEVCR_SLICE_I ? 0
                              : sh.slice_type == HEVCR_SLICE_P ? (sh.cabac_init ? 2 : 1)
                                                               : (sh.cabac_init ? 1 : 2);
        init_type_ = init_type;
        size_t sub = 0;
        start_substream(data, size, starts, sub);
        // a dependent segment goes on with the contexts the previous segment
        // left (the reference keeps one engine for the picture,
        // coding_tree.cpp:579-583); every segment restarts qPY_PREV (588-589)
        if (!sh.dependent) cab_.init_contexts(init_type, slice_qp_);
        uint8_t wpp_saved[C_NUM];
        bool wpp_have = false;
        const int first = static_cast<int>(sh.segment_address);
        last_qpy_ = slice_qp_;
        for (int ts = first_ts; ts < wc * hc; ts++) {
            const int ctb = ts2rs_[static_cast<size_t>(ts)];
            const int rx = ctb % wc, ry = ctb / wc;
            if (pps_.entropy_sync && rx == 0 && ctb != first) {
                // new substream: the arithmetic decoder restarts at its entry point (9.3.2.4/9.3.2.5)
                if (++sub >= starts.size()) throw ParseError(-1, "missing entry point");
                start_substream(data, size, starts, sub);
                // contexts synchronised from the CTB above-right when it exists, else initialised
                if (wc > 1 && wpp_have)
                    cab_.load(wpp_saved);
                else
                    cab_.init_contexts(init_type, slice_qp_);
                last_qpy_ = slice_qp_;
            }
            ctb_x0_ = rx << sps_.log2_ctb;
            ctb_y0_ = ry << sps_.log2_ctb;
            ctb_addr_ = ctb;
            ctb_seg_[static_cast<size_t>(ctb)] = static_cast<uint8_t>(seg_);
            if (sh.sao_luma || sh.sao_chroma) sao(rx, ry);
            coding_quadtree(ctb_x0_, ctb_y0_, sps_.log2_ctb, 0);
            const int end_of_slice = cab_.terminate(Cabac::kEndOfSlice);
            if (pps_.entropy_sync && (rx == 1 || wc == 1)) {
                cab_.save(wpp_saved);
                wpp_have = true;
            }
            next_ts_ = ts + 1;
            if (end_of_slice) break;
            if (ts == wc * hc - 1) throw ParseError(-1, "slice data runs past the last CTB");
            const int nxt = ts2rs_[static_cast<size_t>(ts) + 1];
            if (tile_[static_cast<size_t>(nxt)] != tile_[static_cast<size_t>(ctb)]) {
                // end of a tile: end_of_subset_one_bit, byte alignment, the next
                // substream; the reference keeps the contexts and restarts
                // qPY_PREV (coding_tree.cpp:628-639)
                if (!cab_.terminate(Cabac::kEndOfSubset)) throw ParseError(-1, "end_of_subset_one_bit");
                cab_.restart_aligned();
                last_qpy_ = slice_qp_;
            } else if (pps_.entropy_sync && rx == wc - 1) {
                if (!cab_.terminate(Cabac::kEndOfSubset)) throw ParseError(-1, "end_of_subset_one_bit");
                wpp_have = wpp_have && wc > 1;
            }
        }
        // the segment's parameters as the filters and the engine read them
        hevcr_slice q{};
        q.slice_type = sh.slice_type;
        q.slice_cb_qp_offset = sh.cb_qp_offset;
        q.slice_cr_qp_offset = sh.cr_qp_offset;
        q.deblocking_disabled = sh.deblocking_disabled;
        q.beta_offset_div2 = sh.beta_offset_div2;
        q.tc_offset_div2 = sh.tc_offset_div2;
        q.loop_filter_across_slices = sh.loop_filter_across_slices;
        q.weighted_pred = (sh.slice_type == HEVCR_SLICE_P && pps_.weighted_pred) ||
                          (sh.slice_type == HEVCR_SLICE_B && pps_.weighted_bipred);
        q.luma_log2_wd = sh.luma_log2_wd;
        q.chroma_log2_wd = sh.chroma_log2_wd;
        for (int l = 0; l < 2; l++)
            for (int i = 0; i < HEVCR_MAX_REFS; i++)
                for (int c = 0; c < 3; c++) {
                    q.weights[l][i].w[c] = sh.weights[l][i].w[c];
                    q.weights[l][i].o[c] = sh.weights[l][i].o[c];
                }
        out_.slices.push_back(q);
    }

    // The picture's records after its last segment: CTB segment indices and
    // the motion kept for later pictures.
    void finish() {
        if (!complete()) throw ParseError(-1, "picture with CTBs in no slice segment");
        out_.ctb_slice.assign(ctb_seg_.begin(), ctb_seg_.end());
        store_col();
    }

private:
    static size_t coded_to_rbsp(size_t c, const std::vector<uint32_t>& epb) {
        size_t n = 0;
        for (uint32_t e : epb)
            if (e < c) n++;
        return c - n;
    }
    static size_t rbsp_to_coded(size_t r, const std::vector<uint32_t>& epb) {
        size_t c = r;
        for (uint32_t e : epb) {
            if (e <= c)
                c++;
            else
                break;
        }
        return c;
    }
    void start_substream(const uint8_t* data, size_t size, const std::vector<size_t>& starts, size_t k) {
        const size_t b = starts[k], e = k + 1 < starts.size() ? starts[k + 1] : size;
        if (b > size || e > size || b > e) throw ParseError(-1, "entry point outside the slice data");
        cab_.start(data + b, data + e);
    }

    // -- neighbourhood -------------------------------------------------------
    size_t g4(int x, int y) const { return static_cast<size_t>(y >> 2) * w4_ + (x >> 2); }
    bool inside(int x, int y) const { return x >= 0 && y >= 0 && x < W_ && y < H_; }
    // the CTB holding luma position (x, y) is in the current CTB's slice
    // segment and tile (6.4.1 as the reference applies it: segment indices,
    // coding_tree.cpp:180-196)
    bool same_region(int x, int y) const {
        const int a = (y >> sps_.log2_ctb) * sps_.width_ctbs() + (x >> sps_.log2_ctb);
        return ctb_seg_[static_cast<size_t>(a)] == seg_ && tile_[static_cast<size_t>(a)] == tile_[static_cast<size_t>(ctb_addr_)];
    }
    // decoded (z-scan earlier and in this slice segment and tile): the CU there has been started
    bool avail(int x, int y) const { return inside(x, y) && mode4_[g4(x, y)] != 0 && same_region(x, y); }
    template <typename T, typename V>
    void fill(std::vector<T>& g, int x0, int y0, int w, int h, V v) {
        for (int y = y0; y < y0 + h; y += 4)
            for (int x = x0; x < x0 + w; x += 4) g[g4(x, y)] = static_cast<T>(v);
    }

    // -- SAO (7.3.8.3) -------------------------------------------------------
    void sao(int rx, int ry) {
        const int wc = sps_.width_ctbs();
        hevcr_sao& s = out_.sao[static_cast<size_t>(ry) * wc + rx];
        s = hevcr_sao{};
        // merge candidates in this slice segment (by raster address against the
        // segment's address, as the reference compares, coding_tree.cpp:67-81) and tile
        const int a = ry * wc + rx, seg0 = static_cast<int>(sh_->segment_address);
        const bool left_ok = rx > 0 && a > seg0 && tile_[static_cast<size_t>(a)] == tile_[static_cast<size_t>(a) - 1];
        const bool up_ok = ry > 0 && a - wc >= seg0 && tile_[static_cast<size_t>(a)] == tile_[static_cast<size_t>(a - wc)];
        if (left_ok && cab_.decode(C_SAO_MERGE)) {
            s = out_.sao[static_cast<size_t>(ry) * wc + rx - 1];
            return;
        }
        if (up_ok && cab_.decode(C_SAO_MERGE)) {
            s = out_.sao[static_cast<size_t>(ry - 1) * wc + rx];
            return;
        }
        const int cmax = (1 << (std::min(sps_.bit_depth, 10) - 5)) - 1;
        for (int c = 0; c < 3; c++) {
            if (!((sh_->sao_luma && c == 0) || (sh_->sao_chroma && c > 0))) continue;
            int type;
            if (c == 2) {
                type = s.type[1];
            } else {
                type = cab_.decode(C_SAO_TYPE) ? (cab_.bypass() ? 2 : 1) : 0;
            }
            s.type[c] = static_cast<uint8_t>(type);
            if (!type) continue;
            int abs[4];
            for (int i = 0; i < 4; i++) {
                int v = 0;
                while (v < cmax && cab_.bypass()) v++;
                abs[i] = v;
            }
            if (type == 1) {  // band offset: signs, then the band position
                for (int i = 0; i < 4; i++)
                    if (abs[i] && cab_.bypass()) abs[i] = -abs[i];
                s.band_pos[c] = static_cast<uint8_t>(cab_.bypass_bits(5));
                for (int i = 0; i < 4; i++) s.offset[c][i] = static_cast<int8_t>(abs[i]);
                s.offset[c][4] = 0;
            } else {  // edge offset: categories 1, 2 positive, 3, 4 negative
                if (c == 0) s.eo_class[0] = static_cast<uint8_t>(cab_.bypass_bits(2));
                if (c == 1) s.eo_class[1] = static_cast<uint8_t>(cab_.bypass_bits(2));
                if (c == 2) s.eo_class[2] = s.eo_class[1];
                s.offset[c][0] = static_cast<int8_t>(abs[0]);
                s.offset[c][1] = static_cast<int8_t>(abs[1]);
                s.offset[c][2] = 0;
                s.offset[c][3] = static_cast<int8_t>(-abs[2]);
                s.offset[c][4] = static_cast<int8_t>(-abs[3]);
            }
        }
    }

    // -- coding quadtree / unit (7.3.8.4, 7.3.8.5) ----------------------------
    void coding_quadtree(int x0, int y0, int log2, int depth) {
        const int s = 1 << log2;
        bool split;
        if (x0 + s <= W_ && y0 + s <= H_ && log2 > sps_.log2_min_cb) {
            int inc = 0;
            if (avail(x0 - 1, y0) && depth4_[g4(x0 - 1, y0)] > depth) inc++;
            if (avail(x0, y0 - 1) && depth4_[g4(x0, y0 - 1)] > depth) inc++;
            split = cab_.decode(C_SPLIT_CU + inc) != 0;
        } else {
            split = log2 > sps_.log2_min_cb;
        }
        if (log2 >= log2_qg_) start_qg(x0, y0);
        if (split) {
            const int h = s >> 1;
            for (int k = 0; k < 4; k++) {
                const int x = x0 + (k & 1) * h, y = y0 + (k >> 1) * h;
                if (x < W_ && y < H_) coding_quadtree(x, y, log2 - 1, depth + 1);
            }
        } else {
            coding_unit(x0, y0, log2, depth);
        }
    }

    // Start of a quantization group (8.6.1): qPY_PRED from the left and above
    // groups inside the same CTB, else from the previous CU in decoding order.
    void start_qg(int xq, int yq) {
        qp_delta_coded_ = false;
        qp_delta_ = 0;
        const int prev = last_qpy_;
        const int a = (xq & (sps_.ctb_size() - 1)) ? qp4_[g4(xq - 1, yq)] : prev;
        const int b = (yq & (sps_.ctb_size() - 1)) ? qp4_[g4(xq, yq - 1)] : prev;
        qpy_pred_ = (a + b + 1) >> 1;
    }

    int cu_qpy() const { return ((qpy_pred_ + qp_delta_ + 52 + 2 * qbd_) % (52 + qbd_)) - qbd_; }

    void coding_unit(int x0, int y0, int log2, int depth) {
        const int s = 1 << log2;
        hevcr_cu cu{};
        cu.x = static_cast<uint16_t>(x0);
        cu.y = static_cast<uint16_t>(y0);
        cu.log2_size = static_cast<uint8_t>(log2);
        cur_bypass_ = pps_.transquant_bypass && cab_.decode(C_TQ_BYPASS);
        if (cur_bypass_) cu.flags = HEVCR_CU_BYPASS;
        bool skip = false;
        if (sh_->slice_type != HEVCR_SLICE_I) {
            int inc = 0;
            if (avail(x0 - 1, y0) && mode4_[g4(x0 - 1, y0)] == 1 + HEVCR_MODE_SKIP) inc++;
            if (avail(x0, y0 - 1) && mode4_[g4(x0, y0 - 1)] == 1 + HEVCR_MODE_SKIP) inc++;
            skip = cab_.decode(C_SKIP + inc) != 0;
        }
        const uint32_t cu_idx = static_cast<uint32_t>(out_.cus.size());
        fill(depth4_, x0, y0, s, s, depth);
        int npu = 0;
        bool merge0 = false;
        if (skip) {
            cu.pred_mode = HEVCR_MODE_SKIP;
            cu.part_mode = HEVCR_PART_2Nx2N;
            fill(mode4_, x0, y0, s, s, 1 + HEVCR_MODE_SKIP);
            prediction_unit(x0, y0, s, x0, y0, s, s, 0, HEVCR_PART_2Nx2N, true, depth);
            npu = 1;
        } else {
            int pred = HEVCR_MODE_INTRA;
            if (sh_->slice_type != HEVCR_SLICE_I) pred = cab_.decode(C_PRED_MODE) ? HEVCR_MODE_INTRA : HEVCR_MODE_INTER;
            int part = HEVCR_PART_2Nx2N;
            if (pred != HEVCR_MODE_INTRA || log2 == sps_.log2_min_cb) part = part_mode(pred, log2);
            cu.pred_mode = static_cast<uint8_t>(pred);
            cu.part_mode = static_cast<uint8_t>(part);
            fill(mode4_, x0, y0, s, s, 1 + pred);
            if (pred == HEVCR_MODE_INTRA) {
                if (part == HEVCR_PART_2Nx2N && sps_.pcm && log2 >= sps_.pcm_log2_min && log2 <= sps_.pcm_log2_max &&
                    cab_.terminate(Cabac::kPcmFlag)) {
                    pcm_unit(cu);
                    return;
                }
                intra_modes(cu, part);
            } else {
                const int q = s / 4, h = s / 2;
                int r[4][4], n = 0;
                auto add = [&](int x, int y, int w, int hh) {
                    r[n][0] = x;
                    r[n][1] = y;
                    r[n][2] = w;
                    r[n][3] = hh;
                    n++;
                };
                switch (part) {
                    case HEVCR_PART_2Nx2N: add(x0, y0, s, s); break;
                    case HEVCR_PART_2NxN: add(x0, y0, s, h); add(x0, y0 + h, s, h); break;
                    case HEVCR_PART_Nx2N: add(x0, y0, h, s); add(x0 + h, y0, h, s); break;
                    case HEVCR_PART_2NxnU: add(x0, y0, s, q); add(x0, y0 + q, s, s - q); break;
                    case HEVCR_PART_2NxnD: add(x0, y0, s, s - q); add(x0, y0 + s - q, s, q); break;
                    case HEVCR_PART_nLx2N: add(x0, y0, q, s); add(x0 + q, y0, s - q, s); break;
                    case HEVCR_PART_nRx2N: add(x0, y0, s - q, s); add(x0 + s - q, y0, q, s); break;
                    default:
                        for (int k = 0; k < 4; k++) add(x0 + (k & 1) * h, y0 + (k >> 1) * h, h, h);
                }
                for (int k = 0; k < n; k++) {
                    const bool m = prediction_unit(x0, y0, s, r[k][0], r[k][1], r[k][2], r[k][3], k, part, false, depth);
                    if (k == 0) merge0 = m;
                }
                npu = n;
            }
        }
        cu.num_pu = static_cast<uint8_t>(cu.pred_mode == HEVCR_MODE_INTRA ? 0 : npu);
        out_.cus.push_back(cu);
        cur_cu_ = cu_idx;
        cur_intra_ = cu.pred_mode == HEVCR_MODE_INTRA;
        if (!skip) {
            bool root_cbf = true;
            if (cu.pred_mode != HEVCR_MODE_INTRA && !(cu.part_mode == HEVCR_PART_2Nx2N && merge0))
                root_cbf = cab_.decode(C_RQT_ROOT_CBF) != 0;
            if (root_cbf) {
                intra_split_ = cu.pred_mode == HEVCR_MODE_INTRA && cu.part_mode == HEVCR_PART_NxN;
                max_tr_depth_ = cu.pred_mode == HEVCR_MODE_INTRA ? sps_.max_th_depth_intra + (intra_split_ ? 1 : 0)
                                                                 : sps_.max_th_depth_inter;
                cu_part_ = cu.part_mode;
                transform_tree(x0, y0, x0, y0, log2, 0, 0, false, false);
            }
        }
        const int qpy = cu_qpy();
        out_.cus[cu_idx].qp_y = static_cast<int8_t>(qpy);
        fill(qp4_, x0, y0, s, s, qpy);
        last_qpy_ = qpy;
    }

    // A PCM CU (7.3.8.7; the reference: coding_tree.cpp:852-875, 1425-1468):
    // the samples as read (no shift to the bit depth) in a record of the CU's
    // size, the CU intra and PCM with no bypass flag, its QpY the previous
    // CU's (qPY_PREV is not updated), the arithmetic decoder restarted after
    // the samples and, as the reference does, the contexts initialised again.
    void pcm_unit(hevcr_cu& cu) {
        const int s = 1 << cu.log2_size;
        cu.flags = HEVCR_CU_PCM;
        cu.pred_mode = HEVCR_MODE_INTRA;
        cu.part_mode = HEVCR_PART_2Nx2N;
        cu.intra_luma[0] = 1;
        cu.intra_chroma = 1;
        cu.num_pu = 0;
        cu.qp_y = static_cast<int8_t>(last_qpy_);
        hevcr_tu t{};
        t.x = t.xbase = cu.x;
        t.y = t.ybase = cu.y;
        t.log2_size = cu.log2_size;
        t.cbf = HEVCR_CBF_Y | HEVCR_CBF_CB | HEVCR_CBF_CR;
        t.cu_idx = static_cast<uint32_t>(out_.cus.size());
        cab_.raw_begin();
        for (int c = 0; c < 3; c++) {
            const int n = c ? s / 2 : s, bits = c ? sps_.pcm_bits_chroma : sps_.pcm_bits_luma;
            t.coeff_off[c] = static_cast<uint32_t>(out_.coeffs.size());
            for (int i = 0; i < n * n; i++) out_.coeffs.push_back(static_cast<int16_t>(cab_.raw_bits(bits)));
        }
        cab_.raw_end();
        cab_.init_contexts(init_type_, slice_qp_);
        out_.cus.push_back(cu);
        out_.tus.push_back(t);
        fill(qp4_, cu.x, cu.y, s, s, last_qpy_);
    }

    int part_mode(int pred, int log2) {
        if (pred == HEVCR_MODE_INTRA) return cab_.decode(C_PART_MODE) ? HEVCR_PART_2Nx2N : HEVCR_PART_NxN;
        if (cab_.decode(C_PART_MODE)) return HEVCR_PART_2Nx2N;
        if (log2 == sps_.log2_min_cb) {
            if (cab_.decode(C_PART_MODE + 1)) return HEVCR_PART_2NxN;
            if (log2 == 3) return HEVCR_PART_Nx2N;
            return cab_.decode(C_PART_MODE + 2) ? HEVCR_PART_Nx2N : HEVCR_PART_NxN;
        }
        const int hor = cab_.decode(C_PART_MODE + 1);
        if (!sps_.amp) return hor ? HEVCR_PART_2NxN : HEVCR_PART_Nx2N;
        if (cab_.decode(C_PART_MODE + 3)) return hor ? HEVCR_PART_2NxN : HEVCR_PART_Nx2N;
        const int b = cab_.bypass();
        return hor ? (b ? HEVCR_PART_2NxnD : HEVCR_PART_2NxnU) : (b ? HEVCR_PART_nRx2N : HEVCR_PART_nLx2N);
    }

    // -- intra modes (7.3.8.5, 8.4.2, 8.4.3) ---------------------------------
    void intra_modes(hevcr_cu& cu, int part) {
        const int s = 1 << cu.log2_size;
        const int n = part == HEVCR_PART_NxN ? 4 : 1, ps = n == 4 ? s / 2 : s;
        bool prev[4];
        for (int k = 0; k < n; k++) prev[k] = cab_.decode(C_PREV_INTRA) != 0;
        for (int k = 0; k < n; k++) {
            const int xp = cu.x + (k & 1) * ps, yp = cu.y + (k >> 1) * ps;
            // candidates A (left) and B (above, inside this CTB row)
            auto cand = [&](int x, int y, bool above) {
                if (!avail(x, y) || mode4_[g4(x, y)] != 1 + HEVCR_MODE_INTRA) return 1;
                if (above && y < ((yp >> sps_.log2_ctb) << sps_.log2_ctb)) return 1;
                return static_cast<int>(ipm4_[g4(x, y)]);
            };
            const int a = cand(xp - 1, yp, false), b = cand(xp, yp - 1, true);
            int c[3];
            if (a == b) {
                if (a < 2) {
                    c[0] = 0;
                    c[1] = 1;
                    c[2] = 26;
                } else {
                    c[0] = a;
                    c[1] = 2 + ((a + 29) % 32);
                    c[2] = 2 + ((a - 2 + 1) % 32);
                }
            } else {
                c[0] = a;
                c[1] = b;
                c[2] = (a != 0 && b != 0) ? 0 : ((a != 1 && b != 1) ? 1 : 26);
            }
            int mode;
            if (prev[k]) {
                int idx = 0;
                if (cab_.bypass()) idx = cab_.bypass() ? 2 : 1;
                mode = c[idx];
            } else {
                mode = static_cast<int>(cab_.bypass_bits(5));
                std::sort(c, c + 3);
                for (int i = 0; i < 3; i++)
                    if (mode >= c[i]) mode++;
            }
            cu.intra_luma[k] = static_cast<uint8_t>(mode);
            fill(ipm4_, xp, yp, ps, ps, mode);
        }
        // intra_chroma_pred_mode and the 4:2:0 derivation (Table 8-2)
        int cm = 4;
        if (cab_.decode(C_CHROMA_MODE)) cm = static_cast<int>(cab_.bypass_bits(2));
        const int luma = cu.intra_luma[0];
        static const int kList[4] = {0, 26, 10, 1};
        cu.intra_chroma = static_cast<uint8_t>(cm == 4 ? luma : (kList[cm] == luma ? 34 : kList[cm]));
    }

    // -- prediction unit and motion (7.3.8.6, 8.5.3.2) -----------------------
    bool prediction_unit(int xc, int yc, int cs, int xp, int yp, int w, int h, int part_idx, int part, bool skip,
                         int depth) {
        Mv4 m;
        bool merge = skip;
        if (!skip) merge = cab_.decode(C_MERGE_FLAG) != 0;
        if (merge) {
            int idx = 0;
            if (sh_->max_num_merge_cand > 1 && cab_.decode(C_MERGE_IDX)) {
                idx = 1;
                while (idx < sh_->max_num_merge_cand - 1 && cab_.bypass()) idx++;
            }
            m = merge_candidate(xc, yc, cs, xp, yp, w, h, part_idx, part, idx);
        } else {
            int dir = 0;  // 0: L0, 1: L1, 2: bi
            if (sh_->slice_type == HEVCR_SLICE_B) {
                if (w + h != 12 && cab_.decode(C_INTER_PRED + depth))
                    dir = 2;
                else
                    dir = cab_.decode(C_INTER_PRED + 4);
            }
            int mvd[2][2] = {{0, 0}, {0, 0}}, refi[2] = {-1, -1}, mvp[2] = {0, 0};
            for (int l = 0; l < 2; l++) {
                if (!(dir == 2 || dir == l)) continue;
                refi[l] = 0;
                if (sh_->num_ref_idx[l] > 1) {
                    const int cmax = sh_->num_ref_idx[l] - 1;
                    int v = 0;
                    while (v < cmax) {
                        const int bin = v < 2 ? cab_.decode(C_REF_IDX + v) : cab_.bypass();
                        if (!bin) break;
                        v++;
                    }
                    refi[l] = v;
                }
                if (l == 1 && sh_->mvd_l1_zero && dir == 2) {
                    mvd[1][0] = mvd[1][1] = 0;
                } else {
                    mvd_coding(mvd[l]);
                }
                mvp[l] = cab_.decode(C_MVP);
            }
            m.flags = 0;
            for (int l = 0; l < 2; l++) {
                if (refi[l] < 0) continue;
                if (refi[l] >= refs_.num[l]) throw ParseError(-1, "ref_idx beyond the list");
                int16_t pred[2];
                amvp(xc, yc, cs, xp, yp, w, h, part_idx, l, refi[l], mvp[l], pred);
                for (int c = 0; c < 2; c++) {
                    const int u = (pred[c] + mvd[l][c] + 65536) & 0xffff;
                    m.mv[l][c] = static_cast<int16_t>(u >= 32768 ? u - 65536 : u);
                }
                m.ref[l] = static_cast<int8_t>(refi[l]);
                m.flags |= static_cast<uint8_t>(1 << l);
            }
        }
        // store for neighbours and the record
        for (int y = yp; y < yp + h; y += 4)
            for (int x = xp; x < xp + w; x += 4) {
                mot4_[g4(x, y)] = m;
                mvok4_[g4(x, y)] = 1;
            }
        hevcr_pu pu{};
        pu.x = static_cast<uint16_t>(xp);
        pu.y = static_cast<uint16_t>(yp);
        pu.w = static_cast<uint16_t>(w);
        pu.h = static_cast<uint16_t>(h);
        for (int l = 0; l < 2; l++) {
            const bool used = (m.flags >> l) & 1;
            pu.mv[l][0] = used ? m.mv[l][0] : 0;
            pu.mv[l][1] = used ? m.mv[l][1] : 0;
            pu.ref_idx[l] = used ? m.ref[l] : -1;
        }
        pu.pred_flags = m.flags & 3;
        out_.pus.push_back(pu);
        return merge;
    }

    void mvd_coding(int* mvd) {
        int g0[2], g1[2] = {0, 0};
        g0[0] = cab_.decode(C_MVD_G0);
        g0[1] = cab_.decode(C_MVD_G0);
        if (g0[0]) g1[0] = cab_.decode(C_MVD_G1);
        if (g0[1]) g1[1] = cab_.decode(C_MVD_G1);
        for (int c = 0; c < 2; c++) {
            int v = 0;
            if (g0[c]) {
                v = 1;
                if (g1[c]) v = 2 + static_cast<int>(exp_golomb(1));
                if (cab_.bypass()) v = -v;
            }
            mvd[c] = v;
        }
    }

    uint32_t exp_golomb(int k) {
        uint32_t v = 0;
        while (cab_.bypass_prefix()) {
            v += 1u << k;
            k++;
            if (k > 31) throw ParseError(-1, "exp-Golomb bypass code too long");
        }
        if (k) v += cab_.bypass_bits(k);
        return v;
    }

    // Prediction block availability (6.4.2) for motion.
    bool pb_avail(int xc, int yc, int cs, int xp, int yp, int w, int h, int part_idx, int xn, int yn) const {
        if (!inside(xn, yn) || !mvok4_[g4(xn, yn)] || !same_region(xn, yn)) return false;
        // NxN second partition: the bottom-left one is decoded later
        if ((w << 1) == cs && (h << 1) == cs && part_idx == 1 && yc + h <= yn && xc + w > xn) return false;
        (void)xp;
        (void)yp;
        return mode4_[g4(xn, yn)] != 1 + HEVCR_MODE_INTRA;
    }

    // Temporal candidate (8.5.3.2.8, 8.5.3.2.9) for list l and target reference index refi.
    bool temporal(int xc, int yc, int xp, int yp, int w, int h, int l, int refi, int16_t* out) const {
        (void)xc;
        if (!col_ || col_->col.empty()) return false;
        auto try_at = [&](int x, int y) {
            const Mv4& c = col_->col[static_cast<size_t>(y >> 4) * col_->col_stride + (x >> 4)];
            if (c.flags & 0x80 || !(c.flags & 3)) return false;
            int lc;
            if (!(c.flags & 1))
                lc = 1;
            else if (!(c.flags & 2))
                lc = 0;
            else
                lc = no_backward_ ? l : (sh_->collocated_from_l0 ? 1 : 0);
            const int rc = c.ref[lc];
            const bool col_lt = col_->ref_lt[lc][rc];
            const bool cur_lt = refs_.lt[l][refi];
            if (col_lt != cur_lt) return false;
            const int col_diff = col_->poc - col_->ref_poc[lc][rc];
            const int cur_diff = cur_.poc - refs_.poc[l][refi];
            for (int k = 0; k < 2; k++)
                out[k] = (cur_lt || col_diff == cur_diff || col_diff == 0) ? c.mv[lc][k]
                                                                           : scale_mv(c.mv[lc][k], col_diff, cur_diff);
            return true;
        };
        const int xb = xp + w, yb = yp + h;
        if ((yc >> sps_.log2_ctb) == (yb >> sps_.log2_ctb) && yb < H_ && xb < W_) {
            if (try_at((xb >> 4) << 4, (yb >> 4) << 4)) return true;
        }
        const int xm = xp + (w >> 1), ym = yp + (h >> 1);
        return try_at((xm >> 4) << 4, (ym >> 4) << 4);
    }

    Mv4 merge_candidate(int xc, int yc, int cs, int xp, int yp, int w, int h, int part_idx, int part, int want) {
        // The 8x4 / 4x8 restriction applies to the unit's size after the
        // singleMCLFlag substitution below, as the reference applies it
        // (inter_prediction.cpp:349-355, 468): under a parallel merge level
        // above 4x4 the units of an 8x8 CU keep bi-prediction.
        if (par_mrg_ > 2 && cs == 8) {  // singleMCLFlag: the CU's list serves every PU
            xp = xc;
            yp = yc;
            w = h = cs;
            part_idx = 0;
        }
        Mv4 list[5];
        int n = 0;
        auto in_mer = [&](int xn, int yn) {
            return (xp >> par_mrg_) == (xn >> par_mrg_) && (yp >> par_mrg_) == (yn >> par_mrg_);
        };
        auto spatial = [&](int xn, int yn, Mv4& m) {
            if (in_mer(xn, yn) || !pb_avail(xc, yc, cs, xp, yp, w, h, part_idx, xn, yn)) return false;
            m = mot4_[g4(xn, yn)];
            return true;
        };
        // Pruning compares with a neighbour that is available, whether or not
        // it entered the list itself (8.5.3.2.3).
        Mv4 a1, b1, b0, a0, b2;
        bool av_a1 = spatial(xp - 1, yp + h - 1, a1);
        if (part_idx == 1 && (part == HEVCR_PART_Nx2N || part == HEVCR_PART_nLx2N || part == HEVCR_PART_nRx2N))
            av_a1 = false;
        bool av_b1 = spatial(xp + w - 1, yp - 1, b1);
        if (part_idx == 1 && (part == HEVCR_PART_2NxN || part == HEVCR_PART_2NxnU || part == HEVCR_PART_2NxnD))
            av_b1 = false;
        const bool av_b0 = spatial(xp + w, yp - 1, b0);
        const bool av_a0 = spatial(xp - 1, yp + h, a0);
        const bool av_b2 = spatial(xp - 1, yp - 1, b2);
        const bool fa1 = av_a1;
        const bool fb1 = av_b1 && !(av_a1 && same_motion(a1, b1));
        const bool fb0 = av_b0 && !(av_b1 && same_motion(b1, b0));
        const bool fa0 = av_a0 && !(av_a1 && same_motion(a1, a0));
        const bool fb2 = av_b2 && !(av_a1 && same_motion(a1, b2)) && !(av_b1 && same_motion(b1, b2)) &&
                         (fa0 + fa1 + fb0 + fb1) != 4;
        if (fa1) list[n++] = a1;
        if (fb1) list[n++] = b1;
        if (fb0) list[n++] = b0;
        if (fa0) list[n++] = a0;
        if (fb2) list[n++] = b2;
        if (n > want) return finish_merge(list[want], w, h);
        // temporal
        if (sh_->temporal_mvp && n < sh_->max_num_merge_cand) {
            Mv4 t;
            int16_t mv[2];
            if (temporal(xc, yc, xp, yp, w, h, 0, 0, mv)) {
                t.flags |= 1;
                t.ref[0] = 0;
                t.mv[0][0] = mv[0];
                t.mv[0][1] = mv[1];
            }
            if (sh_->slice_type == HEVCR_SLICE_B && temporal(xc, yc, xp, yp, w, h, 1, 0, mv)) {
                t.flags |= 2;
                t.ref[1] = 0;
                t.mv[1][0] = mv[0];
                t.mv[1][1] = mv[1];
            }
            if (t.flags & 3) list[n++] = t;
        }
        if (n > want) return finish_merge(list[want], w, h);
        // combined bi-predictive candidates (8.5.3.2.4)
        if (sh_->slice_type == HEVCR_SLICE_B && n > 1 && n < sh_->max_num_merge_cand) {
            static const int kL0[12] = {0, 1, 0, 2, 1, 2, 0, 3, 1, 3, 2, 3};
            static const int kL1[12] = {1, 0, 2, 0, 2, 1, 3, 0, 3, 1, 3, 2};
            const int orig = n;
            for (int ci = 0; ci < orig * (orig - 1) && n < sh_->max_num_merge_cand; ci++) {
                const Mv4& p0 = list[kL0[ci]];
                const Mv4& p1 = list[kL1[ci]];
                if (!(p0.flags & 1) || !(p1.flags & 2)) continue;
                const bool same_pic = refs_.poc[0][p0.ref[0]] == refs_.poc[1][p1.ref[1]] &&
                                      refs_.pic[0][p0.ref[0]] == refs_.pic[1][p1.ref[1]];
                if (same_pic && p0.mv[0][0] == p1.mv[1][0] && p0.mv[0][1] == p1.mv[1][1]) continue;
                Mv4 c;
                c.flags = 3;
                c.ref[0] = p0.ref[0];
                c.ref[1] = p1.ref[1];
                c.mv[0][0] = p0.mv[0][0];
                c.mv[0][1] = p0.mv[0][1];
                c.mv[1][0] = p1.mv[1][0];
                c.mv[1][1] = p1.mv[1][1];
                list[n++] = c;
            }
        }
        if (n > want) return finish_merge(list[want], w, h);
        // zero candidates (8.5.3.2.5)
        const int nref = sh_->slice_type == HEVCR_SLICE_P ? refs_.num[0] : std::min(refs_.num[0], refs_.num[1]);
        int zero = 0;
        while (n <= want) {
            Mv4 z;
            const int r = zero < nref ? zero : 0;
            z.flags = sh_->slice_type == HEVCR_SLICE_P ? 1 : 3;
            z.ref[0] = static_cast<int8_t>(r);
            z.ref[1] = static_cast<int8_t>(sh_->slice_type == HEVCR_SLICE_P ? -1 : r);
            list[n++] = z;
            zero++;
        }
        return finish_merge(list[want], w, h);
    }

    // 8x4 / 4x8 prediction units are never bi-predicted (8.5.3.2.2); the
    // list-1 vector stays in the store (see same_motion).
    static Mv4 finish_merge(Mv4 m, int w, int h) {
        if ((m.flags & 3) == 3 && w + h == 12) {
            m.flags = 1;
            m.ref[1] = -1;
        }
        return m;
    }

    // AMVP (8.5.3.2.6, 8.5.3.2.7): predictor mvp_flag of list l for reference refi.
    void amvp(int xc, int yc, int cs, int xp, int yp, int w, int h, int part_idx, int l, int refi, int flag,
              int16_t* out) {
        const int target_poc = refs_.poc[l][refi];
        const bool target_lt = refs_.lt[l][refi];
        int16_t cand[3][2];
        int n = 0;
        auto nb = [&](int xn, int yn, const Mv4*& m) {
            if (!pb_avail(xc, yc, cs, xp, yp, w, h, part_idx, xn, yn)) return false;
            m = &mot4_[g4(xn, yn)];
            return true;
        };
        // same reference picture in either list, no scaling
        auto unscaled = [&](const Mv4& m, int16_t* mv) {
            for (int k = 0; k < 2; k++) {
                const int ll = k == 0 ? l : 1 - l;
                if ((m.flags & (1 << ll)) && refs_.poc[ll][m.ref[ll]] == target_poc &&
                    refs_.lt[ll][m.ref[ll]] == target_lt) {
                    mv[0] = m.mv[ll][0];
                    mv[1] = m.mv[ll][1];
                    return true;
                }
            }
            return false;
        };
        // any reference of the same long-term-ness, scaled when both are short-term
        auto scaled = [&](const Mv4& m, int16_t* mv) {
            for (int k = 0; k < 2; k++) {
                const int ll = k == 0 ? l : 1 - l;
                if ((m.flags & (1 << ll)) && refs_.lt[ll][m.ref[ll]] == target_lt) {
                    const int rpoc = refs_.poc[ll][m.ref[ll]];
                    if (!target_lt && !refs_.lt[ll][m.ref[ll]] && rpoc != target_poc) {
                        const int td = cur_.poc - rpoc, tb = cur_.poc - target_poc;
                        mv[0] = scale_mv(m.mv[ll][0], td, tb);
                        mv[1] = scale_mv(m.mv[ll][1], td, tb);
                    } else {
                        mv[0] = m.mv[ll][0];
                        mv[1] = m.mv[ll][1];
                    }
                    return true;
                }
            }
            return false;
        };
        const Mv4* ma[2] = {nullptr, nullptr};
        const bool av_a0 = nb(xp - 1, yp + h, ma[0]);
        const bool av_a1 = nb(xp - 1, yp + h - 1, ma[1]);
        const bool is_scaled = av_a0 || av_a1;
        int16_t mva[2] = {0, 0}, mvb[2] = {0, 0};
        bool fa = false;
        for (int k = 0; k < 2 && !fa; k++)
            if (ma[k]) fa = unscaled(*ma[k], mva);
        for (int k = 0; k < 2 && !fa; k++)
            if (ma[k]) fa = scaled(*ma[k], mva);
        const Mv4* mb[3] = {nullptr, nullptr, nullptr};
        nb(xp + w, yp - 1, mb[0]);
        nb(xp + w - 1, yp - 1, mb[1]);
        nb(xp - 1, yp - 1, mb[2]);
        bool fb = false;
        for (int k = 0; k < 3 && !fb; k++)
            if (mb[k]) fb = unscaled(*mb[k], mvb);
        if (!is_scaled && fb && !fa) {
            fa = true;
            mva[0] = mvb[0];
            mva[1] = mvb[1];
        }
        if (!is_scaled) {
            fb = false;
            for (int k = 0; k < 3 && !fb; k++)
                if (mb[k]) fb = scaled(*mb[k], mvb);
        }
        if (fa) {
            cand[n][0] = mva[0];
            cand[n][1] = mva[1];
            n++;
        }
        if (fb && !(fa && mva[0] == mvb[0] && mva[1] == mvb[1])) {
            cand[n][0] = mvb[0];
            cand[n][1] = mvb[1];
            n++;
        }
        if (n < 2 && sh_->temporal_mvp) {
            int16_t t[2];
            if (temporal(xc, yc, xp, yp, w, h, l, refi, t)) {
                cand[n][0] = t[0];
                cand[n][1] = t[1];
                n++;
            }
        }
        while (n < 2) {
            cand[n][0] = cand[n][1] = 0;
            n++;
        }
        out[0] = cand[flag][0];
        out[1] = cand[flag][1];
    }

    // -- transform tree / unit (7.3.8.8, 7.3.8.10) ----------------------------
    void transform_tree(int x0, int y0, int xb, int yb, int log2, int depth, int blk, bool cb_parent, bool cr_parent) {
        bool split;
        if (log2 <= sps_.log2_max_tb && log2 > sps_.log2_min_tb && depth < max_tr_depth_ && !(intra_split_ && depth == 0)) {
            split = cab_.decode(C_SPLIT_TF + 5 - log2) != 0;
        } else {
            const bool inter_split = sps_.max_th_depth_inter == 0 && !cur_intra_ && cu_part_ != HEVCR_PART_2Nx2N && depth == 0;
            split = log2 > sps_.log2_max_tb || (intra_split_ && depth == 0) || inter_split;
        }
        bool cbf_cb = cb_parent, cbf_cr = cr_parent;
        if (log2 > 2) {
            cbf_cb = (depth == 0 || cb_parent) ? cab_.decode(C_CBF_CHROMA + depth) != 0 : false;
            cbf_cr = (depth == 0 || cr_parent) ? cab_.decode(C_CBF_CHROMA + depth) != 0 : false;
        }
        if (split) {
            const int h = 1 << (log2 - 1);
            for (int k = 0; k < 4; k++)
                transform_tree(x0 + (k & 1) * h, y0 + (k >> 1) * h, x0, y0, log2 - 1, depth + 1, k, cbf_cb, cbf_cr);
            return;
        }
        bool cbf_luma = true;
        if (cur_intra_ || depth != 0 || cbf_cb || cbf_cr) cbf_luma = cab_.decode(C_CBF_LUMA + (depth == 0 ? 1 : 0)) != 0;
        hevcr_tu t{};
        t.x = static_cast<uint16_t>(x0);
        t.y = static_cast<uint16_t>(y0);
        t.xbase = static_cast<uint16_t>(depth ? xb : x0);
        t.ybase = static_cast<uint16_t>(depth ? yb : y0);
        t.log2_size = static_cast<uint8_t>(log2);
        t.blk_idx = static_cast<uint8_t>(depth ? blk : 0);
        t.cu_idx = cur_cu_;
        const bool chroma_here = log2 > 2 || blk == 3;
        if (cbf_luma || cbf_cb || cbf_cr) {
            if (pps_.cu_qp_delta && !qp_delta_coded_) {
                int v = 0;
                while (v < 5 && cab_.decode(C_QP_DELTA + (v ? 1 : 0))) v++;
                if (v == 5) v += static_cast<int>(exp_golomb(0));
                if (v && cab_.bypass()) v = -v;
                qp_delta_ = v;
                qp_delta_coded_ = true;
                if (v < -(26 + qbd_ / 2) || v > 25 + qbd_ / 2) throw ParseError(-1, "CuQpDeltaVal");
            }
            if (cbf_luma) {
                t.cbf |= HEVCR_CBF_Y;
                t.tskip |= static_cast<uint8_t>(residual(x0, y0, log2, 0, t.coeff_off[0]));
            }
            if (chroma_here) {
                const int xc = log2 > 2 ? x0 : xb, yc = log2 > 2 ? y0 : yb, lc = log2 > 2 ? log2 - 1 : 2;
                if (cbf_cb) {
                    t.cbf |= HEVCR_CBF_CB;
                    t.tskip |= static_cast<uint8_t>(residual(xc, yc, lc, 1, t.coeff_off[1]) << 1);
                }
                if (cbf_cr) {
                    t.cbf |= HEVCR_CBF_CR;
                    t.tskip |= static_cast<uint8_t>(residual(xc, yc, lc, 2, t.coeff_off[2]) << 2);
                }
            }
        }
        out_.tus.push_back(t);
    }

    // -- residual coding (7.3.8.11) ------------------------------------------
    // Appends the dense NxN block (row-major) to the pool; returns transform_skip_flag.
    int residual(int x0, int y0, int log2, int c, uint32_t& off) {
        const int n = 1 << log2;
        off = static_cast<uint32_t>(out_.coeffs.size());
        out_.coeffs.resize(out_.coeffs.size() + static_cast<size_t>(n) * n, 0);
        int16_t* blk = out_.coeffs.data() + off;
        int tskip = 0;
        if (pps_.transform_skip && !cur_bypass_ && log2 <= 2) tskip = cab_.decode(C_TSKIP + (c ? 1 : 0));
        // last significant position
        const int cmax = (log2 << 1) - 1;
        const int coff = c ? 15 : 3 * (log2 - 2) + ((log2 - 1) >> 2), cshift = c ? log2 - 2 : (log2 + 1) >> 2;
        int px = 0, py = 0;
        while (px < cmax && cab_.decode(C_LAST_X + coff + (px >> cshift))) px++;
        while (py < cmax && cab_.decode(C_LAST_Y + coff + (py >> cshift))) py++;
        auto suffix = [&](int p) {
            if (p <= 3) return p;
            const int nb = (p >> 1) - 1;
            return (1 << nb) * (2 + (p & 1)) + static_cast<int>(cab_.bypass_bits(nb));
        };
        int lx = suffix(px), ly = suffix(py);
        // scan order (7.4.9.11): intra 4x4 / 8x8 luma and 4x4 chroma follow the prediction direction
        int scan = 0;
        if (cur_intra_ && (log2 == 2 || (log2 == 3 && c == 0))) {
            const hevcr_cu& cu = out_.cus[cur_cu_];
            int mode;
            if (c == 0) {
                mode = ipm4_[g4(x0, y0)];
            } else {
                mode = cu.intra_chroma;
            }
            if (mode >= 6 && mode <= 14)
                scan = 2;
            else if (mode >= 22 && mode <= 30)
                scan = 1;
        }
        if (scan == 2) std::swap(lx, ly);
        if (lx >= n || ly >= n) throw ParseError(-1, "last significant coefficient outside the block");
        const int lsb = log2 - 2;  // log2 of the sub-block grid
        const uint8_t* sb_scan = kScan.pos[lsb][scan];
        const uint8_t* in_scan = kScan.pos[2][scan];
        int last_sb = 0, last_pos = 0;
        {
            const int xs = lx >> 2, ys = ly >> 2, xi = lx & 3, yi = ly & 3;
            for (int i = 0; i < (1 << (2 * lsb)); i++)
                if (sb_scan[i] == (xs | ys << 4)) last_sb = i;
            for (int i = 0; i < 16; i++)
                if (in_scan[i] == (xi | yi << 4)) last_pos = i;
        }
        uint8_t csbf[8][8] = {};
        int greater1_state = 1;  // greater1Ctx carried between sub-blocks (9.3.4.2.6)
        bool first_sb_done = false;
        const bool sdh = pps_.sign_data_hiding && !cur_bypass_;
        for (int i = last_sb; i >= 0; i--) {
            const int xs = sb_scan[i] & 15, ys = sb_scan[i] >> 4;
            bool infer_dc = false;
            if (i < last_sb && i > 0) {
                int inc = 0;
                if (xs + 1 < (1 << lsb)) inc += csbf[xs + 1][ys];
                if (ys + 1 < (1 << lsb)) inc += csbf[xs][ys + 1];
                csbf[xs][ys] = static_cast<uint8_t>(cab_.decode(C_CSBF + std::min(inc, 1) + (c ? 2 : 0)));
                infer_dc = true;
            } else {
                csbf[xs][ys] = 1;
            }
            const int prev_csbf = ((xs + 1 < (1 << lsb)) ? csbf[xs + 1][ys] : 0) |
                                  (((ys + 1 < (1 << lsb)) ? csbf[xs][ys + 1] : 0) << 1);
            uint8_t sig[16] = {};
            int nsig = 0, sig_pos[16];
            const int start = i == last_sb ? last_pos - 1 : 15;
            if (i == last_sb) {
                sig[last_pos] = 1;
                sig_pos[nsig++] = last_pos;
            }
            if (csbf[xs][ys]) {
                for (int p = start; p >= 0; p--) {
                    const int xp = in_scan[p] & 15, yp = in_scan[p] >> 4;
                    const int xc = (xs << 2) + xp, yc = (ys << 2) + yp;
                    if (p > 0 || !infer_dc) {
                        sig[p] = static_cast<uint8_t>(cab_.decode(C_SIG + sig_ctx(log2, c, scan, xc, yc, xp, yp, xs, ys, prev_csbf)));
                        if (sig[p]) infer_dc = false;
                    } else {
                        sig[p] = 1;  // inferred DC of a coded sub-block with no other coefficient
                    }
                    if (sig[p]) sig_pos[nsig++] = p;
                }
            }
            if (!nsig) continue;
            // greater1 / greater2 flags
            int ctx_set = (i == 0 || c > 0) ? 0 : 2;
            if (first_sb_done && greater1_state == 0) ctx_set++;
            first_sb_done = true;
            int g1ctx = 1;
            uint8_t g1[16] = {}, g2[16] = {};
            int first_g1 = -1;
            for (int k = 0; k < nsig && k < 8; k++) {
                const int p = sig_pos[k];
                g1[p] = static_cast<uint8_t>(cab_.decode(C_GT1 + ctx_set * 4 + g1ctx + (c ? 16 : 0)));
                if (g1[p]) {
                    g1ctx = 0;
                    if (first_g1 < 0) first_g1 = p;
                } else if (g1ctx > 0 && g1ctx < 3) {
                    g1ctx++;
                }
            }
            greater1_state = g1ctx;
            if (first_g1 >= 0) g2[first_g1] = static_cast<uint8_t>(cab_.decode(C_GT2 + ctx_set + (c ? 4 : 0)));
            const int last_sig = sig_pos[0], first_sig = sig_pos[nsig - 1];
            const bool hidden = sdh && (last_sig - first_sig > 3);
            uint8_t sign[16] = {};
            for (int k = 0; k < nsig; k++) {
                const int p = sig_pos[k];
                if (!hidden || p != first_sig) sign[p] = static_cast<uint8_t>(cab_.bypass());
            }
            int rice = 0, sum = 0, nsc = 0;
            for (int k = 0; k < nsig; k++) {
                const int p = sig_pos[k];
                const int base = 1 + g1[p] + g2[p];
                int level = base;
                if (base == ((nsc < 8) ? ((p == first_g1) ? 3 : 2) : 1)) {
                    const int rem = abs_level_remaining(rice);
                    level = base + rem;
                    if (level > 3 * (1 << rice)) rice = std::min(rice + 1, 4);
                }
                sum += level;
                int v = sign[p] ? -level : level;
                if (hidden && p == first_sig && (sum & 1)) v = -v;
                const int xc = (xs << 2) + (in_scan[p] & 15), yc = (ys << 2) + (in_scan[p] >> 4);
                blk[yc * n + xc] = static_cast<int16_t>(v);
                nsc++;
            }
        }
        return tskip;
    }

    static int sig_ctx(int log2, int c, int scan, int xc, int yc, int xp, int yp, int xs, int ys, int prev_csbf) {
        static const uint8_t kMap4[16] = {0, 1, 4, 5, 2, 3, 4, 5, 6, 6, 8, 8, 7, 7, 8, 8};
        int s;
        if (log2 == 2) {
            s = kMap4[(yc << 2) + xc];
        } else if (xc + yc == 0) {
            s = 0;
        } else {
            if (prev_csbf == 0)
                s = (xp + yp == 0) ? 2 : (xp + yp < 3) ? 1 : 0;
            else if (prev_csbf == 1)
                s = yp == 0 ? 2 : (yp == 1 ? 1 : 0);
            else if (prev_csbf == 2)
                s = xp == 0 ? 2 : (xp == 1 ? 1 : 0);
            else
                s = 2;
            if (c == 0) {
                if (xs > 0 || ys > 0) s += 3;
                s += log2 == 3 ? (scan == 0 ? 9 : 15) : 21;
            } else {
                s += log2 == 3 ? 9 : 12;
            }
        }
        return c == 0 ? s : 27 + s;
    }

    int abs_level_remaining(int rice) {
        int prefix = 0;
        while (cab_.bypass_prefix()) {
            if (++prefix > 32) throw ParseError(-1, "coeff_abs_level_remaining prefix too long");
        }
        if (prefix <= 3) return (prefix << rice) + static_cast<int>(cab_.bypass_bits(rice));
        const int k = prefix - 3 + rice;
        return (((1 << (prefix - 3)) + 2) << rice) + static_cast<int>(cab_.bypass_bits(k));
    }

    // Motion kept for later pictures' temporal prediction: the 4x4 block at the
    // top-left of each 16x16 block (8.5.3.2.8), plus this picture's lists.
    void store_col() {
        const int w16 = (W_ + 15) >> 4, h16 = (H_ + 15) >> 4;
        cur_.col_stride = w16;
        cur_.col.assign(static_cast<size_t>(w16) * h16, Mv4{});
        for (int y = 0; y < h16; y++)
            for (int x = 0; x < w16; x++) {
                const size_t g = g4(x << 4, y << 4);
                Mv4 m = mot4_[g];
                if (mode4_[g] == 1 + HEVCR_MODE_INTRA || !mvok4_[g]) m.flags = 0x80;
                cur_.col[static_cast<size_t>(y) * w16 + x] = m;
            }
        for (int l = 0; l < 2; l++) {
            cur_.num_ref[l] = refs_.num[l];
            for (int i = 0; i < refs_.num[l]; i++) {
                cur_.ref_poc[l][i] = refs_.poc[l][i];
                cur_.ref_lt[l][i] = refs_.lt[l][i];
            }
        }
    }

    const Sps& sps_;
    const Pps& pps_;
    const SliceHeader* sh_;  // the segment being parsed
    const RefLists& refs_;
    DpbPic& cur_;
    PictureRecords& out_;
    Engine& cab_;
    DpbPic* col_ = nullptr;
    std::vector<uint16_t> tile_;         // per CTB (raster): tile index
    std::vector<int> rs2ts_, ts2rs_;     // CTB scan
    std::vector<uint8_t> ctb_seg_;       // per CTB: slice segment (set as it is parsed)
    int nseg_ = 0, seg_ = 0, next_ts_ = 0;
    int init_type_ = 0;
    int W_ = 0, H_ = 0, w4_ = 0, h4_ = 0;
    std::vector<uint8_t> mode4_, ipm4_, mvok4_, depth4_;
    std::vector<int8_t> qp4_;
    std::vector<Mv4> mot4_;
    int qbd_ = 0, slice_qp_ = 26, log2_qg_ = 6, par_mrg_ = 2;
    bool no_backward_ = true;
    int ctb_x0_ = 0, ctb_y0_ = 0, ctb_addr_ = 0;
    bool qp_delta_coded_ = false;
    int qp_delta_ = 0, qpy_pred_ = 26, last_qpy_ = 26;
    uint32_t cur_cu_ = 0;
    bool cur_intra_ = false, intra_split_ = false, cur_bypass_ = false;
    int max_tr_depth_ = 0, cu_part_ = 0;
};

}  // namespace sd
}  // namespace hevcf
