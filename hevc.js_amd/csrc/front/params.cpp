// Host front end: parameter sets and slice headers (H.265 7.3.2-7.3.6 and
// their semantics in 7.4.3-7.4.7).
#include "params.hpp"

#include <algorithm>
#include <cstring>

namespace hevcf {

namespace {

constexpr int kUnsupported = -6;  // HEVCR_EUNSUPPORTED

int ceil_log2(uint32_t v) {
    int n = 0;
    while ((1u << n) < v) n++;
    return n;
}

// (x, y) of scan position i of the up-right diagonal scan of a size x size block (6.5.3).
void diag_scan(int size, int* xs, int* ys) {
    int i = 0, x = 0, y = 0;
    while (i < size * size) {
        while (y >= 0) {
            if (x < size && y < size) {
                xs[i] = x;
                ys[i] = y;
                i++;
            }
            y--;
            x++;
        }
        y = x;
        x = 0;
    }
}

void profile_tier_level(Bits& b, int max_sub_layers_minus1, int* profile, int* level) {
    b.u(2);  // general_profile_space
    b.u(1);  // general_tier_flag
    *profile = static_cast<int>(b.u(5));
    b.skip(32 + 4 + 43 + 1);
    *level = static_cast<int>(b.u(8));
    bool prof[8] = {}, lev[8] = {};
    for (int i = 0; i < max_sub_layers_minus1; i++) {
        prof[i] = b.flag();
        lev[i] = b.flag();
    }
    if (max_sub_layers_minus1 > 0)
        for (int i = max_sub_layers_minus1; i < 8; i++) b.u(2);
    for (int i = 0; i < max_sub_layers_minus1; i++) {
        if (prof[i]) b.skip(88);
        if (lev[i]) b.skip(8);
    }
}

// scaling_list_data() (7.3.4). Lists are kept as coded: the 4x4 list and the
// 8x8 base lists in coded order, which is the order perform_dequant indexes
// them by raster position (transform.cpp:343-350 with sps.cpp:105-109).
void scaling_list_data(Bits& b, hevcr_scaling& s) {
    default_scaling(s);
    for (int size = 0; size < 4; size++) {
        const int step = size == 3 ? 3 : 1;
        for (int m = 0; m < 6; m += step) {
            if (!b.flag()) {  // scaling_list_pred_mode_flag == 0: copy (or default for delta 0)
                const uint32_t delta = b.ue();
                if (delta) {
                    const int ref = m - static_cast<int>(delta) * step;
                    if (ref < 0) throw ParseError(-1, "scaling_list_pred_matrix_id_delta");
                    std::memcpy(s.list[size][m], s.list[size][ref], 64);
                    if (size > 1) s.dc[size - 2][m] = s.dc[size - 2][ref];
                }
            } else {
                int next = 8;
                const int n = std::min(64, 1 << (4 + (size << 1)));
                if (size > 1) {
                    next = b.se() + 8;
                    s.dc[size - 2][m] = static_cast<uint8_t>(next);
                }
                for (int i = 0; i < n; i++) {
                    next = (next + b.se() + 256) % 256;
                    s.list[size][m][i] = static_cast<uint8_t>(next);
                }
            }
        }
    }
}

void st_ref_pic_set(Bits& b, int idx, const std::vector<StRps>& sets, int num_sets, StRps& r) {
    bool inter = idx != 0 && b.flag();
    if (inter) {
        int delta_idx = 1;
        if (idx == num_sets) delta_idx = static_cast<int>(b.ue()) + 1;
        if (delta_idx > idx) throw ParseError(-1, "delta_idx_minus1");
        const StRps& R = sets[static_cast<size_t>(idx - delta_idx)];
        const int sign = b.flag() ? 1 : 0;
        const int delta_rps = (1 - 2 * sign) * (static_cast<int>(b.ue()) + 1);
        bool used[33] = {}, use_delta[33] = {};
        for (int j = 0; j <= R.num_delta(); j++) {
            used[j] = b.flag();
            use_delta[j] = used[j] ? true : b.flag();
        }
        // (7-61), (7-62). A predicted set can hold one entry more than the set
        // it is predicted from; more than 16 entries per direction (or in
        // total) is not a valid bitstream and would overrun the arrays.
        int n0 = 0, n1 = 0;
        auto push0 = [&](int d, bool u) {
            if (n0 == 16) throw ParseError(-1, "predicted short-term RPS too large");
            r.delta_s0[n0] = d;
            r.used_s0[n0++] = u;
        };
        auto push1 = [&](int d, bool u) {
            if (n1 == 16) throw ParseError(-1, "predicted short-term RPS too large");
            r.delta_s1[n1] = d;
            r.used_s1[n1++] = u;
        };
        for (int j = R.num_positive - 1; j >= 0; j--) {
            const int d = R.delta_s1[j] + delta_rps;
            if (d < 0 && use_delta[R.num_negative + j]) push0(d, used[R.num_negative + j]);
        }
        if (delta_rps < 0 && use_delta[R.num_delta()]) push0(delta_rps, used[R.num_delta()]);
        for (int j = 0; j < R.num_negative; j++) {
            const int d = R.delta_s0[j] + delta_rps;
            if (d < 0 && use_delta[j]) push0(d, used[j]);
        }
        for (int j = R.num_negative - 1; j >= 0; j--) {
            const int d = R.delta_s0[j] + delta_rps;
            if (d > 0 && use_delta[j]) push1(d, used[j]);
        }
        if (delta_rps > 0 && use_delta[R.num_delta()]) push1(delta_rps, used[R.num_delta()]);
        for (int j = 0; j < R.num_positive; j++) {
            const int d = R.delta_s1[j] + delta_rps;
            if (d > 0 && use_delta[R.num_negative + j]) push1(d, used[R.num_negative + j]);
        }
        r.num_negative = n0;
        r.num_positive = n1;
        if (r.num_delta() > 16) throw ParseError(-1, "predicted short-term RPS too large");
    } else {
        r.num_negative = static_cast<int>(b.ue());
        r.num_positive = static_cast<int>(b.ue());
        if (r.num_negative > 16 || r.num_positive > 16 || r.num_delta() > 16)
            throw ParseError(-1, "short-term RPS too large");
        int poc = 0;
        for (int i = 0; i < r.num_negative; i++) {
            poc -= static_cast<int>(b.ue()) + 1;
            r.delta_s0[i] = poc;
            r.used_s0[i] = b.flag();
        }
        poc = 0;
        for (int i = 0; i < r.num_positive; i++) {
            poc += static_cast<int>(b.ue()) + 1;
            r.delta_s1[i] = poc;
            r.used_s1[i] = b.flag();
        }
    }
}

// 7.4.8: num_negative_pics <= sps_max_dec_pic_buffering_minus1 and
// num_positive_pics <= sps_max_dec_pic_buffering_minus1 - num_negative_pics.
void check_rps_size(const StRps& r, const Sps& sps) {
    const int room = sps.max_dec_pic_buffering - 1;
    if (r.num_negative > room || r.num_positive > room - r.num_negative)
        throw ParseError(-1, "short-term RPS larger than the DPB");
}

void pred_weight_table(Bits& b, const Sps& sps, SliceHeader& sh) {
    sh.luma_log2_wd = static_cast<int>(b.ue());
    if (sh.luma_log2_wd > 7) throw ParseError(-1, "luma_log2_weight_denom");
    sh.chroma_log2_wd = sh.luma_log2_wd + (sps.chroma_format_idc ? b.se() : 0);
    if (sh.chroma_log2_wd < 0 || sh.chroma_log2_wd > 7) throw ParseError(-1, "chroma weight denominator");
    const int nl = sh.slice_type == HEVCR_SLICE_B ? 2 : 1;
    for (int l = 0; l < nl; l++) {
        bool lf[16] = {}, cf[16] = {};
        for (int i = 0; i < sh.num_ref_idx[l]; i++) lf[i] = b.flag();
        if (sps.chroma_format_idc)
            for (int i = 0; i < sh.num_ref_idx[l]; i++) cf[i] = b.flag();
        for (int i = 0; i < sh.num_ref_idx[l]; i++) {
            WeightEntry& w = sh.weights[l][i];
            w.w[0] = static_cast<int16_t>(1 << sh.luma_log2_wd);
            w.o[0] = 0;
            if (lf[i]) {
                w.w[0] = static_cast<int16_t>(w.w[0] + b.se());
                w.o[0] = static_cast<int16_t>(b.se());
            }
            for (int c = 1; c <= 2; c++) {
                w.w[c] = static_cast<int16_t>(1 << sh.chroma_log2_wd);
                w.o[c] = 0;
                if (cf[i]) {
                    w.w[c] = static_cast<int16_t>(w.w[c] + b.se());
                    const int d = b.se();
                    // (7-56): ChromaOffset = Clip3(-128, 127, 128 + d - ((128 * ChromaWeight) >> denom))
                    w.o[c] = static_cast<int16_t>(std::clamp(128 + d - ((128 * w.w[c]) >> sh.chroma_log2_wd), -128, 127));
                }
            }
        }
    }
}

}  // namespace

void default_scaling(hevcr_scaling& s) {
    // Table 7-6: default 8x8 factors in up-right diagonal scan order.
    static const uint8_t kIntra[64] = {16, 16, 16, 16, 16, 16, 16, 16, 16, 16, 17, 16, 17, 16, 17, 18,
                                       17, 18, 18, 17, 18, 21, 19, 20, 21, 20, 19, 21, 24, 22, 22, 24,
                                       24, 22, 22, 24, 25, 25, 27, 30, 27, 25, 25, 29, 31, 35, 35, 31,
                                       29, 36, 41, 44, 41, 36, 47, 54, 54, 47, 65, 70, 65, 88, 88, 115};
    static const uint8_t kInter[64] = {16, 16, 16, 16, 16, 16, 16, 16, 16, 16, 17, 17, 17, 17, 17, 18,
                                       18, 18, 18, 18, 18, 20, 20, 20, 20, 20, 20, 20, 24, 24, 24, 24,
                                       24, 24, 24, 24, 25, 25, 25, 25, 25, 25, 25, 28, 28, 28, 28, 28,
                                       28, 33, 33, 33, 33, 33, 41, 41, 41, 41, 54, 54, 54, 71, 71, 91};
    int xs[64], ys[64];
    diag_scan(8, xs, ys);
    std::memset(&s, 0, sizeof(s));
    for (int m = 0; m < 6; m++) {
        for (int i = 0; i < 16; i++) s.list[0][m][i] = 16;  // Table 7-5: flat 4x4
        for (int size = 1; size < 4; size++) {
            if (size == 3 && m % 3) continue;  // 32x32: matrixId 0 and 3 only
            for (int i = 0; i < 64; i++) s.list[size][m][ys[i] * 8 + xs[i]] = (m < 3 ? kIntra : kInter)[i];
        }
    }
    for (int k = 0; k < 2; k++)
        for (int m = 0; m < 6; m++) s.dc[k][m] = 16;
}

int SliceHeader::num_pic_total_curr() const {
    int n = 0;
    for (int i = 0; i < st_rps.num_negative; i++) n += st_rps.used_s0[i];
    for (int i = 0; i < st_rps.num_positive; i++) n += st_rps.used_s1[i];
    for (int i = 0; i < num_lt; i++) n += lt_used[i];
    return n;
}

void parse_sps(Bits& b, Sps& s) {
    b.u(4);  // sps_video_parameter_set_id
    s.max_sub_layers = static_cast<int>(b.u(3)) + 1;
    b.u(1);  // sps_temporal_id_nesting_flag
    profile_tier_level(b, s.max_sub_layers - 1, &s.profile_idc, &s.level_idc);
    s.id = static_cast<int>(b.ue());
    if (s.id > 15) throw ParseError(-1, "sps id");
    s.chroma_format_idc = static_cast<int>(b.ue());
    if (s.chroma_format_idc > 3) throw ParseError(-1, "chroma_format_idc");
    if (s.chroma_format_idc == 3) b.u(1);
    s.width = static_cast<int>(b.ue());
    s.height = static_cast<int>(b.ue());
    if (b.flag()) {  // conformance_window_flag: offsets in chroma units (4:2:0: 2 luma samples)
        const int sub = s.chroma_format_idc == 1 || s.chroma_format_idc == 2 ? 2 : 1;
        const int subh = s.chroma_format_idc == 1 ? 2 : 1;
        s.conf[0] = static_cast<int>(b.ue()) * sub;
        s.conf[1] = static_cast<int>(b.ue()) * sub;
        s.conf[2] = static_cast<int>(b.ue()) * subh;
        s.conf[3] = static_cast<int>(b.ue()) * subh;
    }
    s.bit_depth = static_cast<int>(b.ue()) + 8;
    s.bit_depth_c = static_cast<int>(b.ue()) + 8;
    s.log2_max_poc_lsb = static_cast<int>(b.ue()) + 4;
    const bool ordering_all = b.flag();
    for (int i = ordering_all ? 0 : s.max_sub_layers - 1; i < s.max_sub_layers; i++) {
        s.max_dec_pic_buffering = static_cast<int>(b.ue()) + 1;
        s.max_num_reorder = static_cast<int>(b.ue());
        s.max_latency_increase_plus1 = static_cast<int>(b.ue());
    }
    s.log2_min_cb = static_cast<int>(b.ue()) + 3;
    s.log2_ctb = s.log2_min_cb + static_cast<int>(b.ue());
    s.log2_min_tb = static_cast<int>(b.ue()) + 2;
    s.log2_max_tb = s.log2_min_tb + static_cast<int>(b.ue());
    s.max_th_depth_inter = static_cast<int>(b.ue());
    s.max_th_depth_intra = static_cast<int>(b.ue());
    if (s.log2_ctb > 6 || s.log2_ctb < 4 || s.log2_max_tb > 5 || s.log2_min_tb > s.log2_max_tb)
        throw ParseError(kUnsupported, "coding block sizes outside CTB 16..64 / TB 4..32");
    s.scaling_list_enabled = b.flag();
    if (s.scaling_list_enabled) {
        if (b.flag())
            scaling_list_data(b, s.scaling);
        else
            default_scaling(s.scaling);
    }
    s.amp = b.flag();
    s.sao = b.flag();
    s.pcm = b.flag();
    if (s.pcm) {
        s.pcm_bits_luma = static_cast<int>(b.u(4)) + 1;
        s.pcm_bits_chroma = static_cast<int>(b.u(4)) + 1;
        s.pcm_log2_min = static_cast<int>(b.ue()) + 3;
        s.pcm_log2_max = s.pcm_log2_min + static_cast<int>(b.ue());
        s.pcm_loop_filter_disabled = b.flag();
    }
    const int nsets = static_cast<int>(b.ue());
    if (nsets > 64) throw ParseError(-1, "num_short_term_ref_pic_sets");
    s.st_rps.assign(static_cast<size_t>(nsets), StRps{});
    for (int i = 0; i < nsets; i++) st_ref_pic_set(b, i, s.st_rps, nsets, s.st_rps[static_cast<size_t>(i)]);
    for (const StRps& r : s.st_rps) check_rps_size(r, s);
    s.long_term_refs_present = b.flag();
    if (s.long_term_refs_present) {
        const int n = static_cast<int>(b.ue());
        if (n > 32) throw ParseError(-1, "num_long_term_ref_pics_sps");
        for (int i = 0; i < n; i++) {
            s.lt_poc_lsb.push_back(b.u(s.log2_max_poc_lsb));
            s.lt_used.push_back(b.flag());
        }
    }
    s.temporal_mvp = b.flag();
    s.strong_intra_smoothing = b.flag();
    // VUI and extensions follow; nothing after this point affects reconstruction in Main / Main 10.
    // Constraints of 7.4.3.2 the parser relies on (a stream breaking them is rejected, not followed):
    const int min_cb = 1 << s.log2_min_cb;
    if (s.width <= 0 || s.height <= 0 || s.width > 8192 || s.height > 4352 || s.width % min_cb || s.height % min_cb)
        throw ParseError(-1, "picture size (a multiple of the minimum CB, at most 8192x4352)");
    if (s.conf[0] + s.conf[1] >= s.width || s.conf[2] + s.conf[3] >= s.height) throw ParseError(-1, "conformance window");
    if (s.log2_min_tb >= s.log2_min_cb || s.log2_max_tb > std::min(s.log2_ctb, 5))
        throw ParseError(-1, "transform block sizes");
    if (s.max_th_depth_inter > s.log2_ctb - s.log2_min_tb || s.max_th_depth_intra > s.log2_ctb - s.log2_min_tb)
        throw ParseError(-1, "max_transform_hierarchy_depth");
    if (s.bit_depth > 16 || s.bit_depth_c > 16 || s.log2_max_poc_lsb > 16 || s.max_dec_pic_buffering > 16 ||
        s.max_num_reorder > s.max_dec_pic_buffering - 1)
        throw ParseError(-1, "SPS value out of range");
    if (s.pcm && (s.pcm_log2_min < 3 || s.pcm_log2_max > std::min(s.log2_ctb, 5))) throw ParseError(-1, "PCM sizes");
    if (s.pcm && (s.pcm_bits_luma > s.bit_depth || s.pcm_bits_chroma > s.bit_depth_c))
        throw ParseError(-1, "PCM sample bit depth above the bit depth");
}

void parse_pps(Bits& b, Pps& p, const SpsTable& sps) {
    p.id = static_cast<int>(b.ue());
    p.sps_id = static_cast<int>(b.ue());
    if (p.id > 63 || p.sps_id > 15) throw ParseError(-1, "pps / sps id");
    p.dependent_slices = b.flag();
    p.output_flag_present = b.flag();
    p.num_extra_slice_header_bits = static_cast<int>(b.u(3));
    p.sign_data_hiding = b.flag();
    p.cabac_init_present = b.flag();
    p.num_ref_idx_default[0] = static_cast<int>(b.ue()) + 1;
    p.num_ref_idx_default[1] = static_cast<int>(b.ue()) + 1;
    p.init_qp = 26 + b.se();
    p.constrained_intra_pred = b.flag();
    p.transform_skip = b.flag();
    p.cu_qp_delta = b.flag();
    if (p.cu_qp_delta) p.diff_cu_qp_delta_depth = static_cast<int>(b.ue());
    p.cb_qp_offset = b.se();
    p.cr_qp_offset = b.se();
    p.slice_chroma_qp_offsets_present = b.flag();
    p.weighted_pred = b.flag();
    p.weighted_bipred = b.flag();
    p.transquant_bypass = b.flag();
    p.tiles = b.flag();
    p.entropy_sync = b.flag();
    if (p.tiles) {
        const uint32_t cols = b.ue() + 1, rows = b.ue() + 1;
        if (cols > HEVCR_MAX_TILE_COLS || rows > HEVCR_MAX_TILE_ROWS) throw ParseError(-1, "tile columns / rows");
        p.tile_cols = static_cast<int>(cols);
        p.tile_rows = static_cast<int>(rows);
        p.tile_uniform = b.flag();
        if (!p.tile_uniform) {
            for (uint32_t i = 0; i + 1 < cols; i++) p.tile_col_width.push_back(static_cast<int>(b.ue()) + 1);
            for (uint32_t i = 0; i + 1 < rows; i++) p.tile_row_height.push_back(static_cast<int>(b.ue()) + 1);
        }
        p.loop_filter_across_tiles = b.flag();
    }
    p.loop_filter_across_slices = b.flag();
    if (b.flag()) {  // deblocking_filter_control_present_flag
        p.deblocking_override_enabled = b.flag();
        p.deblocking_disabled = b.flag();
        if (!p.deblocking_disabled) {
            p.beta_offset_div2 = b.se();
            p.tc_offset_div2 = b.se();
        }
    }
    p.scaling_list_present = b.flag();
    if (p.scaling_list_present) scaling_list_data(b, p.scaling);
    p.lists_modification_present = b.flag();
    p.log2_parallel_merge_level = static_cast<int>(b.ue()) + 2;
    p.slice_header_extension_present = b.flag();
    (void)sps;
}

// entry points, slice header extension and byte_alignment(): the end of every slice segment header
void entry_points_and_end(Bits& b, const Pps& pps, SliceHeader& sh);

void tile_sizes(const Sps& sps, const Pps& pps, std::vector<int>& cols, std::vector<int>& rows) {
    const int wc = sps.width_ctbs(), hc = sps.height_ctbs();
    auto sizes = [](int total, int n, bool uniform, const std::vector<int>& given, std::vector<int>& out) {
        out.assign(static_cast<size_t>(n), 0);
        int sum = 0;
        for (int i = 0; i < n; i++) {
            if (uniform)
                out[static_cast<size_t>(i)] = ((i + 1) * total) / n - (i * total) / n;
            else if (i + 1 < n)
                out[static_cast<size_t>(i)] = given[static_cast<size_t>(i)];
            else
                out[static_cast<size_t>(i)] = total - sum;
            sum += out[static_cast<size_t>(i)];
            if (out[static_cast<size_t>(i)] < 1) throw ParseError(-1, "tile sizes");
        }
    };
    sizes(wc, pps.tiles ? pps.tile_cols : 1, pps.tile_uniform, pps.tile_col_width, cols);
    sizes(hc, pps.tiles ? pps.tile_rows : 1, pps.tile_uniform, pps.tile_row_height, rows);
}

void parse_slice_header(Bits& b, int nal_type, const SpsTable& spst, const PpsTable& ppst, SliceHeader& sh,
                        const SliceHeader* independent) {
    sh = SliceHeader{};
    sh.first_slice = b.flag();
    if (nal_type >= NAL_BLA_W_LP && nal_type <= 23) sh.no_output_of_prior_pics = b.flag();
    sh.pps_id = static_cast<int>(b.ue());
    if (sh.pps_id > 63 || !ppst[static_cast<size_t>(sh.pps_id)]) throw ParseError(-1, "slice refers to a missing PPS");
    const Pps& pps = *ppst[static_cast<size_t>(sh.pps_id)];
    if (!spst[static_cast<size_t>(pps.sps_id)]) throw ParseError(-1, "PPS refers to a missing SPS");
    const Sps& sps = *spst[static_cast<size_t>(pps.sps_id)];
    if (!sh.first_slice) {
        if (pps.dependent_slices) sh.dependent = b.flag();
        sh.segment_address = b.u(ceil_log2(static_cast<uint32_t>(sps.width_ctbs() * sps.height_ctbs())));
        if (sh.segment_address >= static_cast<uint32_t>(sps.width_ctbs() * sps.height_ctbs()))
            throw ParseError(-1, "slice_segment_address");
    }
    if (sh.dependent) {
        // The independent segment's fields, its entry points included. The
        // reference reads no more of a dependent segment's header
        // (slice_header.cpp:141-144): the slice data starts at the next byte
        // boundary, or right there when slice_segment_address ends on one
        // (decoder.cpp:243-244), not after the entry points and
        // byte_alignment() of 7.3.6.1; the slice data is read from the same
        // place here.
        if (!independent || independent->pps_id != sh.pps_id) throw ParseError(-1, "dependent slice segment without its slice");
        const uint32_t address = sh.segment_address;
        sh = *independent;
        sh.first_slice = false;
        sh.dependent = true;
        sh.segment_address = address;
        b.align();
        sh.data_offset = b.pos() / 8;
        return;
    }
    for (int i = 0; i < pps.num_extra_slice_header_bits; i++) b.u(1);
    sh.slice_type = static_cast<int>(b.ue());
    if (sh.slice_type > 2) throw ParseError(-1, "slice_type");
    if (pps.output_flag_present) sh.pic_output = b.flag();
    const bool idr = nal_type == NAL_IDR_W_RADL || nal_type == NAL_IDR_N_LP;
    if (!idr) {
        sh.poc_lsb = b.u(sps.log2_max_poc_lsb);
        const int nsets = static_cast<int>(sps.st_rps.size());
        if (!b.flag()) {  // short_term_ref_pic_set_sps_flag == 0: coded here
            std::vector<StRps> sets = sps.st_rps;
            sets.push_back(StRps{});
            st_ref_pic_set(b, nsets, sets, nsets, sets.back());
            check_rps_size(sets.back(), sps);
            sh.st_rps = sets.back();
        } else {
            int idx = 0;
            if (nsets > 1) idx = static_cast<int>(b.u(ceil_log2(static_cast<uint32_t>(nsets))));
            if (idx >= nsets) throw ParseError(-1, "short_term_ref_pic_set_idx");
            sh.st_rps = sps.st_rps[static_cast<size_t>(idx)];
        }
        if (sps.long_term_refs_present) {
            int nsps = 0;
            if (!sps.lt_poc_lsb.empty()) nsps = static_cast<int>(b.ue());
            const int npics = static_cast<int>(b.ue());
            sh.num_lt = nsps + npics;
            sh.num_lt_sps = nsps;
            if (sh.num_lt > 32) throw ParseError(-1, "long-term pictures");
            for (int i = 0; i < sh.num_lt; i++) {
                if (i < nsps) {
                    int k = 0;
                    if (sps.lt_poc_lsb.size() > 1) k = static_cast<int>(b.u(ceil_log2(static_cast<uint32_t>(sps.lt_poc_lsb.size()))));
                    if (k >= static_cast<int>(sps.lt_poc_lsb.size())) throw ParseError(-1, "lt_idx_sps");
                    sh.lt_poc_lsb[i] = sps.lt_poc_lsb[static_cast<size_t>(k)];
                    sh.lt_used[i] = sps.lt_used[static_cast<size_t>(k)];
                } else {
                    sh.lt_poc_lsb[i] = b.u(sps.log2_max_poc_lsb);
                    sh.lt_used[i] = b.flag();
                }
                sh.lt_msb_present[i] = b.flag();
                if (sh.lt_msb_present[i]) sh.lt_msb_cycle[i] = b.ue();
            }
        }
        if (sps.temporal_mvp) sh.temporal_mvp = b.flag();
    }
    if (sps.sao) {
        sh.sao_luma = b.flag();
        if (sps.chroma_format_idc) sh.sao_chroma = b.flag();
    }
    if (sh.slice_type != HEVCR_SLICE_I) {
        sh.num_ref_idx[0] = pps.num_ref_idx_default[0];
        sh.num_ref_idx[1] = sh.slice_type == HEVCR_SLICE_B ? pps.num_ref_idx_default[1] : 0;
        if (b.flag()) {
            sh.num_ref_idx[0] = static_cast<int>(b.ue()) + 1;
            if (sh.slice_type == HEVCR_SLICE_B) sh.num_ref_idx[1] = static_cast<int>(b.ue()) + 1;
        }
        if (sh.num_ref_idx[0] > 15 || sh.num_ref_idx[1] > 15) throw ParseError(-1, "num_ref_idx_active");
        const int total = sh.num_pic_total_curr();
        if (pps.lists_modification_present && total > 1) {
            const int bits = ceil_log2(static_cast<uint32_t>(total));
            for (int l = 0; l < (sh.slice_type == HEVCR_SLICE_B ? 2 : 1); l++) {
                sh.list_mod[l] = b.flag();
                if (sh.list_mod[l])
                    for (int i = 0; i < sh.num_ref_idx[l]; i++) sh.list_entry[l][i] = b.u(bits);
            }
        }
        if (sh.slice_type == HEVCR_SLICE_B) sh.mvd_l1_zero = b.flag();
        if (pps.cabac_init_present) sh.cabac_init = b.flag();
        if (sh.temporal_mvp) {
            if (sh.slice_type == HEVCR_SLICE_B) sh.collocated_from_l0 = b.flag();
            if ((sh.collocated_from_l0 && sh.num_ref_idx[0] > 1) || (!sh.collocated_from_l0 && sh.num_ref_idx[1] > 1))
                sh.collocated_ref_idx = static_cast<int>(b.ue());
        }
        if ((pps.weighted_pred && sh.slice_type == HEVCR_SLICE_P) ||
            (pps.weighted_bipred && sh.slice_type == HEVCR_SLICE_B))
            pred_weight_table(b, sps, sh);
        sh.max_num_merge_cand = 5 - static_cast<int>(b.ue());
        if (sh.max_num_merge_cand < 1) throw ParseError(-1, "five_minus_max_num_merge_cand");
    }
    sh.qp_delta = b.se();
    if (pps.slice_chroma_qp_offsets_present) {
        sh.cb_qp_offset = b.se();
        sh.cr_qp_offset = b.se();
    }
    bool override_flag = false;
    if (pps.deblocking_override_enabled) override_flag = b.flag();
    sh.deblocking_disabled = pps.deblocking_disabled;
    sh.beta_offset_div2 = pps.beta_offset_div2;
    sh.tc_offset_div2 = pps.tc_offset_div2;
    if (override_flag) {
        sh.deblocking_disabled = b.flag();
        if (!sh.deblocking_disabled) {
            sh.beta_offset_div2 = b.se();
            sh.tc_offset_div2 = b.se();
        }
    }
    sh.loop_filter_across_slices = pps.loop_filter_across_slices;
    if (pps.loop_filter_across_slices && (sh.sao_luma || sh.sao_chroma || !sh.deblocking_disabled))
        sh.loop_filter_across_slices = b.flag();
    entry_points_and_end(b, pps, sh);
}

void entry_points_and_end(Bits& b, const Pps& pps, SliceHeader& sh) {
    if (pps.tiles || pps.entropy_sync) {
        const uint32_t n = b.ue();
        if (n > 1000000) throw ParseError(-1, "num_entry_point_offsets");
        if (n) {
            const int len = static_cast<int>(b.ue()) + 1;
            if (len > 32) throw ParseError(-1, "offset_len_minus1");
            for (uint32_t i = 0; i < n; i++) sh.entry_points.push_back(b.u(len) + 1);
        }
    }
    if (pps.slice_header_extension_present) {
        const uint32_t n = b.ue();
        b.skip(8 * static_cast<size_t>(n));
    }
    // byte_alignment(): a one bit, then zeros up to the byte boundary
    if (!b.flag()) throw ParseError(-1, "slice header byte_alignment");
    b.align();
    sh.data_offset = b.pos() / 8;
}

}  // namespace hevcf
