// Host front end: parameter sets and slice headers (H.265 7.3.2-7.3.6,
// 7.4.3-7.4.7). The reference's counterparts are src/syntax/sps.cpp,
// pps.cpp, slice_header.cpp.
#pragma once

#include <array>
#include <cstdint>
#include <memory>
#include <vector>

#include "bits.hpp"
#include "hevcr_syntax.h"

namespace hevcf {

struct StRps {
    int num_negative = 0, num_positive = 0;
    int delta_s0[16] = {}, delta_s1[16] = {};
    bool used_s0[16] = {}, used_s1[16] = {};
    int num_delta() const { return num_negative + num_positive; }
};

struct Sps {
    int id = 0;
    int max_sub_layers = 1;
    int profile_idc = 0, level_idc = 0;
    int chroma_format_idc = 1;
    int width = 0, height = 0;
    int conf[4] = {0, 0, 0, 0};  // left, right, top, bottom in luma samples
    int bit_depth = 8, bit_depth_c = 8;
    int log2_max_poc_lsb = 4;
    int max_dec_pic_buffering = 1, max_num_reorder = 0, max_latency_increase_plus1 = 0;  // highest sub-layer
    int log2_min_cb = 3, log2_ctb = 4, log2_min_tb = 2, log2_max_tb = 5;
    int max_th_depth_inter = 0, max_th_depth_intra = 0;
    bool scaling_list_enabled = false;
    hevcr_scaling scaling{};
    bool amp = false, sao = false, pcm = false;
    int pcm_log2_min = 0, pcm_log2_max = 0;
    int pcm_bits_luma = 8, pcm_bits_chroma = 8;
    bool pcm_loop_filter_disabled = false;
    std::vector<StRps> st_rps;
    bool long_term_refs_present = false;
    std::vector<uint32_t> lt_poc_lsb;
    std::vector<bool> lt_used;
    bool temporal_mvp = false, strong_intra_smoothing = false;
    // derived
    int ctb_size() const { return 1 << log2_ctb; }
    int width_ctbs() const { return (width + ctb_size() - 1) >> log2_ctb; }
    int height_ctbs() const { return (height + ctb_size() - 1) >> log2_ctb; }
};

struct Pps {
    int id = 0, sps_id = 0;
    bool dependent_slices = false, output_flag_present = false;
    int num_extra_slice_header_bits = 0;
    bool sign_data_hiding = false, cabac_init_present = false;
    int num_ref_idx_default[2] = {1, 1};
    int init_qp = 26;
    bool constrained_intra_pred = false, transform_skip = false;
    bool cu_qp_delta = false;
    int diff_cu_qp_delta_depth = 0;
    int cb_qp_offset = 0, cr_qp_offset = 0;
    bool slice_chroma_qp_offsets_present = false;
    bool weighted_pred = false, weighted_bipred = false;
    bool transquant_bypass = false, tiles = false, entropy_sync = false;
    // tiles (7.4.3.3.1): columns / rows, explicit widths and heights in CTBs
    // (minus the last, derived) unless uniform_spacing_flag
    int tile_cols = 1, tile_rows = 1;
    bool tile_uniform = true, loop_filter_across_tiles = true;
    std::vector<int> tile_col_width, tile_row_height;
    bool loop_filter_across_slices = false;
    bool deblocking_override_enabled = false, deblocking_disabled = false;
    int beta_offset_div2 = 0, tc_offset_div2 = 0;
    bool scaling_list_present = false;
    hevcr_scaling scaling{};
    bool lists_modification_present = false;
    int log2_parallel_merge_level = 2;
    bool slice_header_extension_present = false;
};

struct WeightEntry {
    int16_t w[3] = {0, 0, 0};  // luma, Cb, Cr weights
    int16_t o[3] = {0, 0, 0};  // luma offset, derived chroma offsets (8-bit units)
};

struct SliceHeader {
    bool first_slice = false, no_output_of_prior_pics = false, dependent = false;
    int pps_id = 0;
    uint32_t segment_address = 0;
    int slice_type = HEVCR_SLICE_I;
    bool pic_output = true;
    uint32_t poc_lsb = 0;
    StRps st_rps;                 // the RPS in force (from the SPS or coded here)
    // long-term entries of this picture
    int num_lt = 0, num_lt_sps = 0;
    uint32_t lt_poc_lsb[32] = {};
    bool lt_used[32] = {}, lt_msb_present[32] = {};
    uint32_t lt_msb_cycle[32] = {};
    bool temporal_mvp = false;
    bool sao_luma = false, sao_chroma = false;
    int num_ref_idx[2] = {0, 0};
    bool list_mod[2] = {false, false};
    uint32_t list_entry[2][16] = {};
    bool mvd_l1_zero = false, cabac_init = false;
    bool collocated_from_l0 = true;
    int collocated_ref_idx = 0;
    int luma_log2_wd = 0, chroma_log2_wd = 0;
    WeightEntry weights[2][16];
    int max_num_merge_cand = 5;
    int qp_delta = 0, cb_qp_offset = 0, cr_qp_offset = 0;
    bool deblocking_disabled = false;
    int beta_offset_div2 = 0, tc_offset_div2 = 0;
    bool loop_filter_across_slices = false;
    std::vector<uint32_t> entry_points;   // entry_point_offset_minus1 + 1, coded bytes
    size_t data_offset = 0;               // slice data start, in RBSP bytes
    int slice_qp(const Pps& p) const { return p.init_qp + qp_delta; }
    int num_pic_total_curr() const;
};

using SpsTable = std::array<std::shared_ptr<Sps>, 16>;
using PpsTable = std::array<std::shared_ptr<Pps>, 64>;

void parse_sps(Bits& b, Sps& s);
void parse_pps(Bits& b, Pps& p, const SpsTable& sps);
// Parses a slice segment header; `nal_type` selects the IRAP/IDR fields. A
// dependent slice segment takes the fields it does not code from
// `independent`, the picture's last independent segment (7.4.7.1; the
// reference: decoder.cpp:203-211); without one it is an error.
void parse_slice_header(Bits& b, int nal_type, const SpsTable& sps, const PpsTable& pps, SliceHeader& sh,
                        const SliceHeader* independent = nullptr);

// Column widths and row heights of the picture's tiles in CTBs (6.5.1): one
// each without tiles.
void tile_sizes(const Sps& sps, const Pps& pps, std::vector<int>& cols, std::vector<int>& rows);

// Default scaling lists (Tables 7-5, 7-6) in the stored form the engine's
// hevcr_scaling record indexes (raster positions of the 4x4 / 8x8 base matrix).
void default_scaling(hevcr_scaling& s);

}  // namespace hevcf
