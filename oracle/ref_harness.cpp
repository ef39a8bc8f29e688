// Seeded-syntax harness around the reference decoder's own reconstruction
// code (test infrastructure: builds only where the reference tree is present,
// into oracle/_ref/, and is used to write golden vectors and to pin the
// oracle restatement).
//
// It fills the reference's SPS/PPS/SliceHeader/DPB/DecodingContext from a
// hevcr_frame_syntax and calls the reference's exported functions in the
// reference's order:
//   store_pu_motion + perform_inter_prediction per PU   (coding_tree.cpp:882-952)
//   perform_dequant + perform_transform_inverse +
//   perform_intra_prediction per TU                     (coding_tree.cpp:1256-1419)
//   apply_deblocking, apply_sao                         (decoder.cpp:270-271)
// Restated here, because the reference keeps them file-static or inline in
// the parser: reconstruct_block (coding_tree.cpp:371-396), the chroma QP
// mapping (coding_tree.cpp:1366-1377), the PU prediction copy
// (coding_tree.cpp:888-922) and the filter-grid writes done while parsing
// (coding_tree.cpp:956-1054, 1233-1247).
#include <cstdint>
#include <cstring>
#include <memory>
#include <vector>

#include "common/picture.h"
#include "decoding/cabac_tables.h"
#include "decoding/coding_tree.h"
#include "decoding/dpb.h"
#include "decoding/interpolation.h"
#include "filters/deblocking.h"
#include "filters/sao.h"
#include "hevcr_syntax.h"
#include "syntax/pps.h"
#include "syntax/slice_header.h"
#include "syntax/sps.h"

namespace hevc {
void perform_dequant(DecodingContext& ctx, int x0, int y0, int log2TrafoSize, int cIdx, int qp,
                     const int16_t* coefficients, int16_t* scaled);
void perform_transform_inverse(int log2TrafoSize, int cIdx, bool is_intra, bool transform_skip,
                               int bit_depth, const int16_t* scaled, int16_t* residual);
void perform_intra_prediction(DecodingContext& ctx, int x0, int y0, int log2PredSize, int cIdx,
                              int intra_mode, int16_t* pred_samples);
}  // namespace hevc

using namespace hevc;

namespace {

void fill_sps(SPS& sps, const hevcr_frame_params& p) {
    sps.chroma_format_idc = 1;
    sps.pic_width_in_luma_samples = static_cast<uint32_t>(p.width);
    sps.pic_height_in_luma_samples = static_cast<uint32_t>(p.height);
    sps.bit_depth_luma_minus8 = static_cast<uint32_t>(p.bit_depth - 8);
    sps.bit_depth_chroma_minus8 = static_cast<uint32_t>(p.bit_depth - 8);
    sps.log2_max_pic_order_cnt_lsb_minus4 = 4;
    sps.log2_min_luma_coding_block_size_minus3 = static_cast<uint32_t>(p.min_cb_log2 - 3);
    sps.log2_diff_max_min_luma_coding_block_size = static_cast<uint32_t>(p.ctb_log2 - p.min_cb_log2);
    sps.log2_min_luma_transform_block_size_minus2 = 0;
    sps.log2_diff_max_min_luma_transform_block_size = 3;
    sps.max_transform_hierarchy_depth_inter = 4;
    sps.max_transform_hierarchy_depth_intra = 4;
    sps.amp_enabled_flag = true;
    sps.sample_adaptive_offset_enabled_flag = p.sao_enabled != 0;
    sps.strong_intra_smoothing_enabled_flag = p.strong_intra_smoothing != 0;
    sps.intra_smoothing_disabled_flag = p.intra_smoothing_disabled != 0;
    sps.derive();
}

struct Harness {
    const hevcr_frame_syntax& s;
    const hevcr_frame_params& p;
    SPS sps;
    PPS pps;
    SliceHeader sh;                  // the first slice segment's header (lists, RPS)
    std::vector<SliceHeader> shs;    // one per slice segment
    ShortTermRefPicSet rps;
    DPB dpb;
    Picture* pic = nullptr;
    std::vector<Picture*> ref_pics;

    std::vector<CUInfo> cu_info;
    std::vector<int> mode_y, mode_c;
    std::vector<PUMotionInfo> motion;
    std::vector<uint8_t> cbf, log2tu, ev, eh, slice_idx;
    std::vector<DecodingContext::SaoParams> sao;
    std::vector<uint16_t> sao_backup[3];
    std::vector<const SliceHeader*> sh_ptrs;
    DecodingContext ctx;

    explicit Harness(const hevcr_frame_syntax& syn) : s(syn), p(syn.p) {}

    int setup(const uint16_t* const* ref_planes, int num_ref_pics) {
        fill_sps(sps, p);
        if (s.scaling) {
            sps.scaling_list_enabled_flag = true;
            sps.sps_scaling_list_data_present_flag = true;
            for (int z = 0; z < 4; z++)
                for (int m = 0; m < 6; m++)
                    for (int i = 0; i < 64; i++) sps.scaling_list_data.scaling_list[z][m][i] = s.scaling->list[z][m][i];
            for (int z = 0; z < 2; z++)
                for (int m = 0; m < 6; m++) sps.scaling_list_data.scaling_list_dc[z][m] = s.scaling->dc[z][m];
        }
        pps.pps_cb_qp_offset = p.pps_cb_qp_offset;
        pps.pps_cr_qp_offset = p.pps_cr_qp_offset;
        pps.weighted_pred_flag = p.slice_type == HEVCR_SLICE_P && p.weighted_pred;
        pps.weighted_bipred_flag = p.slice_type == HEVCR_SLICE_B && p.weighted_pred;
        if (s.slices) {  // explicit WP in force per segment: the PPS flag of its slice type
            pps.weighted_pred_flag = pps.weighted_bipred_flag = false;
            for (uint32_t k = 0; k < s.num_slices; k++) {
                const hevcr_slice& q = s.slices[k];
                if (q.weighted_pred && q.slice_type == HEVCR_SLICE_P) pps.weighted_pred_flag = true;
                if (q.weighted_pred && q.slice_type == HEVCR_SLICE_B) pps.weighted_bipred_flag = true;
            }
            for (uint32_t k = 0; k < s.num_slices; k++) {
                const hevcr_slice& q = s.slices[k];
                const bool on = q.slice_type == HEVCR_SLICE_P ? pps.weighted_pred_flag
                                                               : (q.slice_type == HEVCR_SLICE_B && pps.weighted_bipred_flag);
                if (on != (q.weighted_pred != 0)) return -6;  // not expressible with one PPS
            }
        }
        if (s.tiles) {
            pps.tiles_enabled_flag = true;
            pps.uniform_spacing_flag = false;
            pps.num_tile_columns_minus1 = static_cast<uint32_t>(s.tiles->cols - 1);
            pps.num_tile_rows_minus1 = static_cast<uint32_t>(s.tiles->rows - 1);
            pps.column_width_minus1.assign(static_cast<size_t>(s.tiles->cols), 0);
            pps.row_height_minus1.assign(static_cast<size_t>(s.tiles->rows), 0);
            for (int i = 0; i < s.tiles->cols; i++) pps.column_width_minus1[static_cast<size_t>(i)] = s.tiles->col_width[i] - 1u;
            for (int i = 0; i < s.tiles->rows; i++) pps.row_height_minus1[static_cast<size_t>(i)] = s.tiles->row_height[i] - 1u;
            pps.loop_filter_across_tiles_enabled_flag = s.tiles->loop_filter_across_tiles != 0;
        }
        pps.transquant_bypass_enabled_flag = true;
        sps.pcm_enabled_flag = true;
        sps.log2_min_pcm_luma_coding_block_size_minus3 = 0;
        sps.log2_diff_max_min_pcm_luma_coding_block_size = 2;
        sps.pcm_loop_filter_disabled_flag = s.pcm_loop_filter_disabled != 0;
        sps.derive();
        pps.derive_tile_scan(sps);

        sh.slice_type = static_cast<SliceType>(p.slice_type);
        sh.slice_cb_qp_offset = p.slice_cb_qp_offset;
        sh.slice_cr_qp_offset = p.slice_cr_qp_offset;
        sh.slice_deblocking_filter_disabled_flag = p.deblocking_disabled != 0;
        sh.slice_beta_offset_div2 = p.beta_offset_div2;
        sh.slice_tc_offset_div2 = p.tc_offset_div2;
        sh.slice_sao_luma_flag = sh.slice_sao_chroma_flag = p.sao_enabled != 0;
        sh.slice_loop_filter_across_slices_enabled_flag = true;
        PredWeightTable& pwt = sh.pred_weight_table;
        pwt.luma_log2_weight_denom = static_cast<uint32_t>(p.luma_log2_wd);
        pwt.delta_chroma_log2_weight_denom = p.chroma_log2_wd - p.luma_log2_wd;
        for (int l = 0; l < 2; l++)
            for (int i = 0; i < HEVCR_MAX_REFS; i++) {
                auto& rw = l ? pwt.l1[i] : pwt.l0[i];
                const hevcr_weight& w = p.weights[l][i];
                rw.luma_weight = w.w[0];
                rw.luma_offset = w.o[0];
                rw.chroma_weight[0] = w.w[1];
                rw.chroma_weight[1] = w.w[2];
                rw.chroma_offset[0] = w.o[1];
                rw.chroma_offset[1] = w.o[2];
            }

        const int nseg = s.slices ? static_cast<int>(s.num_slices) : 1;
        shs.assign(static_cast<size_t>(nseg), sh);
        for (int k = 0; k < nseg && s.slices; k++) {
            const hevcr_slice& q = s.slices[k];
            SliceHeader& t = shs[static_cast<size_t>(k)];
            t.slice_type = static_cast<SliceType>(q.slice_type);
            t.slice_cb_qp_offset = q.slice_cb_qp_offset;
            t.slice_cr_qp_offset = q.slice_cr_qp_offset;
            t.slice_deblocking_filter_disabled_flag = q.deblocking_disabled != 0;
            t.slice_beta_offset_div2 = q.beta_offset_div2;
            t.slice_tc_offset_div2 = q.tc_offset_div2;
            t.slice_loop_filter_across_slices_enabled_flag = q.loop_filter_across_slices != 0;
            PredWeightTable& tw = t.pred_weight_table;
            tw.luma_log2_weight_denom = static_cast<uint32_t>(q.luma_log2_wd);
            tw.delta_chroma_log2_weight_denom = q.chroma_log2_wd - q.luma_log2_wd;
            for (int l = 0; l < 2; l++)
                for (int i = 0; i < HEVCR_MAX_REFS; i++) {
                    auto& rw = l ? tw.l1[i] : tw.l0[i];
                    const hevcr_weight& qw = q.weights[l][i];
                    rw.luma_weight = qw.w[0];
                    rw.luma_offset = qw.o[0];
                    rw.chroma_weight[0] = qw.w[1];
                    rw.chroma_weight[1] = qw.w[2];
                    rw.chroma_offset[0] = qw.o[1];
                    rw.chroma_offset[1] = qw.o[2];
                }
        }

        const int w = p.width, h = p.height, bd = p.bit_depth;
        // Reference pictures: POC of picture k from the list entries pointing at it.
        std::vector<int> ref_poc(static_cast<size_t>(num_ref_pics), INT32_MIN);
        const int nl = p.slice_type == HEVCR_SLICE_B ? 2 : (p.slice_type == HEVCR_SLICE_P ? 1 : 0);
        for (int l = 0; l < nl; l++)
            for (int i = 0; i < p.num_ref[l]; i++) {
                const int k = p.ref_slot[l][i];
                if (k < 0 || k >= num_ref_pics) return -1;
                ref_poc[static_cast<size_t>(k)] = p.ref_poc[l][i];
            }
        for (int k = 0; k < num_ref_pics; k++) {
            Picture* rp = dpb.alloc_picture(w, h, ChromaFormat::YUV420, bd, bd);
            for (int c = 0; c < 3; c++)
                std::memcpy(rp->planes[c].data(), ref_planes[k * 3 + c], rp->planes[c].size() * 2);
            rp->poc = ref_poc[static_cast<size_t>(k)];
            dpb.mark_current_as_short_term_ref();
            ref_pics.push_back(rp);
        }
        pic = dpb.alloc_picture(w, h, ChromaFormat::YUV420, bd, bd);
        pic->poc = p.poc;
        if (nl > 0) {
            // RPS from the first num_refs distinct pictures of list 0, in list order.
            std::vector<int> before, after;
            std::vector<bool> seen(static_cast<size_t>(num_ref_pics), false);
            for (int i = 0; i < p.num_ref[0]; i++) {
                const int k = p.ref_slot[0][i];
                if (seen[static_cast<size_t>(k)]) continue;
                seen[static_cast<size_t>(k)] = true;
                (ref_poc[static_cast<size_t>(k)] < p.poc ? before : after).push_back(ref_poc[static_cast<size_t>(k)]);
            }
            rps.NumNegativePics = static_cast<uint32_t>(before.size());
            rps.NumPositivePics = static_cast<uint32_t>(after.size());
            for (size_t i = 0; i < before.size(); i++) {
                rps.DeltaPocS0[i] = before[i] - p.poc;
                rps.UsedByCurrPicS0[i] = true;
            }
            for (size_t i = 0; i < after.size(); i++) {
                rps.DeltaPocS1[i] = after[i] - p.poc;
                rps.UsedByCurrPicS1[i] = true;
            }
            sh.active_rps = &rps;
            sh.num_ref_idx_l0_active_minus1 = static_cast<uint32_t>(p.num_ref[0] - 1);
            sh.num_ref_idx_l1_active_minus1 = static_cast<uint32_t>(nl > 1 ? p.num_ref[1] - 1 : 0);
            dpb.derive_rps(sh, sps, NalUnitType::TRAIL_R, p.poc);
            dpb.construct_ref_pic_lists(sh, sps, pps);
            // The reference's list construction must land on exactly the syntax's lists.
            for (int l = 0; l < nl; l++) {
                if ((l ? dpb.num_ref_list1() : dpb.num_ref_list0()) != p.num_ref[l]) return -2;
                for (int i = 0; i < p.num_ref[l]; i++) {
                    Picture* rp = l ? dpb.ref_pic_list1(i) : dpb.ref_pic_list0(i);
                    if (rp != ref_pics[static_cast<size_t>(p.ref_slot[l][i])]) return -3;
                }
            }
        }

        // Grids as allocated per picture (decoder.cpp:117-157).
        const int gw = w / 4, gh = h / 4;
        cu_info.assign(static_cast<size_t>(sps.PicWidthInMinCbsY) * sps.PicHeightInMinCbsY, CUInfo{});
        mode_y.assign(static_cast<size_t>(gw) * gh, 1);
        mode_c.assign(static_cast<size_t>(gw) * gh, 0);
        motion.assign(static_cast<size_t>(gw) * gh, PUMotionInfo{});
        cbf.assign(static_cast<size_t>(gw) * gh, 0);
        log2tu.assign(static_cast<size_t>(gw) * gh, static_cast<uint8_t>(sps.CtbLog2SizeY));
        ev.assign(static_cast<size_t>(gw) * gh, 0);
        eh.assign(static_cast<size_t>(gw) * gh, 0);
        slice_idx.assign(static_cast<size_t>(sps.PicSizeInCtbsY), 0);
        sao.assign(static_cast<size_t>(sps.PicSizeInCtbsY), DecodingContext::SaoParams{});
        for (int i = 0; i < sps.PicSizeInCtbsY; i++)
            for (int c = 0; c < 3; c++) {
                const hevcr_sao& r = s.sao[i];
                sao[static_cast<size_t>(i)].sao_type_idx[c] = r.type[c];
                sao[static_cast<size_t>(i)].sao_band_position[c] = r.band_pos[c];
                sao[static_cast<size_t>(i)].sao_eo_class[c] = r.eo_class[c];
                for (int k = 0; k < 5; k++) sao[static_cast<size_t>(i)].sao_offset_val[c][k] = r.offset[c][k];
            }

        ctx.sps = &sps;
        ctx.pps = &pps;
        ctx.sh = &sh;
        ctx.pic = pic;
        ctx.dpb = &dpb;
        ctx.cu_info = cu_info.data();
        ctx.cu_info_stride = sps.PicWidthInMinCbsY;
        ctx.intra_pred_mode_y = mode_y.data();
        ctx.intra_pred_mode_c = mode_c.data();
        ctx.intra_pred_mode_stride = gw;
        ctx.motion_info = motion.data();
        ctx.motion_info_stride = gw;
        ctx.cbf_luma_grid = cbf.data();
        ctx.log2_tu_size_grid = log2tu.data();
        ctx.edge_flags_v = ev.data();
        ctx.edge_flags_h = eh.data();
        ctx.filter_grid_stride = gw;
        ctx.sao_params = sao.data();
        ctx.sao_params_stride = sps.PicWidthInCtbsY;
        ctx.sao_backup = sao_backup;
        if (s.slices)
            for (int i = 0; i < sps.PicSizeInCtbsY; i++) slice_idx[static_cast<size_t>(i)] = s.ctb_slice[i];
        ctx.slice_idx = slice_idx.data();
        sh_ptrs.resize(shs.size());
        for (size_t k = 0; k < shs.size(); k++) sh_ptrs[k] = &shs[k];
        ctx.slice_headers = sh_ptrs.data();
        ctx.num_slices = static_cast<int>(shs.size());
        return 0;
    }

    // reconstruct_block, coding_tree.cpp:371-396
    void recon_block(int x0, int y0, int log2, int c, const int16_t* pred, const int16_t* res) {
        const int n = 1 << log2, maxv = (1 << p.bit_depth) - 1;
        const int xc = c ? x0 / 2 : x0, yc = c ? y0 / 2 : y0;
        const int pw = c ? p.width / 2 : p.width, ph = c ? p.height / 2 : p.height;
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++) {
                if (xc + i >= pw || yc + j >= ph) continue;
                int v = pred[j * n + i] + res[j * n + i];
                v = v < 0 ? 0 : (v > maxv ? maxv : v);
                pic->sample(c, xc + i, yc + j) = static_cast<uint16_t>(v);
            }
    }

    void cu_grids(const hevcr_cu& cu) {
        const int x0 = cu.x, y0 = cu.y, cb = 1 << cu.log2_size, st = ctx.filter_grid_stride;
        PredMode pm = static_cast<PredMode>(cu.pred_mode);
        for (int j = 0; j < cb / 8; j++)
            for (int i = 0; i < cb / 8; i++) {
                CUInfo& ci = ctx.cu_at(x0 + i * 8, y0 + j * 8);
                ci.pred_mode = pm;
                ci.part_mode = static_cast<PartMode>(cu.part_mode);
                ci.log2CbSize = cu.log2_size;
                ci.cu_transquant_bypass = (cu.flags & HEVCR_CU_BYPASS) != 0;
            }
        for (int dy = 0; dy < cb; dy += 4)
            for (int dx = 0; dx < cb; dx += 4) {
                log2tu[static_cast<size_t>((y0 + dy) / 4 * st + (x0 + dx) / 4)] = cu.log2_size;
                cbf[static_cast<size_t>((y0 + dy) / 4 * st + (x0 + dx) / 4)] = 0;
            }
        auto vline = [&](int px) {
            for (int dy = 0; dy < cb; dy += 4) ev[static_cast<size_t>((y0 + dy) / 4 * st + px / 4)] = 1;
        };
        auto hline = [&](int py) {
            for (int dx = 0; dx < cb; dx += 4) eh[static_cast<size_t>(py / 4 * st + (x0 + dx) / 4)] = 1;
        };
        vline(x0);
        hline(y0);
        const int pm2 = cu.part_mode;
        if (pm2 == HEVCR_PART_Nx2N || pm2 == HEVCR_PART_NxN) vline(x0 + cb / 2);
        if (pm2 == HEVCR_PART_nLx2N) vline(x0 + cb / 4);
        if (pm2 == HEVCR_PART_nRx2N) vline(x0 + 3 * cb / 4);
        if (pm2 == HEVCR_PART_2NxN || pm2 == HEVCR_PART_NxN) hline(y0 + cb / 2);
        if (pm2 == HEVCR_PART_2NxnU) hline(y0 + cb / 4);
        if (pm2 == HEVCR_PART_2NxnD) hline(y0 + 3 * cb / 4);
    }

    void mc_pu(const hevcr_pu& pu) {
        PUMotionInfo mi;
        for (int l = 0; l < 2; l++) {
            mi.mv[l].x = pu.mv[l][0];
            mi.mv[l].y = pu.mv[l][1];
            mi.ref_idx[l] = pu.ref_idx[l];
            mi.pred_flag[l] = (pu.pred_flags >> l) & 1;
        }
        store_pu_motion(ctx, pu.x, pu.y, pu.w, pu.h, mi);
        static thread_local int16_t pred[64 * 64];
        for (int c = 0; c < 3; c++) {
            perform_inter_prediction(ctx, pu.x, pu.y, pu.w, pu.h, c, mi.mv[0], mi.mv[1], mi.ref_idx[0],
                                     mi.ref_idx[1], mi.pred_flag[0], mi.pred_flag[1], pred);
            const int w = c ? pu.w / 2 : pu.w, h = c ? pu.h / 2 : pu.h;
            const int xc = c ? pu.x / 2 : pu.x, yc = c ? pu.y / 2 : pu.y;
            for (int y = 0; y < h; y++)
                for (int x = 0; x < w; x++) pic->sample(c, xc + x, yc + y) = static_cast<uint16_t>(pred[y * w + x]);
        }
    }

    void tu(const hevcr_cu& cu, const hevcr_tu& t) {
        const bool intra = cu.pred_mode == HEVCR_MODE_INTRA;
        const int qpY = cu.qp_y;
        static thread_local int16_t scaled[64 * 64], res[64 * 64], pred[64 * 64], zero[64 * 64];
        const int n = 1 << t.log2_size;
        // Luma
        const bool bypass = (cu.flags & HEVCR_CU_BYPASS) != 0;  // residual = levels (coding_tree.cpp:1292-1305)
        if (t.cbf & HEVCR_CBF_Y) {
            if (bypass) {
                std::memcpy(res, s.coeffs + t.coeff_off[0], sizeof(int16_t) * n * n);
            } else {
                perform_dequant(ctx, t.x, t.y, t.log2_size, 0, qpY + sps.QpBdOffsetY, s.coeffs + t.coeff_off[0], scaled);
                perform_transform_inverse(t.log2_size, 0, intra, (t.tskip & 1) != 0, sps.BitDepthY, scaled, res);
            }
            if (intra) {
                perform_intra_prediction(ctx, t.x, t.y, t.log2_size, 0, ctx.intra_mode_at(t.x, t.y), pred);
            } else {
                for (int y = 0; y < n; y++)
                    for (int x = 0; x < n; x++) pred[y * n + x] = static_cast<int16_t>(pic->sample(0, t.x + x, t.y + y));
            }
            recon_block(t.x, t.y, t.log2_size, 0, pred, res);
        } else if (intra) {
            perform_intra_prediction(ctx, t.x, t.y, t.log2_size, 0, ctx.intra_mode_at(t.x, t.y), pred);
            std::memset(zero, 0, sizeof(zero));
            recon_block(t.x, t.y, t.log2_size, 0, pred, zero);
        }
        // Chroma
        const bool chroma_here = t.log2_size > 2 || t.blk_idx == 3;
        if (chroma_here) {
            const int log2c = t.log2_size > 2 ? t.log2_size - 1 : 2, nc = 1 << log2c;
            const int xC = t.log2_size == 2 ? t.xbase : t.x, yC = t.log2_size == 2 ? t.ybase : t.y;
            for (int c = 1; c <= 2; c++) {
                const bool coded = (t.cbf >> c) & 1;
                if (coded) {
                    const int off = c == 1 ? p.pps_cb_qp_offset + ctx.sh->slice_cb_qp_offset
                                           : p.pps_cr_qp_offset + ctx.sh->slice_cr_qp_offset;
                    int qPi = qpY + off;
                    if (qPi < -sps.QpBdOffsetC) qPi = -sps.QpBdOffsetC;
                    if (qPi > 57) qPi = 57;
                    const int qPc = qPi < 0 ? qPi : (qPi < 58 ? qpChromaTable[qPi] : qPi - 6);
                    if (bypass) {
                        std::memcpy(res, s.coeffs + t.coeff_off[c], sizeof(int16_t) * nc * nc);
                    } else {
                        perform_dequant(ctx, xC, yC, log2c, c, qPc + sps.QpBdOffsetC, s.coeffs + t.coeff_off[c], scaled);
                        perform_transform_inverse(log2c, c, intra, ((t.tskip >> c) & 1) != 0, sps.BitDepthC, scaled, res);
                    }
                    if (intra) {
                        perform_intra_prediction(ctx, xC, yC, log2c, c, ctx.chroma_mode_at(xC, yC), pred);
                    } else {
                        for (int y = 0; y < nc; y++)
                            for (int x = 0; x < nc; x++)
                                pred[y * nc + x] = static_cast<int16_t>(pic->sample(c, xC / 2 + x, yC / 2 + y));
                    }
                    recon_block(xC, yC, log2c, c, pred, res);
                } else if (intra) {
                    perform_intra_prediction(ctx, xC, yC, log2c, c, ctx.chroma_mode_at(xC, yC), pred);
                    std::memset(zero, 0, sizeof(zero));
                    recon_block(xC, yC, log2c, c, pred, zero);
                }
            }
        }
        // TU grid writes (coding_tree.cpp:1233-1247)
        const int st = ctx.filter_grid_stride;
        for (int dy = 0; dy < n; dy += 4)
            for (int dx = 0; dx < n; dx += 4) {
                cbf[static_cast<size_t>((t.y + dy) / 4 * st + (t.x + dx) / 4)] = (t.cbf & HEVCR_CBF_Y) ? 1 : 0;
                log2tu[static_cast<size_t>((t.y + dy) / 4 * st + (t.x + dx) / 4)] = t.log2_size;
            }
        for (int dy = 0; dy < n; dy += 4) ev[static_cast<size_t>((t.y + dy) / 4 * st + t.x / 4)] = 1;
        for (int dx = 0; dx < n; dx += 4) eh[static_cast<size_t>(t.y / 4 * st + (t.x + dx) / 4)] = 1;
    }

    int run(uint16_t* const* out_recon, uint16_t* const* out_dbk, uint16_t* const* out_sao) {
        uint32_t pu_i = 0, tu_i = 0;
        for (uint32_t ci = 0; ci < s.num_cus; ci++) {
            const hevcr_cu& cu = s.cus[ci];
            const int ctb_addr = (cu.y >> sps.CtbLog2SizeY) * sps.PicWidthInCtbsY + (cu.x >> sps.CtbLog2SizeY);
            const int seg = s.slices ? s.ctb_slice[ctb_addr] : 0;
            ctx.sh = &shs[static_cast<size_t>(seg)];
            ctx.current_slice_idx = seg;
            if (cu.flags & HEVCR_CU_PCM) {  // coding_tree.cpp:852-875, 1425-1460: samples, CU info, no grids
                if (tu_i >= s.num_tus || s.tus[tu_i].cu_idx != ci) return -7;
                const hevcr_tu& t = s.tus[tu_i++];
                const int cb = 1 << cu.log2_size;
                for (int c = 0; c < 3; c++) {
                    const int n = c ? cb / 2 : cb, xc = c ? cu.x / 2 : cu.x, yc = c ? cu.y / 2 : cu.y;
                    for (int i = 0; i < n * n; i++)
                        pic->sample(c, xc + i % n, yc + i / n) = static_cast<uint16_t>(s.coeffs[t.coeff_off[c] + i]);
                }
                for (int j = 0; j < cb / 8; j++)
                    for (int i = 0; i < cb / 8; i++) {
                        CUInfo& q = ctx.cu_at(cu.x + i * 8, cu.y + j * 8);
                        q.pred_mode = PredMode::MODE_INTRA;
                        q.log2CbSize = cu.log2_size;
                        q.is_pcm = true;
                        q.qp_y = cu.qp_y;
                    }
                continue;
            }
            cu_grids(cu);
            if (cu.pred_mode == HEVCR_MODE_INTRA) {
                const int cb = 1 << cu.log2_size;
                if (cu.part_mode == HEVCR_PART_NxN) {
                    const int h = cb / 2;
                    for (int k = 0; k < 4; k++) ctx.set_intra_mode(cu.x + (k & 1) * h, cu.y + (k >> 1) * h, h, cu.intra_luma[k]);
                } else {
                    ctx.set_intra_mode(cu.x, cu.y, cb, cu.intra_luma[0]);
                }
                ctx.set_chroma_mode(cu.x, cu.y, cb, cu.intra_chroma);
            } else {
                for (int k = 0; k < cu.num_pu; k++) {
                    if (pu_i >= s.num_pus) return -4;
                    mc_pu(s.pus[pu_i++]);
                }
            }
            while (tu_i < s.num_tus && s.tus[tu_i].cu_idx == ci) tu(cu, s.tus[tu_i++]);
            const int cb = 1 << cu.log2_size;
            for (int j = 0; j < cb / 8; j++)
                for (int i = 0; i < cb / 8; i++) ctx.cu_at(cu.x + i * 8, cu.y + j * 8).qp_y = cu.qp_y;
        }
        if (pu_i != s.num_pus || tu_i != s.num_tus) return -5;
        copy_out(out_recon);
        apply_deblocking(ctx);
        copy_out(out_dbk);
        apply_sao(ctx);
        copy_out(out_sao);
        return 0;
    }

    void copy_out(uint16_t* const* out) {
        if (!out) return;
        for (int c = 0; c < 3; c++)
            if (out[c]) std::memcpy(out[c], pic->planes[c].data(), pic->planes[c].size() * 2);
    }
};

}  // namespace

extern "C" {

// Reconstructs one frame with the reference's code. ref_planes holds 3 planes per
// reference picture (stride == width); out_* are 3 plane buffers or NULL.
int ref_recon_frame(const hevcr_frame_syntax* syn, const uint16_t* const* ref_planes, int num_ref_pics,
                    uint16_t* const* out_recon, uint16_t* const* out_dbk, uint16_t* const* out_sao) {
    if (!syn || syn->sig) return -1;  // dense coefficients only
    auto h = std::make_unique<Harness>(*syn);
    int rc = h->setup(ref_planes, num_ref_pics);
    if (rc) return rc;
    return h->run(out_recon, out_dbk, out_sao);
}

// Dequantisation + inverse transform of one block through perform_dequant /
// perform_transform_inverse (flat scaling list).
int ref_residual_block(int log2, int cidx, int is_intra, int qp_prime, int bit_depth, int tskip,
                       const int16_t* coeffs, int16_t* out) {
    if (log2 < 2 || log2 > 5 || !coeffs || !out) return -1;
    SPS sps;
    hevcr_frame_params p{};
    p.width = 64;
    p.height = 64;
    p.bit_depth = bit_depth;
    p.ctb_log2 = 6;
    p.min_cb_log2 = 3;
    fill_sps(sps, p);
    DecodingContext ctx;
    ctx.sps = &sps;
    int16_t scaled[32 * 32];
    perform_dequant(ctx, 0, 0, log2, cidx, qp_prime, coeffs, scaled);
    perform_transform_inverse(log2, cidx, is_intra != 0, tskip != 0, bit_depth, scaled, out);
    return 0;
}

}  // extern "C"
