// Real-stream tap around the reference decoder (test infrastructure: builds
// only where the reference tree is present, into oracle/_ref/, and is used to
// write the stream golden vectors of tests/golden/).
//
// The reference decoder (src/decoding/decoder.cpp) parses a bitstream and, per
// picture, calls its reconstruction operators from the parser:
//   perform_inter_prediction per PU and component     (coding_tree.cpp:791-912)
//   perform_dequant + perform_transform_inverse per
//   coded transform block                             (coding_tree.cpp:1294-1381)
//   apply_deblocking, apply_sao once per picture      (decoder.cpp:270-271)
// This program links the reference objects with those calls wrapped
// (GNU ld --wrap on their mangled names) and, without changing what the
// reference computes, records per picture the hevcr_frame_syntax that
// describes the same reconstruction:
//   CU records  <- the CUInfo grid and the intra mode grids as the parser left
//                  them (coding_tree.cpp:956-976, 1060-1071, 1124-1135),
//                  CTBs in raster order, CUs in z-order;
//   PU records  <- the motion passed to perform_inter_prediction (luma calls);
//   TU records  <- the transform-tree leaves, read back from
//                  log2_tu_size_grid / cbf_luma_grid (coding_tree.cpp:1230-1239),
//                  with chroma cbf, transform skip and the coefficient blocks
//                  taken from the perform_dequant / perform_transform_inverse
//                  calls;
//   SAO records <- DecodingContext::sao_params;
//   scaling     <- the active ScalingListData (PPS's, else SPS's);
//   parameters  <- SPS / PPS / slice header / DPB reference lists.
// It also writes the picture planes as the reference left them before
// deblocking, after deblocking and after SAO.
//
// Usage: hevc-tap input.265 output.tap
//
// Output: for each picture in decoding order a TapHeader, then (when
// status == 0) hevcr_frame_params (ref_slot[l][i] holds the decoding index of
// the referenced picture), the CU, PU, TU, coefficient and SAO arrays and,
// when scaling lists are on, one hevcr_scaling record, then
// three stages x three planes of uint16 samples (stride == plane width).
// A picture the engine's syntax cannot describe has status != 0, a reason,
// no arrays, and still its planes.
#include <cstdint>
#include <climits>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "common/picture.h"
#include "decoding/cabac_tables.h"
#include "decoding/coding_tree.h"
#include "decoding/decoder.h"
#include "decoding/dpb.h"
#include "hevcr_syntax.h"
#include "syntax/pps.h"
#include "syntax/slice_header.h"
#include "syntax/sps.h"

using namespace hevc;

#define W_DEQUANT _ZN4hevc15perform_dequantERNS_15DecodingContextEiiiiiPKsPs
#define W_TRANSFORM _ZN4hevc25perform_transform_inverseEiibbiPKsPs
#define W_INTER _ZN4hevc24perform_inter_predictionERNS_15DecodingContextEiiiiiRKNS_2MVES4_iibbPs
#define W_DEBLOCK _ZN4hevc16apply_deblockingERNS_15DecodingContextE
#define W_SAO _ZN4hevc9apply_saoERNS_15DecodingContextE
#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define REAL(s) CAT(__real_, s)
#define WRAP(s) CAT(__wrap_, s)

extern "C" {
void REAL(W_DEQUANT)(DecodingContext&, int, int, int, int, int, const int16_t*, int16_t*);
void REAL(W_TRANSFORM)(int, int, bool, bool, int, const int16_t*, int16_t*);
void REAL(W_INTER)(DecodingContext&, int, int, int, int, int, const MV&, const MV&, int, int, bool, bool, int16_t*);
void REAL(W_DEBLOCK)(DecodingContext&);
void REAL(W_SAO)(DecodingContext&);
}

namespace {

struct TapHeader {
    char magic[4];
    int32_t pic_id;
    int32_t cvs_id;
    int32_t status;
    int32_t width, height, bit_depth;
    int32_t crop[4];  // conformance window in luma samples: left, right, top, bottom
    uint32_t num_cus, num_pus, num_tus, num_sao;
    uint32_t has_scaling;  // a hevcr_scaling record follows the SAO records
    uint64_t num_coeffs;
    char reason[64];
};
static_assert(sizeof(TapHeader) == 136, "tap header layout");

struct CoeffCall {
    int x, y, log2, cidx, qp;
    int tskip = 0;
    bool used = false;
    std::vector<int16_t> c;
};

struct InterCall {
    int x, y, w, h;
    MV mv[2];
    int ref_idx[2];
    bool pf[2];
};

// The reference decodes CTB rows on worker threads when a stream uses
// wavefront parallel processing, so the calls are recorded under a lock and
// matched to the syntax by position, never by arrival order.
struct Tap {
    FILE* out = nullptr;
    std::mutex mu;
    std::vector<CoeffCall> coeffs;
    std::vector<InterCall> inter;
    std::unordered_map<const Picture*, int> pic_ids;
    int next_id = 0;
    std::vector<uint16_t> stage[3][3];
    int cur_id = -1;
    std::string unsupported;
};

Tap g_tap;

void snapshot(const Picture& pic, int st) {
    for (int c = 0; c < 3; c++) {
        auto& v = g_tap.stage[st][c];
        v.resize(static_cast<size_t>(pic.width[c]) * pic.height[c]);
        for (int y = 0; y < pic.height[c]; y++)
            std::memcpy(v.data() + static_cast<size_t>(y) * pic.width[c],
                        pic.planes[c].data() + static_cast<size_t>(y) * pic.stride[c],
                        static_cast<size_t>(pic.width[c]) * 2);
    }
}

int chroma_qp_prime(const SPS& sps, const PPS& pps, const SliceHeader& sh, int qpY, int c) {
    // coding_tree.cpp:1366-1377
    const int off = c == 1 ? pps.pps_cb_qp_offset + sh.slice_cb_qp_offset : pps.pps_cr_qp_offset + sh.slice_cr_qp_offset;
    int qPi = qpY + off;
    if (qPi < -sps.QpBdOffsetC) qPi = -sps.QpBdOffsetC;
    if (qPi > 57) qPi = 57;
    const int qPc = qPi < 0 ? qPi : (qPi < 58 ? qpChromaTable[qPi] : qPi - 6);
    return qPc + sps.QpBdOffsetC;
}

struct Builder {
    DecodingContext& ctx;
    const SPS& sps;
    const PPS& pps;
    const SliceHeader& sh;
    std::vector<hevcr_cu> cus;
    std::vector<hevcr_pu> pus;
    std::vector<hevcr_tu> tus;
    std::vector<int16_t> pool;
    std::vector<hevcr_sao> sao;
    hevcr_frame_params p{};
    hevcr_scaling scaling{};
    bool has_scaling = false;
    std::string err;
    std::map<std::tuple<int, int, int, int>, size_t> coeff_at;  // (cidx, x, y, log2) -> call
    std::map<std::tuple<int, int, int, int>, size_t> inter_at;  // (x, y, w, h) -> luma call
    size_t inter_used = 0;

    Builder(DecodingContext& c) : ctx(c), sps(*c.sps), pps(*c.pps), sh(*c.sh) {}

    int log2tu(int x, int y) const { return ctx.log2_tu_size_grid[(y / 4) * ctx.filter_grid_stride + x / 4]; }
    int cbf_y(int x, int y) const { return ctx.cbf_luma_grid[(y / 4) * ctx.filter_grid_stride + x / 4]; }

    bool fail(const char* why) {
        if (err.empty()) err = why;
        return false;
    }

    // Appends the coefficient block of a recorded call; returns its pool offset.
    bool take(int cidx, int x, int y, int log2, int qp_expect, uint32_t& off, int& tskip) {
        auto it = coeff_at.find(std::make_tuple(cidx, x, y, log2));
        if (it == coeff_at.end()) return fail("coded block without a perform_dequant call");
        CoeffCall& k = g_tap.coeffs[it->second];
        if (k.used) return fail("perform_dequant call used twice");
        if (k.qp != qp_expect) return fail("dequantisation QP differs from the CU's QpY");
        k.used = true;
        off = static_cast<uint32_t>(pool.size());
        pool.insert(pool.end(), k.c.begin(), k.c.end());
        tskip = k.tskip;
        return true;
    }

    bool leaf(uint32_t ci, const hevcr_cu& cu, int x, int y, int L, bool root) {
        hevcr_tu t{};
        t.x = static_cast<uint16_t>(x);
        t.y = static_cast<uint16_t>(y);
        t.log2_size = static_cast<uint8_t>(L);
        if (root) {
            t.xbase = t.x;
            t.ybase = t.y;
            t.blk_idx = 0;
        } else {
            t.xbase = static_cast<uint16_t>(x & ~((2 << L) - 1));
            t.ybase = static_cast<uint16_t>(y & ~((2 << L) - 1));
            t.blk_idx = static_cast<uint8_t>(((x >> L) & 1) | (((y >> L) & 1) << 1));
        }
        t.cu_idx = ci;
        const int qpY = cu.qp_y;
        if (cbf_y(x, y)) {
            int ts = 0;
            if (!take(0, x, y, L, qpY + sps.QpBdOffsetY, t.coeff_off[0], ts)) return false;
            t.cbf |= HEVCR_CBF_Y;
            t.tskip |= static_cast<uint8_t>(ts);
        }
        if (L > 2 || t.blk_idx == 3) {
            const int xc = L > 2 ? x : t.xbase, yc = L > 2 ? y : t.ybase, lc = L > 2 ? L - 1 : 2;
            for (int c = 1; c <= 2; c++) {
                if (!coeff_at.count(std::make_tuple(c, xc, yc, lc))) continue;
                int ts = 0;
                if (!take(c, xc, yc, lc, chroma_qp_prime(sps, pps, sh, qpY, c), t.coeff_off[c], ts)) return false;
                t.cbf |= static_cast<uint8_t>(c == 1 ? HEVCR_CBF_CB : HEVCR_CBF_CR);
                t.tskip |= static_cast<uint8_t>(ts << c);
            }
        }
        tus.push_back(t);
        return true;
    }

    // Transform-tree leaves of a CU in z-order, from the TU size grid.
    bool tree(uint32_t ci, const hevcr_cu& cu, int x, int y, int L, bool root) {
        const int g = log2tu(x, y);
        if (g == L) return leaf(ci, cu, x, y, L, root);
        if (g > L || L == 2) return fail("transform tree inconsistent with the TU size grid");
        const int h = 1 << (L - 1);
        for (int k = 0; k < 4; k++)
            if (!tree(ci, cu, x + (k & 1) * h, y + (k >> 1) * h, L - 1, false)) return false;
        return true;
    }

    bool coding_unit(int x0, int y0, int log2) {
        const CUInfo& ci = ctx.cu_at(x0, y0);
        if (ci.is_pcm) return fail("PCM coding unit");
        if (ci.cu_transquant_bypass) return fail("transquant bypass");
        hevcr_cu cu{};
        cu.x = static_cast<uint16_t>(x0);
        cu.y = static_cast<uint16_t>(y0);
        cu.log2_size = static_cast<uint8_t>(log2);
        cu.pred_mode = static_cast<uint8_t>(ci.pred_mode);
        cu.part_mode = static_cast<uint8_t>(ci.part_mode);
        cu.qp_y = static_cast<int8_t>(ci.qp_y);
        const int cb = 1 << log2;
        if (ci.pred_mode == PredMode::MODE_INTRA) {
            if (ci.part_mode == PartMode::PART_NxN) {
                for (int k = 0; k < 4; k++)
                    cu.intra_luma[k] = static_cast<uint8_t>(ctx.intra_mode_at(x0 + (k & 1) * cb / 2, y0 + (k >> 1) * cb / 2));
            } else {
                cu.intra_luma[0] = static_cast<uint8_t>(ctx.intra_mode_at(x0, y0));
            }
            cu.intra_chroma = static_cast<uint8_t>(ctx.chroma_mode_at(x0, y0));
        } else {
            // PU geometry of the partition mode, in partIdx order (coding_tree.cpp:924-951)
            int geo[4][4], n = 0;
            auto add = [&](int x, int y, int w, int h) {
                geo[n][0] = x;
                geo[n][1] = y;
                geo[n][2] = w;
                geo[n][3] = h;
                n++;
            };
            const int q = cb / 4, hf = cb / 2;
            switch (ci.pred_mode == PredMode::MODE_SKIP ? PartMode::PART_2Nx2N : ci.part_mode) {
                case PartMode::PART_2Nx2N: add(x0, y0, cb, cb); break;
                case PartMode::PART_2NxN: add(x0, y0, cb, hf); add(x0, y0 + hf, cb, hf); break;
                case PartMode::PART_Nx2N: add(x0, y0, hf, cb); add(x0 + hf, y0, hf, cb); break;
                case PartMode::PART_2NxnU: add(x0, y0, cb, q); add(x0, y0 + q, cb, cb - q); break;
                case PartMode::PART_2NxnD: add(x0, y0, cb, cb - q); add(x0, y0 + cb - q, cb, q); break;
                case PartMode::PART_nLx2N: add(x0, y0, q, cb); add(x0 + q, y0, cb - q, cb); break;
                case PartMode::PART_nRx2N: add(x0, y0, cb - q, cb); add(x0 + cb - q, y0, q, cb); break;
                case PartMode::PART_NxN:
                    for (int k = 0; k < 4; k++) add(x0 + (k & 1) * hf, y0 + (k >> 1) * hf, hf, hf);
                    break;
            }
            for (int k = 0; k < n; k++) {
                auto it = inter_at.find(std::make_tuple(geo[k][0], geo[k][1], geo[k][2], geo[k][3]));
                if (it == inter_at.end()) return fail("prediction unit without a prediction call");
                const InterCall& m = g_tap.inter[it->second];
                hevcr_pu pu{};
                pu.x = static_cast<uint16_t>(m.x);
                pu.y = static_cast<uint16_t>(m.y);
                pu.w = static_cast<uint16_t>(m.w);
                pu.h = static_cast<uint16_t>(m.h);
                for (int l = 0; l < 2; l++) {
                    pu.mv[l][0] = m.mv[l].x;
                    pu.mv[l][1] = m.mv[l].y;
                    pu.ref_idx[l] = static_cast<int8_t>(m.ref_idx[l]);
                    if (m.pf[l]) pu.pred_flags |= static_cast<uint8_t>(1u << l);
                }
                pus.push_back(pu);
                inter_used++;
            }
            cu.num_pu = static_cast<uint8_t>(n);
        }
        const uint32_t idx = static_cast<uint32_t>(cus.size());
        cus.push_back(cu);
        if (ci.pred_mode == PredMode::MODE_SKIP) return true;
        bool has_tree = ci.pred_mode == PredMode::MODE_INTRA;
        if (!has_tree) {  // rqt_root_cbf: some leaf smaller than the CU, or a coded block inside it
            for (int y = y0; y < y0 + cb && !has_tree; y += 4)
                for (int x = x0; x < x0 + cb && !has_tree; x += 4) has_tree = log2tu(x, y) < log2;
            for (const CoeffCall& k : g_tap.coeffs)
                if (k.x >= x0 && k.y >= y0 && k.x < x0 + cb && k.y < y0 + cb) has_tree = true;
        }
        return !has_tree || tree(idx, cu, x0, y0, log2, true);
    }

    bool quadtree(int x0, int y0, int log2) {
        if (x0 >= static_cast<int>(sps.pic_width_in_luma_samples) ||
            y0 >= static_cast<int>(sps.pic_height_in_luma_samples))
            return true;
        if (ctx.cu_at(x0, y0).log2CbSize == log2) return coding_unit(x0, y0, log2);
        if (log2 <= sps.MinCbLog2SizeY) return fail("CU grid inconsistent");
        const int h = 1 << (log2 - 1);
        for (int k = 0; k < 4; k++)
            if (!quadtree(x0 + (k & 1) * h, y0 + (k >> 1) * h, log2 - 1)) return false;
        return true;
    }

    bool params() {
        const Picture& pic = *ctx.pic;
        if (sps.chroma_format_idc != 1) return fail("chroma format is not 4:2:0");
        if (sps.BitDepthY != sps.BitDepthC || sps.BitDepthY > 10) return fail("bit depth");
        if (pps.tiles_enabled_flag) return fail("tiles");
        if (ctx.num_slices != 1) return fail("more than one slice");
        if (pps.transquant_bypass_enabled_flag) return fail("transquant bypass enabled");
        if (sps.scaling_list_enabled_flag) {  // transform.cpp:308, 339-340
            const ScalingListData& sl =
                pps.pps_scaling_list_data_present_flag ? pps.scaling_list_data : sps.scaling_list_data;
            for (int z = 0; z < 4; z++)
                for (int m = 0; m < 6; m++)
                    for (int i = 0; i < 64; i++) scaling.list[z][m][i] = sl.scaling_list[z][m][i];
            for (int z = 0; z < 2; z++)
                for (int m = 0; m < 6; m++) scaling.dc[z][m] = sl.scaling_list_dc[z][m];
            has_scaling = true;
        }
        p.width = static_cast<int32_t>(sps.pic_width_in_luma_samples);
        p.height = static_cast<int32_t>(sps.pic_height_in_luma_samples);
        if (p.width % 8 || p.height % 8) return fail("picture size not a multiple of 8");
        p.bit_depth = sps.BitDepthY;
        p.ctb_log2 = sps.CtbLog2SizeY;
        p.min_cb_log2 = sps.MinCbLog2SizeY;
        p.slice_type = static_cast<int32_t>(sh.slice_type);
        p.poc = pic.poc;
        p.strong_intra_smoothing = sps.strong_intra_smoothing_enabled_flag;
        p.intra_smoothing_disabled = sps.intra_smoothing_disabled_flag;
        p.sao_enabled = sps.sample_adaptive_offset_enabled_flag;
        p.pps_cb_qp_offset = pps.pps_cb_qp_offset;
        p.pps_cr_qp_offset = pps.pps_cr_qp_offset;
        p.slice_cb_qp_offset = sh.slice_cb_qp_offset;
        p.slice_cr_qp_offset = sh.slice_cr_qp_offset;
        p.deblocking_disabled = sh.slice_deblocking_filter_disabled_flag;
        p.beta_offset_div2 = sh.slice_beta_offset_div2;
        p.tc_offset_div2 = sh.slice_tc_offset_div2;
        p.weighted_pred = (sh.slice_type == SliceType::P && pps.weighted_pred_flag) ||
                          (sh.slice_type == SliceType::B && pps.weighted_bipred_flag);
        const PredWeightTable& pwt = sh.pred_weight_table;
        p.luma_log2_wd = static_cast<int32_t>(pwt.luma_log2_weight_denom);
        p.chroma_log2_wd = static_cast<int32_t>(pwt.luma_log2_weight_denom) + pwt.delta_chroma_log2_weight_denom;
        const int nl = sh.slice_type == SliceType::B ? 2 : (sh.slice_type == SliceType::P ? 1 : 0);
        for (int l = 0; l < nl; l++) {
            const int n = l ? ctx.dpb->num_ref_list1() : ctx.dpb->num_ref_list0();
            if (n > HEVCR_MAX_REFS) return fail("reference list too long");
            p.num_ref[l] = n;
            for (int i = 0; i < n; i++) {
                const Picture* r = l ? ctx.dpb->ref_pic_list1(i) : ctx.dpb->ref_pic_list0(i);
                if (!r) return fail("missing reference picture");
                auto it = g_tap.pic_ids.find(r);
                if (it == g_tap.pic_ids.end()) return fail("reference picture not decoded by this run");
                p.ref_slot[l][i] = it->second;
                p.ref_poc[l][i] = r->poc;
                const auto& rw = l ? pwt.l1[static_cast<size_t>(i)] : pwt.l0[static_cast<size_t>(i)];
                p.weights[l][i].w[0] = rw.luma_weight;
                p.weights[l][i].o[0] = rw.luma_offset;
                for (int c = 0; c < 2; c++) {
                    p.weights[l][i].w[1 + c] = rw.chroma_weight[c];
                    p.weights[l][i].o[1 + c] = rw.chroma_offset[c];
                }
            }
        }
        return true;
    }

    bool build() {
        if (!g_tap.unsupported.empty()) return fail(g_tap.unsupported.c_str());
        if (!params()) return false;
        for (size_t i = 0; i < g_tap.coeffs.size(); i++) {
            const CoeffCall& k = g_tap.coeffs[i];
            if (!coeff_at.emplace(std::make_tuple(k.cidx, k.x, k.y, k.log2), i).second)
                return fail("two coded blocks at one position");
        }
        for (size_t i = 0; i < g_tap.inter.size(); i++) {
            const InterCall& m = g_tap.inter[i];
            if (!inter_at.emplace(std::make_tuple(m.x, m.y, m.w, m.h), i).second)
                return fail("two prediction units at one position");
        }
        const int ctb = sps.CtbLog2SizeY;
        for (int ry = 0; ry < sps.PicHeightInCtbsY; ry++)
            for (int rx = 0; rx < sps.PicWidthInCtbsY; rx++)
                if (!quadtree(rx << ctb, ry << ctb, ctb)) return false;
        if (inter_used != g_tap.inter.size()) return fail("prediction calls outside the CU walk");
        for (const CoeffCall& k : g_tap.coeffs)
            if (!k.used) return fail("coded block outside the transform-tree walk");
        for (int i = 0; i < sps.PicSizeInCtbsY; i++) {
            const auto& s = ctx.sao_params[i];
            hevcr_sao r{};
            for (int c = 0; c < 3; c++) {
                r.type[c] = static_cast<uint8_t>(s.sao_type_idx[c]);
                r.band_pos[c] = static_cast<uint8_t>(s.sao_band_position[c]);
                r.eo_class[c] = static_cast<uint8_t>(s.sao_eo_class[c]);
                for (int k = 0; k < 5; k++) {
                    if (s.sao_offset_val[c][k] < -128 || s.sao_offset_val[c][k] > 127) return fail("SAO offset range");
                    r.offset[c][k] = static_cast<int8_t>(s.sao_offset_val[c][k]);
                }
            }
            sao.push_back(r);
        }
        return true;
    }
};

void emit(DecodingContext& ctx, Builder* b) {
    const Picture& pic = *ctx.pic;
    TapHeader h{};
    std::memcpy(h.magic, "HTAP", 4);
    h.pic_id = g_tap.cur_id;
    h.cvs_id = pic.cvs_id;
    h.width = pic.width[0];
    h.height = pic.height[0];
    h.bit_depth = pic.bit_depth_luma;
    h.crop[0] = pic.conf_win_left;
    h.crop[1] = pic.conf_win_right;
    h.crop[2] = pic.conf_win_top;
    h.crop[3] = pic.conf_win_bottom;
    const bool ok = b && b->err.empty();
    h.status = ok ? 0 : 1;
    if (!ok) std::snprintf(h.reason, sizeof(h.reason), "%s", b ? b->err.c_str() : "no builder");
    if (ok) {
        h.num_cus = static_cast<uint32_t>(b->cus.size());
        h.num_pus = static_cast<uint32_t>(b->pus.size());
        h.num_tus = static_cast<uint32_t>(b->tus.size());
        h.num_sao = static_cast<uint32_t>(b->sao.size());
        h.num_coeffs = b->pool.size();
        h.has_scaling = b->has_scaling ? 1 : 0;
    }
    FILE* f = g_tap.out;
    std::fwrite(&h, sizeof(h), 1, f);
    if (ok) {
        std::fwrite(&b->p, sizeof(b->p), 1, f);
        std::fwrite(b->cus.data(), sizeof(hevcr_cu), b->cus.size(), f);
        std::fwrite(b->pus.data(), sizeof(hevcr_pu), b->pus.size(), f);
        std::fwrite(b->tus.data(), sizeof(hevcr_tu), b->tus.size(), f);
        std::fwrite(b->pool.data(), sizeof(int16_t), b->pool.size(), f);
        std::fwrite(b->sao.data(), sizeof(hevcr_sao), b->sao.size(), f);
        if (b->has_scaling) std::fwrite(&b->scaling, sizeof(b->scaling), 1, f);
    }
    for (int st = 0; st < 3; st++)
        for (int c = 0; c < 3; c++)
            std::fwrite(g_tap.stage[st][c].data(), 2, g_tap.stage[st][c].size(), f);
}

std::unique_ptr<Builder> g_builder;

}  // namespace

extern "C" {

// Index of the last perform_dequant call made by this thread: the
// perform_transform_inverse call that follows it carries transform_skip_flag.
static thread_local size_t t_last_dequant = SIZE_MAX;

void WRAP(W_DEQUANT)(DecodingContext& ctx, int x0, int y0, int log2, int cidx, int qp, const int16_t* coeffs,
                     int16_t* scaled) {
    CoeffCall k;
    k.x = x0;
    k.y = y0;
    k.log2 = log2;
    k.cidx = cidx;
    k.qp = qp;
    k.c.assign(coeffs, coeffs + (1 << (2 * log2)));
    {
        std::lock_guard<std::mutex> g(g_tap.mu);
        g_tap.coeffs.push_back(std::move(k));
        t_last_dequant = g_tap.coeffs.size() - 1;
    }
    REAL(W_DEQUANT)(ctx, x0, y0, log2, cidx, qp, coeffs, scaled);
}

void WRAP(W_TRANSFORM)(int log2, int cidx, bool intra, bool tskip, int bd, const int16_t* scaled, int16_t* res) {
    {
        std::lock_guard<std::mutex> g(g_tap.mu);
        const size_t i = t_last_dequant;
        if (i >= g_tap.coeffs.size() || g_tap.coeffs[i].cidx != cidx || g_tap.coeffs[i].log2 != log2)
            g_tap.unsupported = "transform call without its dequantisation";
        else
            g_tap.coeffs[i].tskip = tskip ? 1 : 0;
        t_last_dequant = SIZE_MAX;
    }
    REAL(W_TRANSFORM)(log2, cidx, intra, tskip, bd, scaled, res);
}

void WRAP(W_INTER)(DecodingContext& ctx, int x, int y, int w, int h, int cidx, const MV& mv0, const MV& mv1, int r0,
                   int r1, bool p0, bool p1, int16_t* pred) {
    if (cidx == 0) {
        InterCall m{x, y, w, h, {mv0, mv1}, {r0, r1}, {p0, p1}};
        std::lock_guard<std::mutex> g(g_tap.mu);
        g_tap.inter.push_back(m);
    }
    REAL(W_INTER)(ctx, x, y, w, h, cidx, mv0, mv1, r0, r1, p0, p1, pred);
}

// Called once per picture on the decoding thread, after every slice's CTBs.
void WRAP(W_DEBLOCK)(DecodingContext& ctx) {
    g_tap.cur_id = g_tap.next_id++;
    g_builder = std::make_unique<Builder>(ctx);
    g_builder->build();
    snapshot(*ctx.pic, 0);
    REAL(W_DEBLOCK)(ctx);
    snapshot(*ctx.pic, 1);
}

void WRAP(W_SAO)(DecodingContext& ctx) {
    REAL(W_SAO)(ctx);
    snapshot(*ctx.pic, 2);
    emit(ctx, g_builder.get());
    g_tap.pic_ids[ctx.pic] = g_tap.cur_id;
    g_tap.coeffs.clear();
    g_tap.inter.clear();
    g_tap.unsupported.clear();
    g_builder.reset();
}

}  // extern "C"

int main(int argc, char** argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s input.265 output.tap\n", argv[0]);
        return 2;
    }
    std::ifstream in(argv[1], std::ios::binary | std::ios::ate);
    if (!in) {
        std::fprintf(stderr, "cannot open %s\n", argv[1]);
        return 1;
    }
    std::vector<uint8_t> data(static_cast<size_t>(in.tellg()));
    in.seekg(0);
    in.read(reinterpret_cast<char*>(data.data()), static_cast<std::streamsize>(data.size()));
    g_tap.out = std::fopen(argv[2], "wb");
    if (!g_tap.out) {
        std::fprintf(stderr, "cannot write %s\n", argv[2]);
        return 1;
    }
    Decoder dec;
    const auto st = dec.decode(data.data(), data.size());
    std::fclose(g_tap.out);
    if (st == DecodeStatus::ERROR) {
        std::fprintf(stderr, "decode failed\n");
        return 1;
    }
    std::printf("%d pictures\n", g_tap.next_id);
    return 0;
}
