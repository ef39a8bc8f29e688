// Host runtime of the reconstruction engine: device DPB per stream, syntax
// validation and upload, derived work lists, batched stage launches (direct
// or replayed from a captured hipGraph) and stage timing.
// Exposes the C-ABI of include/hevcr.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <set>
#include <string>
#include <utility>
#include <vector>

#include "engine.hpp"
#include "hevcr.h"

namespace hevcr {

void launch_expand(hipStream_t s, const FrameDev* fr, const uint32_t* cu_prefix, int nframes, uint32_t total_cus);
template <typename T>
void launch_mc(hipStream_t, const FrameDev*, const McTile*, uint32_t);
template <typename T>
void launch_residual(hipStream_t, const FrameDev*, const ResBlock* const*, const uint32_t*);
template <typename T>
void launch_intra(hipStream_t, const FrameDev*, const IntraItem*, uint32_t);
template <typename T>
void launch_deblock(hipStream_t, const FrameDev*, const uint32_t*, int, uint32_t);
template <typename T>
void launch_sao(hipStream_t, const FrameDev*, const uint32_t*, int, uint32_t);
void launch_copy(hipStream_t s, const void* src, void* dst, size_t bytes);

}  // namespace hevcr

using namespace hevcr;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int last_launch_error() {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(HEVCR_EDEVICE, "kernel launch: %s", hipGetErrorString(e));
    return HEVCR_OK;
}

#define HIPCHK(expr)                                                                                 \
    do {                                                                                             \
        hipError_t e_ = (expr);                                                                      \
        if (e_ != hipSuccess) return fail(HEVCR_EDEVICE, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    int reserve(size_t bytes) {
        if (bytes <= cap) return HEVCR_OK;
        if (p) {
            HIPCHK(hipFree(p));
            p = nullptr;
            cap = 0;
        }
        size_t want = std::max<size_t>(bytes, 256);
        HIPCHK(hipMalloc(&p, want));
        cap = want;
        return HEVCR_OK;
    }
    template <typename U>
    U* as() const {
        return static_cast<U*>(p);
    }
};

int upload(DevBuf& b, const void* src, size_t bytes, hipStream_t s) {
    int rc = b.reserve(bytes);
    if (rc) return rc;
    if (bytes) HIPCHK(hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, s));
    return HEVCR_OK;
}

struct Picture {
    DevBuf mem;
    void* plane[3] = {nullptr, nullptr, nullptr};
};

struct Stream {
    int w, h, bd, bpp, nslots;
    int stride[3];
    size_t plane_bytes[3];
    std::vector<std::unique_ptr<Picture>> slots;
};

}  // namespace

struct hevcr_batch;

struct hevcr_frame {
    int stream_id = 0;
    hevcr_frame_params p{};
    FrameConst fc{};
    int bpp = 1;
    bool sao_active = false;
    bool has_intra = false;
    uint32_t ncu = 0, npu = 0, ntu = 0, nctb = 0;
    DevBuf cus, pus, tus, coeffs, sao, ops, op_begin, cu_first_tu, cu_first_pu;
    DevBuf info4, mot, resid;  // per-frame side information and intra residual planes
    Picture scratch;           // pre-SAO reconstruction
    // host work lists, concatenated into batches
    std::vector<McTile> tiles;
    std::vector<ResBlock> res[4];
    std::vector<std::vector<uint16_t>> diag_ctus;  // CTUs with intra work, per diagonal x + 2y
    double bytes[HEVCR_NUM_STAGES] = {};
    hevcr_batch* plain = nullptr;  // cached one-frame batch of the plain form
    int plain_slot = -1, plain_stop = -1;
};

struct hevcr_engine {
    int device = 0;
    hipStream_t stream = nullptr;
    std::map<int, std::unique_ptr<Stream>> streams;
    int next_stream = 1;
    bool profiling = false;
    struct Pending {
        int stage;
        hipEvent_t a, b;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> free_events;
    double stage_ms[HEVCR_NUM_STAGES] = {};
    uint64_t stage_n[HEVCR_NUM_STAGES] = {};
    DevBuf kat_coeffs, kat_out, kat_blocks, kat_fdev;
};

namespace {

// -- stage timing -----------------------------------------------------------

hipEvent_t get_event(hevcr_engine* e) {
    if (!e->free_events.empty()) {
        hipEvent_t ev = e->free_events.back();
        e->free_events.pop_back();
        return ev;
    }
    hipEvent_t ev = nullptr;
    if (hipEventCreate(&ev) != hipSuccess) return nullptr;
    return ev;
}

struct StageTimer {
    hevcr_engine* e;
    int stage;
    hipEvent_t a = nullptr;
    StageTimer(hevcr_engine* eng, int st, bool on) : e(eng), stage(st) {
        if (on) {
            a = get_event(e);
            if (a) (void)hipEventRecord(a, e->stream);
        }
    }
    ~StageTimer() {
        if (!a) return;
        hipEvent_t b = get_event(e);
        if (!b) return;
        (void)hipEventRecord(b, e->stream);
        e->pending.push_back({stage, a, b});
    }
};

int collect(hevcr_engine* e) {
    for (auto& p : e->pending) {
        HIPCHK(hipEventSynchronize(p.b));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, p.a, p.b));
        e->stage_ms[p.stage] += ms;
        e->stage_n[p.stage] += 1;
        e->free_events.push_back(p.a);
        e->free_events.push_back(p.b);
    }
    e->pending.clear();
    return HEVCR_OK;
}

// -- syntax validation --------------------------------------------------------

uint32_t zorder(int x, int y) {  // 4x4 units inside a CTB
    uint32_t z = 0;
    for (int i = 0; i < 8; i++) z |= (((uint32_t)x >> i) & 1u) << (2 * i) | (((uint32_t)y >> i) & 1u) << (2 * i + 1);
    return z;
}

// Decoding-order key of a block: CTB raster address, then z-order of 4x4 units.
uint64_t dec_key(const hevcr_frame_params& p, int x, int y) {
    const int wctb = (p.width + (1 << p.ctb_log2) - 1) >> p.ctb_log2;
    const uint64_t ctb = (uint64_t)(y >> p.ctb_log2) * wctb + (x >> p.ctb_log2);
    const int m = (1 << p.ctb_log2) - 1;
    return (ctb << 16) | zorder((x & m) >> 2, (y & m) >> 2);
}

// Expected PU rectangles of an inter partition (coding_tree.cpp:925-952).
int part_rects(int part, int x0, int y0, int s, int r[4][4]) {
    const int h = s / 2, q = s / 4;
    switch (part) {
        case HEVCR_PART_2Nx2N: r[0][0] = x0; r[0][1] = y0; r[0][2] = s; r[0][3] = s; return 1;
        case HEVCR_PART_2NxN:
            r[0][0] = x0; r[0][1] = y0; r[0][2] = s; r[0][3] = h;
            r[1][0] = x0; r[1][1] = y0 + h; r[1][2] = s; r[1][3] = h; return 2;
        case HEVCR_PART_Nx2N:
            r[0][0] = x0; r[0][1] = y0; r[0][2] = h; r[0][3] = s;
            r[1][0] = x0 + h; r[1][1] = y0; r[1][2] = h; r[1][3] = s; return 2;
        case HEVCR_PART_NxN:
            for (int k = 0; k < 4; k++) { r[k][0] = x0 + (k & 1) * h; r[k][1] = y0 + (k >> 1) * h; r[k][2] = h; r[k][3] = h; }
            return 4;
        case HEVCR_PART_2NxnU:
            r[0][0] = x0; r[0][1] = y0; r[0][2] = s; r[0][3] = q;
            r[1][0] = x0; r[1][1] = y0 + q; r[1][2] = s; r[1][3] = s - q; return 2;
        case HEVCR_PART_2NxnD:
            r[0][0] = x0; r[0][1] = y0; r[0][2] = s; r[0][3] = s - q;
            r[1][0] = x0; r[1][1] = y0 + s - q; r[1][2] = s; r[1][3] = q; return 2;
        case HEVCR_PART_nLx2N:
            r[0][0] = x0; r[0][1] = y0; r[0][2] = q; r[0][3] = s;
            r[1][0] = x0 + q; r[1][1] = y0; r[1][2] = s - q; r[1][3] = s; return 2;
        case HEVCR_PART_nRx2N:
            r[0][0] = x0; r[0][1] = y0; r[0][2] = s - q; r[0][3] = s;
            r[1][0] = x0 + s - q; r[1][1] = y0; r[1][2] = q; r[1][3] = s; return 2;
        default: return 0;
    }
}

int validate(const Stream& st, const hevcr_frame_syntax* syn) {
    const hevcr_frame_params& p = syn->p;
    if (p.width != st.w || p.height != st.h || p.bit_depth != st.bd)
        return fail(HEVCR_EINVAL, "frame %dx%d/%d does not match stream %dx%d/%d", p.width, p.height, p.bit_depth, st.w,
                    st.h, st.bd);
    if (p.min_cb_log2 != 3) return fail(HEVCR_EUNSUPPORTED, "min_cb_log2 %d (only 3)", p.min_cb_log2);
    if (p.ctb_log2 < 4 || p.ctb_log2 > 6) return fail(HEVCR_EINVAL, "ctb_log2 %d", p.ctb_log2);
    if (p.slice_type < 0 || p.slice_type > 2) return fail(HEVCR_EINVAL, "slice_type %d", p.slice_type);
    const int nl = p.slice_type == HEVCR_SLICE_B ? 2 : (p.slice_type == HEVCR_SLICE_P ? 1 : 0);
    for (int l = 0; l < 2; l++) {
        const int want_min = l < nl ? 1 : 0;
        if (p.num_ref[l] < want_min || p.num_ref[l] > HEVCR_MAX_REFS || (l >= nl && p.num_ref[l] != 0))
            return fail(HEVCR_EINVAL, "num_ref[%d] = %d", l, p.num_ref[l]);
        for (int i = 0; i < p.num_ref[l]; i++)
            if (p.ref_slot[l][i] < 0 || p.ref_slot[l][i] >= st.nslots)
                return fail(HEVCR_EINVAL, "ref_slot[%d][%d] = %d", l, i, p.ref_slot[l][i]);
    }
    if (p.weighted_pred && (p.luma_log2_wd < 0 || p.luma_log2_wd > 7 || p.chroma_log2_wd < 0 || p.chroma_log2_wd > 7))
        return fail(HEVCR_EINVAL, "weight denominators %d/%d", p.luma_log2_wd, p.chroma_log2_wd);
    if (p.beta_offset_div2 < -6 || p.beta_offset_div2 > 6 || p.tc_offset_div2 < -6 || p.tc_offset_div2 > 6)
        return fail(HEVCR_EINVAL, "deblocking offsets");
    for (int v : {p.pps_cb_qp_offset, p.pps_cr_qp_offset, p.slice_cb_qp_offset, p.slice_cr_qp_offset})
        if (v < -12 || v > 12) return fail(HEVCR_EINVAL, "chroma qp offset %d", v);
    if (!syn->sao && p.sao_enabled) return fail(HEVCR_EINVAL, "sao records missing");
    if ((syn->num_cus && !syn->cus) || (syn->num_pus && !syn->pus) || (syn->num_tus && !syn->tus) ||
        (syn->num_coeffs && !syn->coeffs))
        return fail(HEVCR_EINVAL, "null syntax array");

    const int qbd = 6 * (p.bit_depth - 8);
    const int w8 = p.width / 8, h8 = p.height / 8, gw = p.width / 4;
    std::vector<uint8_t> cov8((size_t)w8 * h8, 0), cov4((size_t)gw * (p.height / 4), 0);
    uint64_t prev_key = 0;
    uint32_t pi = 0, ti = 0;
    for (uint32_t ci = 0; ci < syn->num_cus; ci++) {
        const hevcr_cu& cu = syn->cus[ci];
        const int s = 1 << cu.log2_size;
        if (cu.log2_size < 3 || cu.log2_size > p.ctb_log2 || (cu.x & (s - 1)) || (cu.y & (s - 1)) ||
            cu.x + s > p.width || cu.y + s > p.height)
            return fail(HEVCR_EINVAL, "cu %u geometry (%d,%d) log2 %d", ci, cu.x, cu.y, cu.log2_size);
        const uint64_t key = dec_key(p, cu.x, cu.y);
        if (ci && key <= prev_key) return fail(HEVCR_EINVAL, "cu %u out of decoding order", ci);
        prev_key = key;
        if (cu.qp_y < -qbd || cu.qp_y > 51) return fail(HEVCR_EINVAL, "cu %u qp %d", ci, cu.qp_y);
        if (cu.pred_mode > 2) return fail(HEVCR_EINVAL, "cu %u pred_mode", ci);
        if (p.slice_type == HEVCR_SLICE_I && cu.pred_mode != HEVCR_MODE_INTRA)
            return fail(HEVCR_EINVAL, "cu %u inter in an I slice", ci);
        for (int j = 0; j < s / 8; j++)
            for (int i = 0; i < s / 8; i++) {
                uint8_t& c = cov8[(size_t)(cu.y / 8 + j) * w8 + cu.x / 8 + i];
                if (c) return fail(HEVCR_EINVAL, "cu %u overlaps", ci);
                c = 1;
            }
        const bool intra = cu.pred_mode == HEVCR_MODE_INTRA;
        if (intra) {
            if (cu.part_mode != HEVCR_PART_2Nx2N && !(cu.part_mode == HEVCR_PART_NxN && cu.log2_size == 3))
                return fail(HEVCR_EINVAL, "cu %u intra partition %d", ci, cu.part_mode);
            const int npu = cu.part_mode == HEVCR_PART_NxN ? 4 : 1;
            for (int k = 0; k < npu; k++)
                if (cu.intra_luma[k] > 34) return fail(HEVCR_EINVAL, "cu %u luma mode", ci);
            if (cu.intra_chroma > 34) return fail(HEVCR_EINVAL, "cu %u chroma mode", ci);
        } else {
            int r[4][4];
            if (cu.pred_mode == HEVCR_MODE_SKIP && cu.part_mode != HEVCR_PART_2Nx2N)
                return fail(HEVCR_EINVAL, "cu %u skip partition", ci);
            const int n = part_rects(cu.part_mode, cu.x, cu.y, s, r);
            if (!n || n != cu.num_pu) return fail(HEVCR_EINVAL, "cu %u partition %d / %d PUs", ci, cu.part_mode, cu.num_pu);
            for (int k = 0; k < n; k++, pi++) {
                if (pi >= syn->num_pus) return fail(HEVCR_EINVAL, "pu list too short");
                const hevcr_pu& pu = syn->pus[pi];
                if (pu.x != r[k][0] || pu.y != r[k][1] || pu.w != r[k][2] || pu.h != r[k][3])
                    return fail(HEVCR_EINVAL, "pu %u does not match its CU partition", pi);
                if (!(pu.pred_flags & 3) || (pu.pred_flags & ~3u)) return fail(HEVCR_EINVAL, "pu %u pred flags", pi);
                for (int l = 0; l < 2; l++)
                    if ((pu.pred_flags >> l) & 1) {
                        if (l >= nl || pu.ref_idx[l] < 0 || pu.ref_idx[l] >= p.num_ref[l])
                            return fail(HEVCR_EINVAL, "pu %u ref_idx[%d] = %d", pi, l, pu.ref_idx[l]);
                        if (p.weighted_pred) {
                            const hevcr_weight& w = p.weights[l][pu.ref_idx[l]];
                            for (int c = 0; c < 3; c++)
                                if (w.w[c] < -128 || w.w[c] > 255 || w.o[c] < -128 || w.o[c] > 127)
                                    return fail(HEVCR_EINVAL, "pu %u weight out of range", pi);
                        }
                    }
            }
        }
        // transform units of this CU
        int tcount = 0;
        for (; ti < syn->num_tus && syn->tus[ti].cu_idx == ci; ti++, tcount++) {
            const hevcr_tu& t = syn->tus[ti];
            const int n = 1 << t.log2_size;
            if (cu.pred_mode == HEVCR_MODE_SKIP) return fail(HEVCR_EINVAL, "tu %u in a skip CU", ti);
            if (t.log2_size < 2 || t.log2_size > 5 || (t.x & (n - 1)) || (t.y & (n - 1)) || t.x < cu.x || t.y < cu.y ||
                t.x + n > cu.x + s || t.y + n > cu.y + s || t.blk_idx > 3)
                return fail(HEVCR_EINVAL, "tu %u geometry", ti);
            if (t.log2_size == 2 && (t.xbase != (t.x & ~7) || t.ybase != (t.y & ~7)))
                return fail(HEVCR_EINVAL, "tu %u base position", ti);
            if (t.cbf & ~7u) return fail(HEVCR_EINVAL, "tu %u cbf", ti);
            if (t.tskip & ~t.cbf) return fail(HEVCR_EINVAL, "tu %u transform skip without residual", ti);
            const bool chroma_here = t.log2_size > 2 || t.blk_idx == 3;
            const int nc = t.log2_size > 2 ? n / 2 : 4;
            for (int c = 0; c < 3; c++) {
                if (!((t.cbf >> c) & 1)) continue;
                if (c && !chroma_here) continue;
                const int bn = c ? nc : n;
                if ((t.tskip >> c) & 1 && bn != 4) return fail(HEVCR_EUNSUPPORTED, "tu %u transform skip above 4x4", ti);
                if ((uint64_t)t.coeff_off[c] + (uint64_t)bn * bn > syn->num_coeffs)
                    return fail(HEVCR_EINVAL, "tu %u coefficients out of range", ti);
            }
            for (int j = 0; j < n / 4; j++)
                for (int i = 0; i < n / 4; i++) {
                    uint8_t& c = cov4[(size_t)(t.y / 4 + j) * gw + t.x / 4 + i];
                    if (c) return fail(HEVCR_EINVAL, "tu %u overlaps", ti);
                    c = 1;
                }
        }
        if (ti < syn->num_tus && syn->tus[ti].cu_idx < ci) return fail(HEVCR_EINVAL, "tu %u out of order", ti);
        if (intra) {  // intra CUs are fully covered by their transform tree
            for (int j = 0; j < s / 4; j++)
                for (int i = 0; i < s / 4; i++)
                    if (!cov4[(size_t)(cu.y / 4 + j) * gw + cu.x / 4 + i])
                        return fail(HEVCR_EINVAL, "intra cu %u not covered by TUs", ci);
        }
        (void)tcount;
    }
    if (pi != syn->num_pus) return fail(HEVCR_EINVAL, "%u PU records left over", syn->num_pus - pi);
    if (ti != syn->num_tus) return fail(HEVCR_EINVAL, "%u TU records left over", syn->num_tus - ti);
    for (uint8_t c : cov8)
        if (!c) return fail(HEVCR_EINVAL, "picture not covered by CUs");
    if (p.sao_enabled) {
        const int ctb = 1 << p.ctb_log2;
        const int nctb = ((p.width + ctb - 1) / ctb) * ((p.height + ctb - 1) / ctb);
        const int maxo = (1 << ((p.bit_depth < 10 ? p.bit_depth : 10) - 5)) - 1;
        for (int i = 0; i < nctb; i++) {
            const hevcr_sao& sr = syn->sao[i];
            for (int c = 0; c < 3; c++) {
                if (sr.type[c] > 2 || sr.band_pos[c] > 31 || sr.eo_class[c] > 3) return fail(HEVCR_EINVAL, "sao ctb %d", i);
                for (int k = 0; k < 5; k++)
                    if (sr.offset[c][k] < -maxo || sr.offset[c][k] > maxo) return fail(HEVCR_EINVAL, "sao ctb %d offset", i);
            }
        }
    }
    return HEVCR_OK;
}

int chroma_qp_prime(const hevcr_frame_params& p, int qpy, int c) {
    const int qbd = 6 * (p.bit_depth - 8);
    const int off = c == 1 ? p.pps_cb_qp_offset + p.slice_cb_qp_offset : p.pps_cr_qp_offset + p.slice_cr_qp_offset;
    int qpi = std::min(57, std::max(-qbd, qpy + off));
    int qpc = qpi;
    if (qpi >= 30) qpc = qpi > 42 ? qpi - 6 : (qpi - 30 < 4 ? 29 + (qpi - 30) : 33 + ((qpi - 34) >> 1));
    return qpc + qbd;
}

FrameConst make_const(const Stream& st, const hevcr_frame_params& p) {
    FrameConst fc;
    std::memset(&fc, 0, sizeof(fc));
    fc.W = p.width;
    fc.H = p.height;
    fc.bd = p.bit_depth;
    fc.maxv = (1 << p.bit_depth) - 1;
    fc.ctb_log2 = p.ctb_log2;
    fc.wctb = (p.width + (1 << p.ctb_log2) - 1) >> p.ctb_log2;
    fc.hctb = (p.height + (1 << p.ctb_log2) - 1) >> p.ctb_log2;
    fc.gw = p.width / 4;
    fc.gh = p.height / 4;
    for (int c = 0; c < 3; c++) fc.stride[c] = st.stride[c];
    fc.slice_type = p.slice_type;
    fc.weighted = p.weighted_pred && p.slice_type != HEVCR_SLICE_I;
    fc.luma_log2_wd = p.luma_log2_wd;
    fc.chroma_log2_wd = p.chroma_log2_wd;
    for (int l = 0; l < 2; l++) {
        fc.num_ref[l] = p.num_ref[l];
        for (int i = 0; i < kMaxRefs; i++) {
            fc.ref_poc[l][i] = p.ref_poc[l][i];
            for (int c = 0; c < 3; c++) {
                fc.wt[l][i][c] = p.weights[l][i].w[c];
                fc.off[l][i][c] = p.weights[l][i].o[c];
            }
        }
    }
    fc.pps_cb_qp_offset = p.pps_cb_qp_offset;
    fc.pps_cr_qp_offset = p.pps_cr_qp_offset;
    fc.beta_offset_div2 = p.beta_offset_div2;
    fc.tc_offset_div2 = p.tc_offset_div2;
    fc.strong_intra_smoothing = p.strong_intra_smoothing;
    fc.intra_smoothing_disabled = p.intra_smoothing_disabled;
    return fc;
}

int alloc_picture(Stream& st, Picture& pic) {
    const size_t total = st.plane_bytes[0] + st.plane_bytes[1] + st.plane_bytes[2];
    int rc = pic.mem.reserve(total);
    if (rc) return rc;
    char* base = static_cast<char*>(pic.mem.p);
    pic.plane[0] = base;
    pic.plane[1] = base + st.plane_bytes[0];
    pic.plane[2] = base + st.plane_bytes[0] + st.plane_bytes[1];
    return HEVCR_OK;
}

Stream* find_stream(hevcr_engine* e, int id) {
    auto it = e->streams.find(id);
    return it == e->streams.end() ? nullptr : it->second.get();
}


// -- intra neighbour availability (intra_prediction.cpp:102-133) -------------

// The reference's availability of component sample (rx, ry) for the block
// whose top-left luma sample is (xl, yl): inside the picture, and in an
// earlier CTB in raster order or in the same CTB at a smaller z-scan address
// of 4x4 luma blocks (one slice, no tiles).
bool sample_avail(const FrameConst& fc, int c, int xl, int yl, int rx, int ry) {
    const int sub = c ? 1 : 0;
    if (rx < 0 || ry < 0 || rx >= (fc.W >> sub) || ry >= (fc.H >> sub)) return false;
    const int lx = rx << sub, ly = ry << sub;
    const int cx = xl >> fc.ctb_log2, cy = yl >> fc.ctb_log2, nx = lx >> fc.ctb_log2, ny = ly >> fc.ctb_log2;
    if (nx != cx || ny != cy) return ny * fc.wctb + nx < cy * fc.wctb + cx;
    const int ox = cx << fc.ctb_log2, oy = cy << fc.ctb_log2;
    return zorder((lx - ox) >> 2, (ly - oy) >> 2) < zorder((xl - ox) >> 2, (yl - oy) >> 2);
}

// Fills the five availability counts of an intra block of size n at
// component position (xc, yc). Availability is constant over each 4x4 luma
// block (4 luma / 2 chroma samples along an edge), so one test per unit.
int intra_counts(const FrameConst& fc, int c, int xc, int yc, int n, IntraOp& op) {
    const int sub = c ? 1 : 0, unit = c ? 2 : 4;
    const int xl = xc << sub, yl = yc << sub;
    auto av = [&](int rx, int ry) { return sample_avail(fc, c, xl, yl, rx, ry); };
    int nbl = 0, ntr = 0;
    bool gap = false;
    for (int k = 0; k < n; k += unit) {  // bottom-left, nearest row first
        if (av(xc - 1, yc + n + k)) {
            if (gap) return -1;
            nbl += unit;
        } else {
            gap = true;
        }
    }
    gap = false;
    for (int k = 0; k < n; k += unit) {  // top-right, nearest column first
        if (av(xc + n + k, yc - 1)) {
            if (gap) return -1;
            ntr += unit;
        } else {
            gap = true;
        }
    }
    const bool left = av(xc - 1, yc), top = av(xc, yc - 1);
    for (int k = unit; k < n; k += unit)
        if (av(xc - 1, yc + k) != left || av(xc + k, yc - 1) != top) return -1;
    op.nbl = (uint8_t)nbl;
    op.ntr = (uint8_t)ntr;
    op.left = left ? 1 : 0;
    op.top = top ? 1 : 0;
    op.corner = av(xc - 1, yc - 1) ? 1 : 0;
    return 0;
}

// -- frame preparation --------------------------------------------------------

int prepare_frame(hevcr_engine* e, Stream& st, int stream_id, const hevcr_frame_syntax* syn, hevcr_frame* f) {
    const hevcr_frame_params& p = syn->p;
    f->stream_id = stream_id;
    f->p = p;
    f->fc = make_const(st, p);
    f->bpp = st.bpp;
    f->ncu = syn->num_cus;
    f->npu = syn->num_pus;
    f->ntu = syn->num_tus;
    const FrameConst& fc = f->fc;
    const int b = st.bpp;
    const int nctb = fc.wctb * fc.hctb;
    f->nctb = (uint32_t)nctb;

    // SAO activity: the reference skips the whole pass when no CTB has SAO (sao.cpp:33-39)
    f->sao_active = false;
    if (p.sao_enabled)
        for (int i = 0; i < nctb && !f->sao_active; i++)
            for (int c = 0; c < 3; c++) f->sao_active |= syn->sao[i].type[c] != 0;

    // per-CU record ranges
    std::vector<uint32_t> first_tu(syn->num_cus + 1), first_pu(syn->num_cus + 1);
    {
        uint32_t ti = 0, pi = 0;
        for (uint32_t ci = 0; ci < syn->num_cus; ci++) {
            first_tu[ci] = ti;
            first_pu[ci] = pi;
            while (ti < syn->num_tus && syn->tus[ti].cu_idx == ci) ti++;
            if (syn->cus[ci].pred_mode != HEVCR_MODE_INTRA) pi += syn->cus[ci].num_pu;
        }
        first_tu[syn->num_cus] = ti;
        first_pu[syn->num_cus] = pi;
    }

    // MC tiles (<= 16x16 luma) of every PU
    f->tiles.clear();
    f->tiles.reserve(syn->num_pus * 2);
    double mc_bytes = 0;
    for (uint32_t i = 0; i < syn->num_pus; i++) {
        const hevcr_pu& pu = syn->pus[i];
        for (int ty = 0; ty < pu.h; ty += 16)
            for (int tx = 0; tx < pu.w; tx += 16) {
                McTile t;
                t.frame = 0;
                t.x = (uint16_t)(pu.x + tx);
                t.y = (uint16_t)(pu.y + ty);
                t.w = (uint8_t)std::min(16, pu.w - tx);
                t.h = (uint8_t)std::min(16, pu.h - ty);
                std::memcpy(t.mv, pu.mv, sizeof(t.mv));
                t.ref[0] = pu.ref_idx[0];
                t.ref[1] = pu.ref_idx[1];
                t.flags = pu.pred_flags;
                t.pad = 0;
                f->tiles.push_back(t);
            }
        const int nlists = (pu.pred_flags & 1) + ((pu.pred_flags >> 1) & 1);
        const double fp = (double)(pu.w + 7) * (pu.h + 7) + 2.0 * (pu.w / 2 + 3) * (pu.h / 2 + 3);
        mc_bytes += b * (nlists * fp + 1.5 * pu.w * pu.h);
    }
    std::memset(f->bytes, 0, sizeof(f->bytes));
    f->bytes[HEVCR_STAGE_MC] = mc_bytes;

    // residual blocks per size; intra operations per CTU and component
    for (auto& r : f->res) r.clear();
    std::vector<std::vector<IntraOp>> ops_c[3];
    for (int c = 0; c < 3; c++) ops_c[c].resize((size_t)nctb);
    std::vector<uint8_t> ctu_has_intra((size_t)nctb, 0);
    const int qbd = 6 * (p.bit_depth - 8);
    double res_bytes = 0, intra_bytes = 0;
    for (uint32_t ti = 0; ti < syn->num_tus; ti++) {
        const hevcr_tu& t = syn->tus[ti];
        const hevcr_cu& cu = syn->cus[t.cu_idx];
        const bool intra = cu.pred_mode == HEVCR_MODE_INTRA;
        const int ctu = (t.y >> p.ctb_log2) * fc.wctb + (t.x >> p.ctb_log2);
        const bool chroma_here = t.log2_size > 2 || t.blk_idx == 3;
        const int l2c = t.log2_size > 2 ? t.log2_size - 1 : 2;
        const int xb = t.log2_size == 2 ? t.xbase : t.x, yb = t.log2_size == 2 ? t.ybase : t.y;
        for (int c = 0; c < 3; c++) {
            if (c && !chroma_here) continue;
            const int l2 = c ? l2c : t.log2_size;
            const int xc = c ? xb / 2 : t.x, yc = c ? yb / 2 : t.y;
            const bool coded = (t.cbf >> c) & 1;
            if (coded) {
                ResBlock rb;
                rb.coeff_off = t.coeff_off[c];
                rb.x = (uint16_t)xc;
                rb.y = (uint16_t)yc;
                rb.log2 = (uint8_t)l2;
                rb.cidx = (uint8_t)c;
                rb.qp = (uint8_t)(c ? chroma_qp_prime(p, cu.qp_y, c) : cu.qp_y + qbd);
                rb.flags = (uint8_t)((intra ? kResIntra : kResToPlane) | (intra && c == 0 && l2 == 2 ? kResDst : 0) |
                                     (((t.tskip >> c) & 1) ? kResTskip : 0));
                rb.frame = 0;
                rb.pad = 0;
                f->res[l2 - 2].push_back(rb);
                res_bytes += (2.0 + (intra ? 2.0 : 2.0 * b)) * (1 << (2 * l2));
            }
            if (intra) {
                IntraOp op;
                op.x = (uint16_t)xc;
                op.y = (uint16_t)yc;
                op.log2 = (uint8_t)l2;
                if (c == 0) {
                    const int s = 1 << cu.log2_size, h = s / 2;
                    const int k = cu.part_mode == HEVCR_PART_NxN ? (t.y >= cu.y + h) * 2 + (t.x >= cu.x + h) : 0;
                    op.mode = cu.intra_luma[k];
                } else {
                    op.mode = cu.intra_chroma;
                }
                op.cbf = coded ? 1 : 0;
                if (intra_counts(fc, c, xc, yc, 1 << l2, op))
                    return fail(HEVCR_ERROR, "tu %u: neighbour availability is not prefix-shaped", ti);
                ops_c[c][(size_t)ctu].push_back(op);
                ctu_has_intra[(size_t)ctu] = 1;
                intra_bytes += (coded ? 2.0 : 0.0) * (1 << (2 * l2));
            }
        }
    }
    f->bytes[HEVCR_STAGE_RESIDUAL] = res_bytes;

    std::vector<IntraOp> ops;
    std::vector<uint32_t> op_begin((size_t)nctb * 3 + 1, 0);
    for (int ctu = 0; ctu < nctb; ctu++)
        for (int c = 0; c < 3; c++) {
            if (ops_c[c][(size_t)ctu].size() > 256)  // k_intra stages at most 256 operations per CTU
                return fail(HEVCR_ERROR, "ctu %d: %zu intra blocks", ctu, ops_c[c][(size_t)ctu].size());
            op_begin[(size_t)ctu * 3 + c] = (uint32_t)ops.size();
            ops.insert(ops.end(), ops_c[c][(size_t)ctu].begin(), ops_c[c][(size_t)ctu].end());
        }
    op_begin[(size_t)nctb * 3] = (uint32_t)ops.size();
    f->has_intra = !ops.empty();
    const int ndiag = fc.wctb + 2 * (fc.hctb - 1);
    f->diag_ctus.assign((size_t)ndiag, {});
    uint32_t nctu_intra = 0;
    for (int y = 0; y < fc.hctb; y++)
        for (int x = 0; x < fc.wctb; x++)
            if (ctu_has_intra[(size_t)(y * fc.wctb + x)]) {
                f->diag_ctus[(size_t)(x + 2 * y)].push_back((uint16_t)(y * fc.wctb + x));
                nctu_intra++;
            }
    const int ctbs = 1 << p.ctb_log2;
    intra_bytes += (double)nctu_intra * (2.0 * b + 2.0) * 1.5 * ctbs * ctbs;
    f->bytes[HEVCR_STAGE_INTRA] = intra_bytes;
    const double S = 1.5 * p.width * p.height;
    f->bytes[HEVCR_STAGE_DEBLOCK] = p.deblocking_disabled ? 0 : 2 * S * b + 4.0 * fc.gw * fc.gh;
    f->bytes[HEVCR_STAGE_SAO] = f->sao_active ? 2 * S * b + 16.0 * nctb : 0;
    f->bytes[HEVCR_STAGE_EXPAND] = syn->num_cus * sizeof(hevcr_cu) + syn->num_tus * sizeof(hevcr_tu) +
                                   syn->num_pus * sizeof(hevcr_pu) + 4.0 * fc.gw * fc.gh +
                                   (syn->num_pus ? 12.0 * fc.gw * fc.gh : 0);

    // device side: records, side information, residual planes, scratch picture
    hipStream_t s = e->stream;
    const size_t cells = (size_t)fc.gw * fc.gh;
    int rc;
    if ((rc = upload(f->cus, syn->cus, syn->num_cus * sizeof(hevcr_cu), s)) ||
        (rc = upload(f->pus, syn->pus, syn->num_pus * sizeof(hevcr_pu), s)) ||
        (rc = upload(f->tus, syn->tus, syn->num_tus * sizeof(hevcr_tu), s)) ||
        (rc = upload(f->coeffs, syn->coeffs, syn->num_coeffs * sizeof(int16_t), s)) ||
        (rc = upload(f->sao, syn->sao, p.sao_enabled ? (size_t)nctb * sizeof(hevcr_sao) : 0, s)) ||
        (rc = upload(f->ops, ops.data(), ops.size() * sizeof(IntraOp), s)) ||
        (rc = upload(f->op_begin, op_begin.data(), op_begin.size() * sizeof(uint32_t), s)) ||
        (rc = upload(f->cu_first_tu, first_tu.data(), first_tu.size() * sizeof(uint32_t), s)) ||
        (rc = upload(f->cu_first_pu, first_pu.data(), first_pu.size() * sizeof(uint32_t), s)) ||
        (rc = f->info4.reserve(cells * 4)) || (rc = f->mot.reserve(syn->num_pus ? cells * sizeof(Motion4) : 0)))
        return rc;
    if (f->has_intra && (rc = f->resid.reserve((st.plane_bytes[0] + 2 * st.plane_bytes[1]) / st.bpp * 2))) return rc;
    if (f->sao_active && (rc = alloc_picture(st, f->scratch))) return rc;
    HIPCHK(hipStreamSynchronize(s));  // host vectors go out of scope
    return HEVCR_OK;
}

void frame_dev(const Stream& st, const hevcr_frame* f, Picture& slot, bool full, FrameDev& d) {
    std::memset(&d, 0, sizeof(d));
    d.fc = f->fc;
    const bool use_sao = full && f->sao_active;
    for (int c = 0; c < 3; c++) {
        d.rec[c] = use_sao ? f->scratch.plane[c] : slot.plane[c];
        d.out[c] = slot.plane[c];
    }
    if (f->has_intra) {
        int16_t* base = f->resid.as<int16_t>();
        d.res[0] = base;
        d.res[1] = base + st.plane_bytes[0] / st.bpp;
        d.res[2] = d.res[1] + st.plane_bytes[1] / st.bpp;
    }
    for (int l = 0; l < 2; l++)
        for (int i = 0; i < f->p.num_ref[l]; i++)
            for (int c = 0; c < 3; c++) d.ref[l][i][c] = st.slots[(size_t)f->p.ref_slot[l][i]]->plane[c];
    d.info4 = f->info4.as<uint32_t>();
    d.mot = f->mot.as<Motion4>();
    d.cus = f->cus.as<hevcr_cu>();
    d.tus = f->tus.as<hevcr_tu>();
    d.pus = f->pus.as<hevcr_pu>();
    d.cu_first_tu = f->cu_first_tu.as<uint32_t>();
    d.cu_first_pu = f->cu_first_pu.as<uint32_t>();
    d.coeffs = f->coeffs.as<int16_t>();
    d.sao = f->sao.as<hevcr_sao>();
    d.ops = f->ops.as<IntraOp>();
    d.op_begin = f->op_begin.as<uint32_t>();
}

}  // namespace

// -- batches --------------------------------------------------------------------

// Frames of one pixel type launched together.
struct BatchGroup {
    int bpp = 1;
    std::vector<int> members;
    DevBuf fdev, cu_prefix, tile_prefix, ctb_prefix, tiles, res[4], items;
    uint32_t total_cus = 0, ntiles = 0, nres[4] = {0, 0, 0, 0}, total_tiles = 0, total_ctbs = 0;
    std::vector<uint32_t> diag_start;  // offsets of each diagonal in `items` (size ndiag + 1)
};

struct hevcr_batch {
    std::vector<hevcr_frame*> frames;
    std::vector<int> slots;
    int stop = HEVCR_RUN_FULL;
    BatchGroup groups[2];
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    double bytes[HEVCR_NUM_STAGES] = {};
    ~hevcr_batch() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

namespace {

int build_group(hevcr_engine* e, hevcr_batch* b, BatchGroup& g) {
    hipStream_t s = e->stream;
    const int n = (int)g.members.size();
    std::vector<FrameDev> devs((size_t)n);
    std::vector<uint32_t> cu_prefix(1, 0), tile_prefix(1, 0), ctb_prefix(1, 0);
    std::vector<McTile> tiles;
    std::vector<ResBlock> res[4];
    size_t ndiag = 0;
    for (int k = 0; k < n; k++) {
        const int fi = g.members[(size_t)k];
        hevcr_frame* f = b->frames[(size_t)fi];
        Stream& st = *e->streams[f->stream_id];
        frame_dev(st, f, *st.slots[(size_t)b->slots[(size_t)fi]], b->stop == HEVCR_RUN_FULL, devs[(size_t)k]);
        cu_prefix.push_back(cu_prefix.back() + f->ncu);
        const bool dbk = b->stop != HEVCR_RUN_STOP_RECON && !f->p.deblocking_disabled;
        const uint32_t ntiles = (uint32_t)(((f->fc.W + 4 + 63) >> 6) * ((f->fc.H + 4 + 63) >> 6));
        tile_prefix.push_back(tile_prefix.back() + (dbk ? ntiles : 0));
        const bool sao = b->stop == HEVCR_RUN_FULL && f->sao_active;
        ctb_prefix.push_back(ctb_prefix.back() + (sao ? f->nctb : 0));
        for (McTile t : f->tiles) {
            t.frame = (uint16_t)k;
            tiles.push_back(t);
        }
        for (int q = 0; q < 4; q++)
            for (ResBlock r : f->res[q]) {
                r.frame = (uint16_t)k;
                res[q].push_back(r);
            }
        ndiag = std::max(ndiag, f->diag_ctus.size());
    }
    std::vector<IntraItem> items;
    g.diag_start.assign(1, 0);
    for (size_t dgi = 0; dgi < ndiag; dgi++) {
        for (int k = 0; k < n; k++) {
            const hevcr_frame* f = b->frames[(size_t)g.members[(size_t)k]];
            if (dgi >= f->diag_ctus.size()) continue;
            for (uint16_t ctu : f->diag_ctus[dgi]) items.push_back({(uint16_t)k, ctu});
        }
        g.diag_start.push_back((uint32_t)items.size());
    }
    g.total_cus = cu_prefix.back();
    g.total_tiles = tile_prefix.back();
    g.total_ctbs = ctb_prefix.back();
    g.ntiles = (uint32_t)tiles.size();
    int rc;
    if ((rc = upload(g.fdev, devs.data(), devs.size() * sizeof(FrameDev), s)) ||
        (rc = upload(g.cu_prefix, cu_prefix.data(), cu_prefix.size() * 4, s)) ||
        (rc = upload(g.tile_prefix, tile_prefix.data(), tile_prefix.size() * 4, s)) ||
        (rc = upload(g.ctb_prefix, ctb_prefix.data(), ctb_prefix.size() * 4, s)) ||
        (rc = upload(g.tiles, tiles.data(), tiles.size() * sizeof(McTile), s)) ||
        (rc = upload(g.items, items.data(), items.size() * sizeof(IntraItem), s)))
        return rc;
    for (int q = 0; q < 4; q++) {
        g.nres[q] = (uint32_t)res[q].size();
        if ((rc = upload(g.res[q], res[q].data(), res[q].size() * sizeof(ResBlock), s))) return rc;
    }
    HIPCHK(hipStreamSynchronize(s));
    return HEVCR_OK;
}

int build_batch(hevcr_engine* e, hevcr_frame* const* frames, const int* slots, int n, int stop, hevcr_batch** out) {
    if (!e || !frames || !slots || !out || n < 1 || n > 65535) return fail(HEVCR_EINVAL, "batch arguments");
    if (stop < HEVCR_RUN_FULL || stop > HEVCR_RUN_STOP_DEBLOCK) return fail(HEVCR_EINVAL, "stop %d", stop);
    std::set<std::pair<int, int>> written;
    for (int i = 0; i < n; i++) {
        hevcr_frame* f = frames[i];
        if (!f) return fail(HEVCR_EINVAL, "null frame %d", i);
        auto it = e->streams.find(f->stream_id);
        if (it == e->streams.end()) return fail(HEVCR_EINVAL, "frame %d: stream %d is gone", i, f->stream_id);
        if (slots[i] < 0 || slots[i] >= it->second->nslots) return fail(HEVCR_EINVAL, "frame %d: slot %d", i, slots[i]);
        if (!written.insert({f->stream_id, slots[i]}).second)
            return fail(HEVCR_EINVAL, "frame %d: slot %d of stream %d written twice", i, slots[i], f->stream_id);
    }
    for (int i = 0; i < n; i++) {
        const hevcr_frame* f = frames[i];
        for (int l = 0; l < 2; l++)
            for (int k = 0; k < f->p.num_ref[l]; k++)
                if (written.count({f->stream_id, f->p.ref_slot[l][k]}))
                    return fail(HEVCR_EINVAL, "frame %d references slot %d, which this batch writes", i, f->p.ref_slot[l][k]);
    }
    auto b = std::make_unique<hevcr_batch>();
    b->frames.assign(frames, frames + n);
    b->slots.assign(slots, slots + n);
    b->stop = stop;
    b->groups[0].bpp = 1;
    b->groups[1].bpp = 2;
    for (int i = 0; i < n; i++) {
        b->groups[frames[i]->bpp == 1 ? 0 : 1].members.push_back(i);
        for (int st = 0; st < HEVCR_NUM_STAGES; st++) b->bytes[st] += frames[i]->bytes[st];
    }
    if (stop != HEVCR_RUN_FULL) b->bytes[HEVCR_STAGE_SAO] = 0;
    if (stop == HEVCR_RUN_STOP_RECON) b->bytes[HEVCR_STAGE_DEBLOCK] = 0;
    for (auto& g : b->groups)
        if (!g.members.empty()) {
            int rc = build_group(e, b.get(), g);
            if (rc) return rc;
        }
    *out = b.release();
    return HEVCR_OK;
}

template <typename T>
void launch_stage(hevcr_engine* e, BatchGroup& g, int stage) {
    hipStream_t s = e->stream;
    const FrameDev* fr = g.fdev.as<FrameDev>();
    const int n = (int)g.members.size();
    switch (stage) {
        case HEVCR_STAGE_EXPAND: launch_expand(s, fr, g.cu_prefix.as<uint32_t>(), n, g.total_cus); break;
        case HEVCR_STAGE_MC: launch_mc<T>(s, fr, g.tiles.as<McTile>(), g.ntiles); break;
        case HEVCR_STAGE_RESIDUAL: {
            const ResBlock* lists[4] = {g.res[0].as<ResBlock>(), g.res[1].as<ResBlock>(), g.res[2].as<ResBlock>(),
                                        g.res[3].as<ResBlock>()};
            launch_residual<T>(s, fr, lists, g.nres);
            break;
        }
        case HEVCR_STAGE_DEBLOCK: launch_deblock<T>(s, fr, g.tile_prefix.as<uint32_t>(), n, g.total_tiles); break;
        case HEVCR_STAGE_SAO: launch_sao<T>(s, fr, g.ctb_prefix.as<uint32_t>(), n, g.total_ctbs); break;
        default: break;
    }
}

template <typename T>
void launch_intra_diag(hevcr_engine* e, BatchGroup& g, size_t dgi) {
    if (dgi + 1 >= g.diag_start.size()) return;
    const uint32_t a = g.diag_start[dgi], z = g.diag_start[dgi + 1];
    if (z > a) launch_intra<T>(e->stream, g.fdev.as<FrameDev>(), g.items.as<IntraItem>() + a, z - a);
}

int launch_batch(hevcr_engine* e, hevcr_batch* b, bool timed) {
    for (int stage = 0; stage < HEVCR_NUM_STAGES; stage++) {
        if (stage == HEVCR_STAGE_SAO && b->stop != HEVCR_RUN_FULL) break;
        if (stage == HEVCR_STAGE_DEBLOCK && b->stop == HEVCR_RUN_STOP_RECON) break;
        bool any = false;
        for (auto& g : b->groups) {
            if (g.members.empty()) continue;
            switch (stage) {
                case HEVCR_STAGE_EXPAND: any |= g.total_cus > 0; break;
                case HEVCR_STAGE_MC: any |= g.ntiles > 0; break;
                case HEVCR_STAGE_RESIDUAL: any |= (g.nres[0] | g.nres[1] | g.nres[2] | g.nres[3]) > 0; break;
                case HEVCR_STAGE_INTRA: any |= g.diag_start.back() > 0; break;
                case HEVCR_STAGE_DEBLOCK: any |= g.total_tiles > 0; break;
                case HEVCR_STAGE_SAO: any |= g.total_ctbs > 0; break;
            }
        }
        if (!any) continue;
        StageTimer t(e, stage, timed);
        if (stage == HEVCR_STAGE_INTRA) {
            size_t nd = 0;
            for (auto& g : b->groups) nd = std::max(nd, g.diag_start.size());
            for (size_t dgi = 0; dgi + 1 < nd; dgi++) {
                if (!b->groups[0].members.empty()) launch_intra_diag<uint8_t>(e, b->groups[0], dgi);
                if (!b->groups[1].members.empty()) launch_intra_diag<uint16_t>(e, b->groups[1], dgi);
            }
        } else {
            if (!b->groups[0].members.empty()) launch_stage<uint8_t>(e, b->groups[0], stage);
            if (!b->groups[1].members.empty()) launch_stage<uint16_t>(e, b->groups[1], stage);
        }
    }
    return last_launch_error();
}

int run_batch(hevcr_engine* e, hevcr_batch* b, bool graph) {
    HIPCHK(hipSetDevice(e->device));
    for (size_t i = 0; i < b->frames.size(); i++)
        if (!e->streams.count(b->frames[i]->stream_id)) return fail(HEVCR_EINVAL, "a stream of this batch is gone");
    if (!graph) return launch_batch(e, b, e->profiling);
    if (!b->gexec) {
        HIPCHK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
        const int rc = launch_batch(e, b, false);
        hipGraph_t gr = nullptr;
        const hipError_t ec = hipStreamEndCapture(e->stream, &gr);
        if (rc) {
            if (gr) (void)hipGraphDestroy(gr);
            return rc;
        }
        if (ec != hipSuccess) return fail(HEVCR_EDEVICE, "hipStreamEndCapture: %s", hipGetErrorString(ec));
        b->graph = gr;
        HIPCHK(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
    }
    StageTimer t(e, HEVCR_STAGE_EXPAND, false);
    HIPCHK(hipGraphLaunch(b->gexec, e->stream));
    return HEVCR_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------

extern "C" {

const char* hevcr_last_error(void) { return g_err.c_str(); }

int hevcr_engine_create(int device, hevcr_engine** out) {
    if (!out) return fail(HEVCR_EINVAL, "null out");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(HEVCR_EDEVICE, "no GPU visible");
    if (device < 0 || device >= n) return fail(HEVCR_EINVAL, "device %d of %d", device, n);
    HIPCHK(hipSetDevice(device));
    auto* e = new (std::nothrow) hevcr_engine();
    if (!e) return fail(HEVCR_ENOMEM, "engine allocation");
    e->device = device;
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) {
        delete e;
        return fail(HEVCR_EDEVICE, "hipStreamCreate failed");
    }
    *out = e;
    return HEVCR_OK;
}

void hevcr_engine_destroy(hevcr_engine* e) {
    if (!e) return;
    (void)hipSetDevice(e->device);
    (void)hipStreamSynchronize(e->stream);
    collect(e);
    for (auto ev : e->free_events) (void)hipEventDestroy(ev);
    e->streams.clear();
    (void)hipStreamDestroy(e->stream);
    delete e;
}

int hevcr_engine_sync(hevcr_engine* e) {
    if (!e) return fail(HEVCR_EINVAL, "null engine");
    HIPCHK(hipSetDevice(e->device));
    HIPCHK(hipStreamSynchronize(e->stream));
    return HEVCR_OK;
}

int hevcr_stream_create(hevcr_engine* e, int width, int height, int bit_depth, int dpb_slots, int* stream_id) {
    if (!e || !stream_id) return fail(HEVCR_EINVAL, "null argument");
    if (width < 8 || height < 8 || (width & 7) || (height & 7) || width > 8192 || height > 4352)
        return fail(HEVCR_EINVAL, "picture size %dx%d", width, height);
    if (bit_depth < 8 || bit_depth > 10) return fail(HEVCR_EUNSUPPORTED, "bit depth %d", bit_depth);
    if (dpb_slots < 1 || dpb_slots > 128) return fail(HEVCR_EINVAL, "dpb slots %d", dpb_slots);
    HIPCHK(hipSetDevice(e->device));
    auto st = std::make_unique<Stream>();
    st->w = width;
    st->h = height;
    st->bd = bit_depth;
    st->bpp = bit_depth > 8 ? 2 : 1;
    st->nslots = dpb_slots;
    st->stride[0] = (width + 63) & ~63;
    st->stride[1] = st->stride[2] = (width / 2 + 63) & ~63;
    st->plane_bytes[0] = (size_t)st->stride[0] * height * st->bpp;
    st->plane_bytes[1] = st->plane_bytes[2] = (size_t)st->stride[1] * (height / 2) * st->bpp;
    for (int i = 0; i < dpb_slots; i++) {
        st->slots.push_back(std::make_unique<Picture>());
        int rc = alloc_picture(*st, *st->slots.back());
        if (rc) return rc;
        HIPCHK(hipMemsetAsync(st->slots.back()->mem.p, 0, st->slots.back()->mem.cap, e->stream));
    }
    HIPCHK(hipStreamSynchronize(e->stream));
    *stream_id = e->next_stream++;
    e->streams[*stream_id] = std::move(st);
    return HEVCR_OK;
}

int hevcr_stream_destroy(hevcr_engine* e, int stream_id) {
    if (!e) return fail(HEVCR_EINVAL, "null engine");
    auto it = e->streams.find(stream_id);
    if (it == e->streams.end()) return fail(HEVCR_EINVAL, "stream %d", stream_id);
    HIPCHK(hipSetDevice(e->device));
    HIPCHK(hipStreamSynchronize(e->stream));
    e->streams.erase(it);
    return HEVCR_OK;
}

static int picture_copy(hevcr_engine* e, int stream_id, int slot, uint16_t* const* host, bool up) {
    Stream* st = e ? find_stream(e, stream_id) : nullptr;
    if (!st) return fail(HEVCR_EINVAL, "stream %d", stream_id);
    if (slot < 0 || slot >= st->nslots) return fail(HEVCR_EINVAL, "slot %d", slot);
    for (int c = 0; c < 3; c++)
        if (!host[c]) return fail(HEVCR_EINVAL, "null plane");
    HIPCHK(hipSetDevice(e->device));
    Picture& pic = *st->slots[(size_t)slot];
    for (int c = 0; c < 3; c++) {
        const int w = c ? st->w / 2 : st->w, h = c ? st->h / 2 : st->h;
        std::vector<uint8_t> tmp((size_t)st->stride[c] * h * st->bpp, 0);
        if (up) {
            const int maxv = (1 << st->bd) - 1;
            for (int y = 0; y < h; y++)
                for (int x = 0; x < w; x++) {
                    const uint16_t v = host[c][(size_t)y * w + x];
                    if (v > maxv) return fail(HEVCR_EINVAL, "sample above the bit depth");
                    if (st->bpp == 1)
                        tmp[(size_t)y * st->stride[c] + x] = (uint8_t)v;
                    else
                        reinterpret_cast<uint16_t*>(tmp.data())[(size_t)y * st->stride[c] + x] = v;
                }
            HIPCHK(hipMemcpyAsync(pic.plane[c], tmp.data(), tmp.size(), hipMemcpyHostToDevice, e->stream));
            HIPCHK(hipStreamSynchronize(e->stream));
        } else {
            HIPCHK(hipMemcpyAsync(tmp.data(), pic.plane[c], tmp.size(), hipMemcpyDeviceToHost, e->stream));
            HIPCHK(hipStreamSynchronize(e->stream));
            for (int y = 0; y < h; y++)
                for (int x = 0; x < w; x++)
                    host[c][(size_t)y * w + x] = st->bpp == 1 ? tmp[(size_t)y * st->stride[c] + x]
                                                              : reinterpret_cast<uint16_t*>(tmp.data())[(size_t)y * st->stride[c] + x];
        }
    }
    return HEVCR_OK;
}

int hevcr_picture_upload(hevcr_engine* e, int stream_id, int slot, const uint16_t* y, const uint16_t* cb,
                         const uint16_t* cr) {
    uint16_t* h[3] = {const_cast<uint16_t*>(y), const_cast<uint16_t*>(cb), const_cast<uint16_t*>(cr)};
    return picture_copy(e, stream_id, slot, h, true);
}

int hevcr_picture_download(hevcr_engine* e, int stream_id, int slot, uint16_t* y, uint16_t* cb, uint16_t* cr) {
    uint16_t* h[3] = {y, cb, cr};
    return picture_copy(e, stream_id, slot, h, false);
}

int hevcr_frame_prepare(hevcr_engine* e, int stream_id, const hevcr_frame_syntax* syn, hevcr_frame** out) {
    if (!e || !syn || !out) return fail(HEVCR_EINVAL, "null argument");
    *out = nullptr;
    Stream* st = find_stream(e, stream_id);
    if (!st) return fail(HEVCR_EINVAL, "stream %d", stream_id);
    int rc = validate(*st, syn);
    if (rc) return rc;
    HIPCHK(hipSetDevice(e->device));
    auto f = std::make_unique<hevcr_frame>();
    if ((rc = prepare_frame(e, *st, stream_id, syn, f.get()))) return rc;
    *out = f.release();
    return HEVCR_OK;
}

int hevcr_frame_run(hevcr_engine* e, hevcr_frame* f, int out_slot, int stop) {
    if (!e || !f) return fail(HEVCR_EINVAL, "null argument");
    Stream* st = find_stream(e, f->stream_id);
    if (!st) return fail(HEVCR_EINVAL, "stream %d is gone", f->stream_id);
    if (out_slot < 0 || out_slot >= st->nslots) return fail(HEVCR_EINVAL, "slot %d", out_slot);
    if (stop < HEVCR_RUN_FULL || stop > HEVCR_RUN_STOP_DEBLOCK) return fail(HEVCR_EINVAL, "stop %d", stop);
    for (int l = 0; l < 2; l++)
        for (int i = 0; i < f->p.num_ref[l]; i++)
            if (f->p.ref_slot[l][i] == out_slot) return fail(HEVCR_EINVAL, "reference slot %d is the output slot", out_slot);
    HIPCHK(hipSetDevice(e->device));
    if (!f->plain || f->plain_slot != out_slot || f->plain_stop != stop) {
        delete f->plain;
        f->plain = nullptr;
        hevcr_frame* one[1] = {f};
        int rc = build_batch(e, one, &out_slot, 1, stop, &f->plain);
        if (rc) return rc;
        f->plain_slot = out_slot;
        f->plain_stop = stop;
    }
    return run_batch(e, f->plain, false);
}

void hevcr_frame_release(hevcr_engine* e, hevcr_frame* f) {
    if (!f) return;
    if (e) {
        (void)hipSetDevice(e->device);
        (void)hipStreamSynchronize(e->stream);
    }
    delete f->plain;
    delete f;
}

int hevcr_frame_bytes(const hevcr_frame* f, double* bytes) {
    if (!f || !bytes) return fail(HEVCR_EINVAL, "null argument");
    for (int i = 0; i < HEVCR_NUM_STAGES; i++) bytes[i] = f->bytes[i];
    if (!f->sao_active) bytes[HEVCR_STAGE_SAO] = 0;
    return HEVCR_OK;
}

int hevcr_decode_frame(hevcr_engine* e, int stream_id, const hevcr_frame_syntax* syn, int out_slot) {
    hevcr_frame* f = nullptr;
    int rc = hevcr_frame_prepare(e, stream_id, syn, &f);
    if (rc) return rc;
    rc = hevcr_frame_run(e, f, out_slot, HEVCR_RUN_FULL);
    hevcr_frame_release(e, f);
    return rc;
}

int hevcr_batch_create(hevcr_engine* e, hevcr_frame* const* frames, const int* out_slots, int n, int stop,
                       hevcr_batch** out) {
    if (out) *out = nullptr;
    if (!e) return fail(HEVCR_EINVAL, "null engine");
    HIPCHK(hipSetDevice(e->device));
    return build_batch(e, frames, out_slots, n, stop, out);
}

int hevcr_batch_run(hevcr_engine* e, hevcr_batch* b, int use_graph) {
    if (!e || !b) return fail(HEVCR_EINVAL, "null argument");
    return run_batch(e, b, use_graph != 0);
}

void hevcr_batch_release(hevcr_engine* e, hevcr_batch* b) {
    if (!b) return;
    if (e) {
        (void)hipSetDevice(e->device);
        (void)hipStreamSynchronize(e->stream);
    }
    delete b;
}

int hevcr_batch_bytes(const hevcr_batch* b, double* bytes) {
    if (!b || !bytes) return fail(HEVCR_EINVAL, "null argument");
    for (int i = 0; i < HEVCR_NUM_STAGES; i++) bytes[i] = b->bytes[i];
    return HEVCR_OK;
}

int hevcr_set_profiling(hevcr_engine* e, int on) {
    if (!e) return fail(HEVCR_EINVAL, "null engine");
    e->profiling = on != 0;
    return HEVCR_OK;
}

int hevcr_stage_times(hevcr_engine* e, double* ms, uint64_t* launches) {
    if (!e) return fail(HEVCR_EINVAL, "null engine");
    HIPCHK(hipSetDevice(e->device));
    int rc = collect(e);
    if (rc) return rc;
    for (int i = 0; i < HEVCR_NUM_STAGES; i++) {
        if (ms) ms[i] = e->stage_ms[i];
        if (launches) launches[i] = e->stage_n[i];
    }
    return HEVCR_OK;
}

int hevcr_reset_stage_times(hevcr_engine* e) {
    if (!e) return fail(HEVCR_EINVAL, "null engine");
    HIPCHK(hipSetDevice(e->device));
    int rc = collect(e);
    std::memset(e->stage_ms, 0, sizeof(e->stage_ms));
    std::memset(e->stage_n, 0, sizeof(e->stage_n));
    return rc;
}

int hevcr_residual_blocks(hevcr_engine* e, int bit_depth, uint32_t n, const uint8_t* log2, const uint8_t* cidx,
                          const uint8_t* qp, const uint8_t* flags, const uint32_t* off, const int16_t* coeffs,
                          uint64_t num_coeffs, int16_t* out) {
    if (!e || (n && (!log2 || !cidx || !qp || !flags || !off || !coeffs || !out)))
        return fail(HEVCR_EINVAL, "null argument");
    if (bit_depth < 8 || bit_depth > 10) return fail(HEVCR_EUNSUPPORTED, "bit depth %d", bit_depth);
    // Blocks are laid out one below the other in a 32-sample-wide residual plane.
    std::vector<ResBlock> lists[4];
    std::vector<uint32_t> row(n);
    uint32_t rows = 0;
    for (uint32_t i = 0; i < n; i++) {
        if (log2[i] < 2 || log2[i] > 5 || cidx[i] > 2 || qp[i] > 51 + 6 * (bit_depth - 8) || (flags[i] & ~3u))
            return fail(HEVCR_EINVAL, "block %u parameters", i);
        if ((flags[i] & 2) && log2[i] != 2) return fail(HEVCR_EUNSUPPORTED, "block %u transform skip above 4x4", i);
        if ((uint64_t)off[i] + (1u << (2 * log2[i])) > num_coeffs) return fail(HEVCR_EINVAL, "block %u range", i);
        row[i] = rows;
        rows += 1u << log2[i];
        if (rows > 65535) return fail(HEVCR_EINVAL, "too many blocks in one call");
        ResBlock rb;
        rb.coeff_off = off[i];
        rb.x = 0;
        rb.y = (uint16_t)row[i];
        rb.log2 = log2[i];
        rb.cidx = 0;  // all blocks go to the single residual plane
        rb.qp = qp[i];
        rb.flags = (uint8_t)(((flags[i] & 1) ? kResIntra : 0) | ((flags[i] & 1) && cidx[i] == 0 && log2[i] == 2 ? kResDst : 0) |
                             ((flags[i] & 2) ? kResTskip : 0));
        rb.frame = 0;
        rb.pad = 0;
        lists[log2[i] - 2].push_back(rb);
    }
    HIPCHK(hipSetDevice(e->device));
    hipStream_t s = e->stream;
    int rc;
    if ((rc = upload(e->kat_coeffs, coeffs, num_coeffs * 2, s)) || (rc = e->kat_out.reserve((size_t)std::max(rows, 1u) * 32 * 2)))
        return rc;
    size_t total = 0;
    for (auto& l : lists) total += l.size();
    if ((rc = e->kat_blocks.reserve(std::max<size_t>(total, 1) * sizeof(ResBlock)))) return rc;
    FrameDev d;
    std::memset(&d, 0, sizeof(d));
    d.fc.bd = bit_depth;
    d.fc.maxv = (1 << bit_depth) - 1;
    d.fc.stride[0] = 32;
    d.coeffs = e->kat_coeffs.as<int16_t>();
    d.res[0] = e->kat_out.as<int16_t>();
    if ((rc = upload(e->kat_fdev, &d, sizeof(d), s))) return rc;
    const ResBlock* ptr[4];
    uint32_t cnt[4];
    size_t at = 0;
    for (int k = 0; k < 4; k++) {
        ptr[k] = e->kat_blocks.as<ResBlock>() + at;
        cnt[k] = (uint32_t)lists[k].size();
        if (cnt[k])
            HIPCHK(hipMemcpyAsync(const_cast<ResBlock*>(ptr[k]), lists[k].data(), cnt[k] * sizeof(ResBlock),
                                  hipMemcpyHostToDevice, s));
        at += cnt[k];
    }
    launch_residual<uint16_t>(s, e->kat_fdev.as<FrameDev>(), ptr, cnt);
    HIPCHK(hipGetLastError());
    std::vector<int16_t> plane((size_t)rows * 32);
    HIPCHK(hipMemcpyAsync(plane.data(), e->kat_out.p, plane.size() * 2, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    for (uint32_t i = 0; i < n; i++) {
        const int nn = 1 << log2[i];
        for (int y = 0; y < nn; y++)
            for (int x = 0; x < nn; x++) out[off[i] + (size_t)y * nn + x] = plane[(size_t)(row[i] + y) * 32 + x];
    }
    return HEVCR_OK;
}

int hevcr_bench_copy(hevcr_engine* e, uint64_t bytes, int reps) {
    if (!e || bytes < 16 || (bytes & 15) || reps < 1)
        return fail(HEVCR_EINVAL, "bytes %llu reps %d", (unsigned long long)bytes, reps);
    HIPCHK(hipSetDevice(e->device));
    DevBuf a, b;
    int rc;
    if ((rc = a.reserve(bytes)) || (rc = b.reserve(bytes))) return rc;
    HIPCHK(hipMemsetAsync(a.p, 1, bytes, e->stream));
    for (int i = 0; i < reps; i++) launch_copy(e->stream, a.p, b.p, bytes);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(e->stream));
    return HEVCR_OK;
}

}  // extern "C"
