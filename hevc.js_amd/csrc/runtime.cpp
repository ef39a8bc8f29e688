// Host runtime of the reconstruction engine: device DPB per stream, syntax
// validation, derived work lists kept in device memory per prepared frame,
// batched stage launches (direct or replayed from a captured hipGraph) and
// stage timing. Exposes the C-ABI of include/hevcr.h.
#include "runtime.hpp"

#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <set>
#include <utility>

namespace hevcr {

namespace {
thread_local std::string g_err;
}

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

const char* last_error() { return g_err.c_str(); }

int DevBuf::reserve(size_t bytes) {
    if (bytes <= cap) return HEVCR_OK;
    release();
    const size_t want = std::max<size_t>(bytes, 256);
    HIPCHK(hipMalloc(&p, want));
    cap = want;
    return HEVCR_OK;
}

int PinnedBuf::reserve(size_t bytes) {
    if (bytes <= cap) return HEVCR_OK;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
    const size_t want = std::max<size_t>(bytes, 4096);
    HIPCHK(hipHostMalloc(&p, want, hipHostMallocDefault));
    cap = want;
    return HEVCR_OK;
}

BlobPool::~BlobPool() {
    if (all.empty()) return;
    (void)hipSetDevice(device);
    for (void* q : all) (void)hipFree(q);
}

int BlobPool::get(size_t bytes, void** out, size_t* cap) {
    bytes = std::max<size_t>(bytes, 256);
    {
        std::lock_guard<std::mutex> g(m);
        auto it = free.lower_bound(bytes);
        if (it != free.end() && it->first <= bytes + bytes / 2) {  // a block at most 1.5x the size
            *cap = it->first;
            *out = it->second;
            free.erase(it);
            return HEVCR_OK;
        }
    }
    // a new block with some room, so that frames of similar size share blocks
    const size_t want = ((bytes + bytes / 4) + 65535) & ~(size_t)65535;
    void* q = nullptr;
    HIPCHK(hipMalloc(&q, want));
    std::lock_guard<std::mutex> g(m);
    all.push_back(q);
    *out = q;
    *cap = want;
    return HEVCR_OK;
}

void BlobPool::put(void* q, size_t cap) {
    std::lock_guard<std::mutex> g(m);
    free.emplace(cap, q);
}

UploadCtx::~UploadCtx() {
    if (s) (void)hipStreamSynchronize(s);
    for (int k = 0; k < 2; k++) {
        if (ev[k]) (void)hipEventDestroy(ev[k]);
        if (pinned[k]) (void)hipHostFree(pinned[k]);
    }
    if (s) (void)hipStreamDestroy(s);
}

int UploadCtx::init(int dev) {
    device = dev;
    HIPCHK(hipSetDevice(dev));
    HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    for (int k = 0; k < 2; k++)  // waiting threads sleep
        HIPCHK(hipEventCreateWithFlags(&ev[k], hipEventDisableTiming | hipEventBlockingSync));
    return HEVCR_OK;
}

// Staging buffer k with room for `bytes`; the copy out of it issued two
// frames earlier has finished when this returns.
int UploadCtx::reserve(int k, size_t bytes) {
    HIPCHK(hipEventSynchronize(ev[k]));
    if (bytes <= cap[k]) return HEVCR_OK;
    if (pinned[k]) (void)hipHostFree(pinned[k]);
    pinned[k] = nullptr;
    cap[k] = 0;
    const size_t want = std::max<size_t>(bytes + bytes / 2, 1 << 20);
    HIPCHK(hipHostMalloc(&pinned[k], want, hipHostMallocDefault));
    cap[k] = want;
    return HEVCR_OK;
}

double now_seconds() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

Stream* find_stream(hevcr_engine* e, int id) {
    auto it = e->streams.find(id);
    return it == e->streams.end() ? nullptr : it->second.get();
}

}  // namespace hevcr

using namespace hevcr;

hevcr_frame::~hevcr_frame() {
    delete plain;
    if (upload_ev) {  // the blob goes back to the pool only once the upload into it has finished
        (void)hipEventSynchronize(upload_ev);
        (void)hipEventDestroy(upload_ev);
    }
}

hevcr_batch::~hevcr_batch() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    if (staged) (void)hipEventDestroy(staged);
}

namespace {

inline size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

int last_launch_error() {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(HEVCR_EDEVICE, "kernel launch: %s", hipGetErrorString(e));
    return HEVCR_OK;
}

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

// Timer slots: the stages, then the kernels (kKernel + HEVCR_KERNEL_*).
constexpr int kKernel = HEVCR_NUM_STAGES;

int collect(hevcr_engine* e) {
    for (auto& p : e->pending) {
        HIPCHK(hipEventSynchronize(p.b));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, p.a, p.b));
        e->timer_ms[p.stage] += ms;
        e->timer_n[p.stage] += 1;
        e->free_events.push_back(p.a);
        e->free_events.push_back(p.b);
    }
    e->pending.clear();
    return HEVCR_OK;
}

// -- syntax validation --------------------------------------------------------

uint32_t spread_bits(uint32_t v) {  // bit i of v (i < 8) to bit 2i
    v &= 0xffu;
    v = (v | (v << 4)) & 0x0f0fu;
    v = (v | (v << 2)) & 0x3333u;
    return (v | (v << 1)) & 0x5555u;
}
uint32_t zorder(int x, int y) {  // 4x4 units inside a CTB: bits of x and y interleaved, x in the even bits
    return spread_bits((uint32_t)x) | (spread_bits((uint32_t)y) << 1);
}

// Slice segments and tiles of a picture, per CTB raster address: the CTB
// scan (CtbAddrRsToTs, pps.cpp derive_tile_scan), each CTB's tile and slice
// segment, and each segment's parameters (a picture without segment records
// is one segment with p's parameters, one without tiles one tile).
struct PicLayout {
    int wctb = 0, hctb = 0, nctb = 0;
    bool multi = false;               // several slice segments or several tiles
    std::vector<uint32_t> rs2ts;      // empty: raster scan
    std::vector<uint16_t> tile;       // empty: one tile
    const uint8_t* seg = nullptr;     // null: every CTB in segment 0
    std::vector<hevcr_slice> segs;
    int lf_tiles = 1;                 // loop_filter_across_tiles_enabled_flag
    int seg_of(int a) const { return seg ? seg[a] : 0; }
    int tile_of(int a) const { return tile.empty() ? 0 : tile[a]; }
    uint32_t ts(int a) const { return rs2ts.empty() ? (uint32_t)a : rs2ts[(size_t)a]; }
};

hevcr_slice slice_of_params(const hevcr_frame_params& p) {
    hevcr_slice q;
    std::memset(&q, 0, sizeof(q));
    q.slice_type = p.slice_type;
    q.slice_cb_qp_offset = p.slice_cb_qp_offset;
    q.slice_cr_qp_offset = p.slice_cr_qp_offset;
    q.deblocking_disabled = p.deblocking_disabled;
    q.beta_offset_div2 = p.beta_offset_div2;
    q.tc_offset_div2 = p.tc_offset_div2;
    q.loop_filter_across_slices = 1;
    q.weighted_pred = p.weighted_pred;
    q.luma_log2_wd = p.luma_log2_wd;
    q.chroma_log2_wd = p.chroma_log2_wd;
    std::memcpy(q.weights, p.weights, sizeof(q.weights));
    return q;
}

// Fills L from the frame's tile and slice-segment records, checking their
// shape: tile columns / rows that add up to the picture, segments that cover
// consecutive runs of the CTB scan in order, starting with segment 0.
int build_layout(const hevcr_frame_syntax* syn, PicLayout& L) {
    const hevcr_frame_params& p = syn->p;
    const int ctb = 1 << p.ctb_log2;
    L.wctb = (p.width + ctb - 1) / ctb;
    L.hctb = (p.height + ctb - 1) / ctb;
    L.nctb = L.wctb * L.hctb;
    L.rs2ts.clear();
    L.tile.clear();
    L.lf_tiles = 1;
    const hevcr_tiles* t = syn->tiles;
    if (t && (t->cols > 1 || t->rows > 1)) {
        if (t->cols < 1 || t->cols > HEVCR_MAX_TILE_COLS || t->rows < 1 || t->rows > HEVCR_MAX_TILE_ROWS)
            return fail(HEVCR_EINVAL, "tiles %dx%d", t->cols, t->rows);
        int sw = 0, sh = 0;
        for (int i = 0; i < t->cols; i++) {
            if (!t->col_width[i]) return fail(HEVCR_EINVAL, "tile column %d has no CTB", i);
            sw += t->col_width[i];
        }
        for (int i = 0; i < t->rows; i++) {
            if (!t->row_height[i]) return fail(HEVCR_EINVAL, "tile row %d has no CTB", i);
            sh += t->row_height[i];
        }
        if (sw != L.wctb || sh != L.hctb) return fail(HEVCR_EINVAL, "tile columns / rows do not cover the picture");
        L.tile.resize((size_t)L.nctb);
        L.rs2ts.resize((size_t)L.nctb);
        uint32_t ts = 0;
        int y0 = 0;
        for (int r = 0; r < t->rows; r++) {
            int x0 = 0;
            for (int c = 0; c < t->cols; c++) {
                for (int y = y0; y < y0 + t->row_height[r]; y++)
                    for (int x = x0; x < x0 + t->col_width[c]; x++) {
                        L.tile[(size_t)(y * L.wctb + x)] = (uint16_t)(r * t->cols + c);
                        L.rs2ts[(size_t)(y * L.wctb + x)] = ts++;
                    }
                x0 += t->col_width[c];
            }
            y0 += t->row_height[r];
        }
        L.lf_tiles = t->loop_filter_across_tiles != 0;
    }
    L.segs.clear();
    L.seg = nullptr;
    if (syn->slices) {
        if (syn->num_slices < 1 || syn->num_slices > HEVCR_MAX_SLICES || !syn->ctb_slice)
            return fail(HEVCR_EINVAL, "%u slice segments", syn->num_slices);
        L.seg = syn->ctb_slice;
        L.segs.assign(syn->slices, syn->slices + syn->num_slices);
        // along the CTB scan: 0 first, then each segment's run in order, all of them
        thread_local std::vector<uint8_t> tls_by_ts;
        std::vector<uint8_t>& by_ts = tls_by_ts;
        by_ts.resize((size_t)L.nctb);
        for (int a = 0; a < L.nctb; a++) by_ts[L.ts(a)] = L.seg[a];
        if (by_ts[0] != 0) return fail(HEVCR_EINVAL, "the first CTB is not in slice segment 0");
        for (int i = 1; i < L.nctb; i++)
            if (by_ts[(size_t)i] != by_ts[(size_t)i - 1] && by_ts[(size_t)i] != by_ts[(size_t)i - 1] + 1)
                return fail(HEVCR_EINVAL, "slice segments out of order in the CTB scan");
        if (by_ts[(size_t)L.nctb - 1] + 1u != syn->num_slices) return fail(HEVCR_EINVAL, "slice segments without CTBs");
        if (L.segs[0].slice_type != p.slice_type)
            return fail(HEVCR_EINVAL, "slice segment 0 is not of the picture's slice type");
    } else {
        L.segs.assign(1, slice_of_params(p));
    }
    L.multi = L.segs.size() > 1 || !L.tile.empty();
    return HEVCR_OK;
}

// Decoding-order key of a block: CTB scan address, then z-order of 4x4 units.
uint64_t dec_key(const PicLayout& L, const hevcr_frame_params& p, int x, int y) {
    const uint64_t ctb = L.ts((y >> p.ctb_log2) * L.wctb + (x >> p.ctb_log2));
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

// Factor m(x, y) of every block size and matrixId, as perform_dequant picks it
// (transform.cpp:308-356): the 4x4 list, or the 8x8 base list replicated over
// r x r squares (r = N / 8) with the DC factor at (0, 0) from 16x16 up, where a
// stored 0 reads as 16.
void scaling_factors(const hevcr_scaling& sl, uint8_t* out) {
    for (int q = 0; q < 4; q++) {
        const int n = 4 << q, r = n / 8;
        for (int k = 0; k < 6; k++)
            for (int y = 0; y < n; y++)
                for (int x = 0; x < n; x++) {
                    int m;
                    if (q == 0)
                        m = sl.list[0][k][y * 4 + x];
                    else if (q >= 2 && x == 0 && y == 0)
                        m = sl.dc[q - 2][k] ? sl.dc[q - 2][k] : 16;
                    else
                        m = sl.list[q][k][(y / r) * 8 + x / r];
                    out[kScaleOff[q] + k * n * n + y * n + x] = (uint8_t)m;
                }
    }
}

int chroma_qp_prime(const hevcr_frame_params& p, const hevcr_slice& q, int qpy, int c);

// A frame's coded residual blocks by size and plane, with their algorithmic
// bytes. Validation emits the blocks of inter CUs as it checks their
// transform units (one pass over the records instead of two); preparation
// adds those of intra and PCM CUs. Each list has room for one block per
// transform unit (a plane has at most one coded block per unit, and a PCM
// CU one record), so blocks are written without a capacity check; `n` holds
// how many are filled.
struct ResLists {
    std::vector<ResBlock> blocks[4][3];  // [log2 size - 2][plane]
    uint32_t n[4][3] = {};
    // every transform unit as the expansion kernel reads it: x / 4 (bits
    // 0-11), y / 4 (12-23), log2 size - 2 (24-25), coded luma (26)
    std::vector<uint32_t> tuw;
    // per CU (and one past the last): its first transform unit and PU record
    std::vector<uint32_t> first_tu, first_pu;
    bool any_intra = false;
    uint64_t in[4] = {0, 0, 0, 0};  // coefficient bytes read by the emitted inter blocks, per size
    double bytes = 0, kbytes[4] = {0, 0, 0, 0};
    void reset(size_t ntus) {
        for (int q = 0; q < 4; q++)
            for (int c = 0; c < 3; c++) {
                if (blocks[q][c].size() < ntus) blocks[q][c].resize(ntus);
                n[q][c] = 0;
            }
        tuw.resize(ntus);
        any_intra = false;
        bytes = 0;
        for (int q = 0; q < 4; q++) {
            kbytes[q] = 0;
            in[q] = 0;
        }
    }
    // the inter blocks' bytes once they are all emitted: coefficients read,
    // prediction read and written
    void close_inter(int bpp) {
        for (int q = 0; q < 4; q++) {
            const uint64_t cnt = (uint64_t)n[q][0] + n[q][1] + n[q][2];
            kbytes[q] = (double)in[q] + 2.0 * bpp * (double)(cnt << (2 * (q + 2)));
            bytes += kbytes[q];
        }
    }
    void add(const ResBlock& rb, double b) {
        const int q = rb.log2 - 2;
        blocks[q][rb.cidx][n[q][rb.cidx]++] = rb;
        bytes += b;
        kbytes[q] += b;
    }
    uint32_t count(int q) const { return n[q][0] + n[q][1] + n[q][2]; }
};

// out: where the residual blocks of inter CUs go (cleared first), or null.
int validate(const Stream& st, const hevcr_frame_syntax* syn, ResLists* out) {
    const hevcr_frame_params& p = syn->p;
    thread_local PicLayout tls_layout;
    PicLayout& L = tls_layout;
    if (int rc = build_layout(syn, L)) return rc;
    if (p.width != st.w || p.height != st.h || p.bit_depth != st.bd)
        return fail(HEVCR_EINVAL, "frame %dx%d/%d does not match stream %dx%d/%d", p.width, p.height, p.bit_depth, st.w,
                    st.h, st.bd);
    if (p.ctb_log2 < 4 || p.ctb_log2 > 6) return fail(HEVCR_EINVAL, "ctb_log2 %d", p.ctb_log2);
    if (p.min_cb_log2 < 3 || p.min_cb_log2 > p.ctb_log2) return fail(HEVCR_EINVAL, "min_cb_log2 %d", p.min_cb_log2);
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
    for (int v : {p.pps_cb_qp_offset, p.pps_cr_qp_offset})
        if (v < -12 || v > 12) return fail(HEVCR_EINVAL, "chroma qp offset %d", v);
    // per slice segment (p's own fields when it has no segment records)
    for (size_t k = 0; k < L.segs.size(); k++) {
        const hevcr_slice& q = L.segs[k];
        if (q.slice_type < 0 || q.slice_type > 2) return fail(HEVCR_EINVAL, "slice segment %zu type %d", k, q.slice_type);
        const bool wp = q.weighted_pred && q.slice_type != HEVCR_SLICE_I;
        if (wp && (q.luma_log2_wd < 0 || q.luma_log2_wd > 7 || q.chroma_log2_wd < 0 || q.chroma_log2_wd > 7))
            return fail(HEVCR_EINVAL, "weight denominators %d/%d", q.luma_log2_wd, q.chroma_log2_wd);
        if (q.beta_offset_div2 < -6 || q.beta_offset_div2 > 6 || q.tc_offset_div2 < -6 || q.tc_offset_div2 > 6)
            return fail(HEVCR_EINVAL, "deblocking offsets");
        for (int v : {q.slice_cb_qp_offset, q.slice_cr_qp_offset})
            if (v < -12 || v > 12) return fail(HEVCR_EINVAL, "chroma qp offset %d", v);
    }
    if (!syn->sao && p.sao_enabled) return fail(HEVCR_EINVAL, "sao records missing");
    if ((syn->num_cus && !syn->cus) || (syn->num_pus && !syn->pus) || (syn->num_tus && !syn->tus) ||
        (!syn->sig && syn->num_coeffs && !syn->coeffs))
        return fail(HEVCR_EINVAL, "null syntax array");

    const int qbd = 6 * (p.bit_depth - 8);
    // Coverage without bitmaps: an aligned square of 4x4 cells is one
    // contiguous run of z-order indices, so blocks whose keys increase and
    // whose runs do not overlap are disjoint, and disjoint blocks inside an
    // area cover it exactly when their areas add up to it (the picture for
    // CUs, an intra CU for its transform units).
    if (p.width % (1 << p.min_cb_log2) || p.height % (1 << p.min_cb_log2))
        return fail(HEVCR_EINVAL, "picture %dx%d is not a multiple of the minimum CB size", p.width, p.height);
    uint64_t cu_area = 0;
    const uint32_t* const sig = syn->sig;
    const uint64_t num_sig = sig ? syn->num_sig : 0;
    uint64_t prev_key = 0;
    uint32_t pi = 0, ti = 0;
    const int cmask = (1 << p.ctb_log2) - 1;
    if (out) {
        out->reset(syn->num_tus);
        out->first_tu.resize((size_t)syn->num_cus + 1);
        out->first_pu.resize((size_t)syn->num_cus + 1);
    }
    for (uint32_t ci = 0; ci < syn->num_cus; ci++) {
        const hevcr_cu& cu = syn->cus[ci];
        const int s = 1 << cu.log2_size;
        if (cu.log2_size < p.min_cb_log2 || cu.log2_size > p.ctb_log2 || (cu.x & (s - 1)) || (cu.y & (s - 1)) ||
            cu.x + s > p.width || cu.y + s > p.height)
            return fail(HEVCR_EINVAL, "cu %u geometry (%d,%d) log2 %d", ci, cu.x, cu.y, cu.log2_size);
        const uint64_t key = dec_key(L, p, cu.x, cu.y);
        if (ci && key < prev_key) return fail(HEVCR_EINVAL, "cu %u out of decoding order or overlapping", ci);
        prev_key = key + (uint64_t)(s >> 2) * (uint64_t)(s >> 2);  // where its run of cells ends
        cu_area += (uint64_t)s * (uint64_t)s;
        if (cu.qp_y < -qbd || cu.qp_y > 51) return fail(HEVCR_EINVAL, "cu %u qp %d", ci, cu.qp_y);
        if (cu.pred_mode > 2) return fail(HEVCR_EINVAL, "cu %u pred_mode", ci);
        const hevcr_slice& sq = L.segs[(size_t)L.seg_of((cu.y >> p.ctb_log2) * L.wctb + (cu.x >> p.ctb_log2))];
        if (sq.slice_type == HEVCR_SLICE_I && cu.pred_mode != HEVCR_MODE_INTRA)
            return fail(HEVCR_EINVAL, "cu %u inter in an I slice", ci);
        const int snl = sq.slice_type == HEVCR_SLICE_B ? 2 : (sq.slice_type == HEVCR_SLICE_P ? 1 : 0);
        if (cu.flags & ~(HEVCR_CU_PCM | HEVCR_CU_BYPASS)) return fail(HEVCR_EINVAL, "cu %u flags", ci);
        const bool pcm = cu.flags & HEVCR_CU_PCM;
        if (pcm && (cu.pred_mode != HEVCR_MODE_INTRA || cu.part_mode != HEVCR_PART_2Nx2N || cu.log2_size > 5))
            return fail(HEVCR_EINVAL, "cu %u: PCM in a CU other than intra 2Nx2N up to 32x32", ci);
        const bool intra = cu.pred_mode == HEVCR_MODE_INTRA;
        if (out) {
            out->first_tu[ci] = ti;
            out->first_pu[ci] = pi;
            out->any_intra |= intra;
        }
        if (intra) {
            if (cu.part_mode != HEVCR_PART_2Nx2N && !(cu.part_mode == HEVCR_PART_NxN && cu.log2_size == p.min_cb_log2))
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
                        // the segment's slice type decides its lists; their entries are the picture's
                        if (l >= nl || l >= snl || pu.ref_idx[l] < 0 || pu.ref_idx[l] >= p.num_ref[l])
                            return fail(HEVCR_EINVAL, "pu %u ref_idx[%d] = %d", pi, l, pu.ref_idx[l]);
                        if (sq.weighted_pred) {
                            const hevcr_weight& w = sq.weights[l][pu.ref_idx[l]];
                            for (int c = 0; c < 3; c++)
                                if (w.w[c] < -128 || w.w[c] > 255 || w.o[c] < -128 || w.o[c] > 127)
                                    return fail(HEVCR_EINVAL, "pu %u weight out of range", pi);
                        }
                    }
            }
        }
        // transform units of this CU; an inter CU's coded blocks go to `out`
        // as they are checked (QPs once per CU)
        ResLists* const emit = (out && !intra && !pcm) ? out : nullptr;
        uint8_t eqp[3] = {0, 0, 0}, efl = 0;
        if (emit) {
            eqp[0] = (uint8_t)(cu.qp_y + qbd);
            eqp[1] = (uint8_t)chroma_qp_prime(p, sq, cu.qp_y, 1);
            eqp[2] = (uint8_t)chroma_qp_prime(p, sq, cu.qp_y, 2);
            efl = (uint8_t)(kResToPlane | ((cu.flags & HEVCR_CU_BYPASS) ? kResBypass : 0));
        }
        int tcount = 0;
        uint64_t prev_tu = 0, tu_cells = 0;
        if (pcm) {  // one record holding the samples: the CU's size, all three blocks, values 0..maxv
            if (ti >= syn->num_tus || syn->tus[ti].cu_idx != ci ||
                (ti + 1 < syn->num_tus && syn->tus[ti + 1].cu_idx == ci))
                return fail(HEVCR_EINVAL, "pcm cu %u needs exactly one sample record", ci);
            const hevcr_tu& t = syn->tus[ti];
            if (t.x != cu.x || t.y != cu.y || t.log2_size != cu.log2_size || t.cbf != 7 || t.tskip || t.blk_idx)
                return fail(HEVCR_EINVAL, "pcm cu %u sample record", ci);
            const int maxv = (1 << p.bit_depth) - 1;
            for (int c = 0; c < 3; c++) {
                const int bn = c ? s / 2 : s;
                bool bad = false;
                if (!syn->sig) {
                    if ((uint64_t)t.coeff_off[c] + (uint64_t)bn * bn > syn->num_coeffs) continue;  // checked below
                    const int16_t* q = syn->coeffs + t.coeff_off[c];
                    for (int k = 0; k < bn * bn; k++) bad |= q[k] < 0 || q[k] > maxv;
                } else if (t.coeff_off[c] < syn->num_sig && t.coeff_off[c] + 1 + (uint64_t)syn->sig[t.coeff_off[c]] <= syn->num_sig) {
                    const uint32_t k = syn->sig[t.coeff_off[c]];
                    for (uint32_t j = 1; j <= k; j++) {
                        const int v = (int16_t)(syn->sig[t.coeff_off[c] + j] >> 16);
                        bad |= v < 0 || v > maxv;
                    }
                }
                if (bad) return fail(HEVCR_EINVAL, "pcm cu %u sample beyond the bit depth", ci);
            }
        }
        for (; ti < syn->num_tus && syn->tus[ti].cu_idx == ci; ti++, tcount++) {
            const hevcr_tu& t = syn->tus[ti];
            const int n = 1 << t.log2_size;
            // a CU's units in z-scan order (inside the CU's CTB, checked below: the z-index there suffices)
            const uint64_t tkey = zorder((t.x & cmask) >> 2, (t.y & cmask) >> 2);
            if (tcount && tkey < prev_tu) return fail(HEVCR_EINVAL, "tu %u out of z-scan order or overlapping", ti);
            prev_tu = tkey + (uint64_t)(n >> 2) * (uint64_t)(n >> 2);
            tu_cells += (uint64_t)(n >> 2) * (uint64_t)(n >> 2);
            if (cu.pred_mode == HEVCR_MODE_SKIP) return fail(HEVCR_EINVAL, "tu %u in a skip CU", ti);
            if (t.log2_size < 2 || t.log2_size > 5 || (t.x & (n - 1)) || (t.y & (n - 1)) || t.x < cu.x || t.y < cu.y ||
                t.x + n > cu.x + s || t.y + n > cu.y + s || t.blk_idx > 3)
                return fail(HEVCR_EINVAL, "tu %u geometry", ti);
            if (t.log2_size == 2 && (t.xbase != (t.x & ~7) || t.ybase != (t.y & ~7)))
                return fail(HEVCR_EINVAL, "tu %u base position", ti);
            if (t.cbf & ~7u) return fail(HEVCR_EINVAL, "tu %u cbf", ti);
            if (t.tskip & ~t.cbf) return fail(HEVCR_EINVAL, "tu %u transform skip without residual", ti);
            const int nc = t.log2_size > 2 ? n / 2 : 4;
            // the coded blocks (chroma of 4x4 luma blocks sits with the fourth),
            // checked and emitted without a branch per block (coded-block
            // patterns do not predict): an uncoded block's slot is written and
            // not counted
            const unsigned coded = t.cbf & (t.log2_size == 2 && t.blk_idx != 3 ? 1u : 7u);
            uint32_t bad = ((t.tskip & coded & 1u) != 0 && n != 4) | ((t.tskip & coded & 6u) != 0 && nc != 4);
            uint32_t kin[3];  // coefficient bytes each coded block reads
            for (int c = 0; c < 3; c++) {
                const uint32_t bit = (coded >> c) & 1u;
                const uint64_t bb = c ? (uint64_t)(nc * nc) : (uint64_t)(n * n);
                const uint64_t at = t.coeff_off[c];
                if (sig) {  // compact list: a count, then that many words inside the array
                    const bool inside = at < num_sig;
                    const uint32_t k = num_sig ? sig[inside ? at : 0] : 0;
                    bad |= bit & (uint32_t)(!inside | (k > bb) | (at + 1 + k > num_sig));
                    kin[c] = bit ? 4u * (1u + k) : 0u;
                } else {
                    bad |= bit & (uint32_t)(at + bb > syn->num_coeffs);
                    kin[c] = bit ? (uint32_t)(2 * bb) : 0u;
                }
            }
            if (bad) {  // which check failed
                for (int c = 0; c < 3; c++) {
                    if (!((coded >> c) & 1u)) continue;
                    const int bn = c ? nc : n;
                    if ((t.tskip >> c) & 1 && bn != 4)
                        return fail(HEVCR_EUNSUPPORTED, "tu %u transform skip above 4x4", ti);
                    const uint64_t at = t.coeff_off[c];
                    if (!sig) {
                        if (at + (uint64_t)bn * bn > syn->num_coeffs)
                            return fail(HEVCR_EINVAL, "tu %u coefficients out of range", ti);
                    } else if (at >= num_sig) {
                        return fail(HEVCR_EINVAL, "tu %u compact coefficients out of range", ti);
                    } else if (sig[at] > (uint32_t)(bn * bn) || at + 1 + sig[at] > num_sig) {
                        return fail(HEVCR_EINVAL, "tu %u compact coefficient count %u", ti, sig[at]);
                    }
                }
                return fail(HEVCR_EINVAL, "tu %u coded blocks", ti);
            }
            // (the positions of compact lists are not checked here, a word per
            // coefficient being a third of this function's time: the residual
            // kernels take them modulo the block size)
            if (emit) {
                const int l2c = t.log2_size > 2 ? t.log2_size - 1 : 2;
                const int xc = (t.log2_size == 2 ? t.xbase : t.x) / 2, yc = (t.log2_size == 2 ? t.ybase : t.y) / 2;
                for (int c = 0; c < 3; c++) {
                    const int l2 = c ? l2c : t.log2_size;
                    uint32_t& cnt = emit->n[l2 - 2][c];
                    ResBlock& rb = emit->blocks[l2 - 2][c][cnt];
                    rb.coeff_off = t.coeff_off[c];
                    rb.x = (uint16_t)(c ? xc : t.x);
                    rb.y = (uint16_t)(c ? yc : t.y);
                    rb.log2 = (uint8_t)l2;
                    rb.cidx = (uint8_t)c;
                    rb.qp = eqp[c];
                    rb.flags = (uint8_t)(efl | (((t.tskip >> c) & 1) ? kResTskip : 0));
                    cnt += (coded >> c) & 1u;
                    emit->in[l2 - 2] += kin[c];
                }
            }
            if (out)
                out->tuw[ti] = (uint32_t)(t.x >> 2) | (uint32_t)(t.y >> 2) << 12 | (uint32_t)(t.log2_size - 2) << 24 |
                               (uint32_t)(t.cbf & HEVCR_CBF_Y) << 26;
        }
        if (ti < syn->num_tus && syn->tus[ti].cu_idx < ci) return fail(HEVCR_EINVAL, "tu %u out of order", ti);
        if (intra && tu_cells != (uint64_t)(s >> 2) * (uint64_t)(s >> 2))  // fully covered by their transform tree
            return fail(HEVCR_EINVAL, "intra cu %u not covered by TUs", ci);
        (void)tcount;
    }
    if (out) {
        out->close_inter(st.bpp);
        out->first_tu[syn->num_cus] = ti;
        out->first_pu[syn->num_cus] = pi;
    }
    if (pi != syn->num_pus) return fail(HEVCR_EINVAL, "%u PU records left over", syn->num_pus - pi);
    if (ti != syn->num_tus) return fail(HEVCR_EINVAL, "%u TU records left over", syn->num_tus - ti);
    if (cu_area != (uint64_t)p.width * (uint64_t)p.height) return fail(HEVCR_EINVAL, "picture not covered by CUs");
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

int chroma_qp_prime(const hevcr_frame_params& p, const hevcr_slice& q, int qpy, int c) {
    const int qbd = 6 * (p.bit_depth - 8);
    const int off = c == 1 ? p.pps_cb_qp_offset + q.slice_cb_qp_offset : p.pps_cr_qp_offset + q.slice_cr_qp_offset;
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
    fc.deblock = p.deblocking_disabled ? 0 : 1;
    return fc;
}

// Pictures carry padding after their last plane. The MC kernels' interior
// path loads whole fixed-size footprints (23 luma rows x 32 samples, 11 chroma
// rows x 16 samples from an aligned start) for every tile up to 16x16, while
// only the rows the tile's own footprint needs are tested against the
// picture: below the last chroma plane a 16x4 tile reads up to 6 rows and 16
// samples more. Eight chroma rows plus 256 bytes cover that.
size_t picture_pad(const Stream& st) { return (size_t)8 * st.stride[2] * st.bpp + 256; }

int alloc_picture(Stream& st, Picture& pic) {
    const size_t total = st.plane_bytes[0] + st.plane_bytes[1] + st.plane_bytes[2] + picture_pad(st);
    int rc = pic.mem.reserve(total);
    if (rc) return rc;
    char* base = static_cast<char*>(pic.mem.p);
    pic.plane[0] = base;
    pic.plane[1] = base + st.plane_bytes[0];
    pic.plane[2] = base + st.plane_bytes[0] + st.plane_bytes[1];
    return HEVCR_OK;
}



// -- intra neighbour availability (intra_prediction.cpp:102-133) -------------

// The reference's availability of a neighbour sample: inside the picture, and
// in an earlier CTB in raster order or in the same CTB at a smaller z-scan
// address of 4x4 luma blocks (one slice, no tiles). With transform units
// visited in decoding order (CUs in order, a CU's units in z-scan order, which
// validation checks), that is exactly "its 4x4 luma cell is already decoded":
// a map of decoded cells, set as each intra transform unit and each inter CU
// is passed, answers it with one bit test (cells outside the picture are never
// set).
struct DecodedMap {
    std::vector<uint64_t> bits;
    size_t ws = 0;
    int gw = 0, gh = 0;
    void reset(int w4, int h4) {
        gw = w4;
        gh = h4;
        ws = (size_t)(w4 + 63) / 64;
        bits.assign(ws * (size_t)h4, 0);
    }
    bool test(int cx, int cy) const {
        return cx >= 0 && cy >= 0 && cx < gw && cy < gh && ((bits[(size_t)cy * ws + (size_t)(cx >> 6)] >> (cx & 63)) & 1);
    }
    // an n x n square of cells at (cx, cy), cx a multiple of n <= 16 (inside one word)
    void mark(int cx, int cy, int n) {
        const uint64_t m = ((1ull << n) - 1) << (cx & 63);
        uint64_t* w = &bits[(size_t)cy * ws + (size_t)(cx >> 6)];
        for (int j = 0; j < n; j++) w[(size_t)j * ws] |= m;
    }
};

// Fills the available reference run [lo, hi) of an intra block of size n at
// component position (xc, yc), and its hole flags. Availability is constant
// over each 4x4 luma block (4 luma / 2 chroma samples along an edge), so one
// test per unit. One slice segment and one tile (cross == null): the left
// column, the top row and the corner of an aligned block are in blocks
// earlier in z-scan (or in the left / upper CTBs, earlier in raster order),
// so they are available exactly when they lie inside the picture, and the
// available entries form one run. Otherwise cross(cx, cy) says whether cell
// (cx, cy), outside the block's CTB, lies in the same slice segment and tile
// (intra_prediction.cpp:108-124): each part is available when decoded and,
// across the CTB boundary, in the same segment and tile. Returns false for a
// pattern that is not one run with at most the hole the kernels take.
template <typename Cross>
bool intra_avail(const DecodedMap& dm, int c, int xc, int yc, int n, int ctb_log2, const Cross* cross, IntraOp& op) {
    const int sub = c ? 1 : 0, unit = c ? 2 : 4;
    const int nu = n / unit;                                 // units (4x4 luma cells) along an edge
    const int cx = (xc << sub) >> 2, cy = (yc << sub) >> 2;  // the block's first cell
    // line indices: [0, n) bottom-left extension (far to near), [n, 2n) left
    // column (bottom to top), 2n corner, [2n + 1, 3n + 1) top row, then the
    // top-right extension
    if (!cross) {
        const bool left = xc > 0, top = yc > 0;
        int bl = 0, tr = 0;  // available units of the bottom-left and top-right extensions
        if (left)
            while (bl < nu && dm.test(cx - 1, cy + nu + bl)) bl++;
        if (top)
            while (tr < nu && dm.test(cx + nu + tr, cy - 1)) tr++;
        if (left) {
            op.lo = (uint8_t)(n - bl * unit);
            op.hi = (uint8_t)(top ? 3 * n + 1 + tr * unit : 2 * n);
        } else if (top) {
            op.lo = (uint8_t)(2 * n + 1);
            op.hi = (uint8_t)(3 * n + 1 + tr * unit);
        } else {
            op.lo = op.hi = 0;
        }
        return true;
    }
    const int cs = ctb_log2 - 2;  // cells per CTB side, log2
    auto avail = [&](int qx, int qy) {
        if (!dm.test(qx, qy)) return false;
        return ((qx >> cs) == (cx >> cs) && (qy >> cs) == (cy >> cs)) || (*cross)(qx, qy);
    };
    // availability of the 4n + 1 entries by unit: bottom-left / left units
    // from the top down, the corner, the top / top-right units left to right
    uint64_t a = 0;  // bit u: unit u of the line (u = 0 the bottom-most bottom-left unit)
    int u = 0;
    for (int k = 2 * nu - 1; k >= 0; k--, u++)
        if (avail(cx - 1, cy + k)) a |= 1ull << u;
    const int uc = u++;  // the corner
    if (avail(cx - 1, cy - 1)) a |= 1ull << uc;
    for (int k = 0; k < 2 * nu; k++, u++)
        if (avail(cx + k, cy - 1)) a |= 1ull << u;
    op.flags &= (uint8_t)~(kIntraGapCorner | kIntraGapTop);
    if (!a) {
        op.lo = op.hi = 0;
        return true;
    }
    auto entry = [&](int v) { return v < uc ? v * unit : (v == uc ? 2 * n : 2 * n + 1 + (v - uc - 1) * unit); };
    auto entry_end = [&](int v) { return v < uc ? (v + 1) * unit : (v == uc ? 2 * n + 1 : 2 * n + 1 + (v - uc) * unit); };
    const int u0 = __builtin_ctzll(a), u1 = 63 - __builtin_clzll(a);  // first and last available unit
    const uint64_t run = (u1 == 63 ? ~0ull : ((1ull << (u1 + 1)) - 1)) & ~((1ull << u0) - 1);
    op.lo = (uint8_t)entry(u0);
    op.hi = (uint8_t)entry_end(u1);
    const uint64_t hole = run & ~a;
    if (!hole) return true;
    // the one hole the kernels take: the corner, or the corner and the top row,
    // with the unit below it (the left column's top) available
    const uint64_t corner = 1ull << uc, top_row = ((1ull << nu) - 1) << (uc + 1);
    if (hole == corner)
        op.flags |= kIntraGapCorner;
    else if (hole == (corner | top_row))
        op.flags |= kIntraGapTop;
    else
        return false;
    return true;
}

uint8_t intra_filter_flags(const FrameConst& fc, int c, int log2, int mode) {
    if (c != 0 || fc.intra_smoothing_disabled || mode == 1 || log2 <= 2) return 0;
    const int dv = std::min(std::abs(mode - 26), std::abs(mode - 10));
    const int thr = log2 == 3 ? 7 : (log2 == 4 ? 1 : 0);
    if (!(dv > thr)) return 0;
    return (uint8_t)(kIntraFilter | (fc.strong_intra_smoothing && log2 == 5 ? kIntraStrong : 0));
}


// Schedules one CTU component's intra blocks (given in decoding order) into
// steps. A block depends on every block of the same CTU holding one of its
// available reference samples and gets a level one above the latest of them;
// blocks of one level write disjoint samples and read none that another of the
// level writes. Within a level, 4x4 blocks form steps of up to four (run side
// by side, 16 lanes each; one per step when !group4), larger blocks a step
// each. Reorders `ops` into step order and sets `span` per block: at a step's
// first block the step's size, plus kIntraGroupStep for 4x4 steps, else 0.
// cx0, cy0: the CTU's component origin, S its size.
void intra_steps(std::vector<IntraOp>& ops, int cx0, int cy0, int S, std::vector<uint8_t>& span, bool group4) {
    span.clear();
    if (ops.empty()) return;  // most CTBs of inter pictures
    const int G = S / 4;  // cells of 4 x 4 component samples
    thread_local std::vector<uint16_t> tls_lmap;  // reused: this runs once per CTB and component
    thread_local std::vector<int> tls_level;
    std::vector<uint16_t>& lmap = tls_lmap;  // per cell: 1 + the level of the block holding it, 0: none
    std::vector<int>& level = tls_level;
    lmap.assign((size_t)G * G, 0);
    level.assign(ops.size(), 0);
    int maxl = 0;
    for (size_t k = 0; k < ops.size(); k++) {
        const IntraOp& op = ops[k];
        const int n = 1 << op.log2, x0 = op.x - cx0, y0 = op.y - cy0;
        // the cells holding the available entries of the reference line, taken
        // segment by segment (samples of a neighbour CTU are done before: skipped)
        int lv = 0;
        auto visit = [&](int px, int py) {
            if (px < 0 || py < 0 || px >= S || py >= S) return;
            lv = std::max(lv, (int)lmap[(size_t)((py >> 2) * G + (px >> 2))]);
        };
        const int lo = op.lo, hi = op.hi;
        if (lo < 2 * n) {  // left column and its extension: entry j is sample row y0 + 2n - 1 - j
            const int ya = y0 + 2 * n - std::min(hi, 2 * n), yb = y0 + 2 * n - 1 - lo;
            for (int py = ya & ~3; py <= yb; py += 4) visit(x0 - 1, std::max(py, ya));
        }
        if (lo <= 2 * n && 2 * n < hi) visit(x0 - 1, y0 - 1);  // the corner
        if (hi > 2 * n + 1) {  // top row and its extension: entry j is sample column x0 + j - 2n - 1
            const int xa = x0 + std::max(lo, 2 * n + 1) - 2 * n - 1, xb = x0 + hi - 2 * n - 2;
            for (int px = xa & ~3; px <= xb; px += 4) visit(std::max(px, xa), y0 - 1);
        }
        level[k] = lv;
        maxl = std::max(maxl, lv);
        for (int cy = y0 >> 2; cy < (y0 + n) >> 2; cy++)
            for (int cx = x0 >> 2; cx < (x0 + n) >> 2; cx++) lmap[(size_t)cy * G + cx] = (uint16_t)(lv + 1);
    }
    // blocks by level (a counting sort keeps decoding order inside a level);
    // within a level the 4x4 blocks come first, in steps of up to four
    // consecutive blocks, then the larger blocks, a step each
    thread_local std::vector<uint32_t> tls_start, tls_order;
    std::vector<uint32_t>& start = tls_start;
    std::vector<uint32_t>& order = tls_order;
    start.assign((size_t)(2 * (maxl + 1) + 1), 0);
    for (size_t k = 0; k < ops.size(); k++) start[(size_t)(2 * level[k] + (ops[k].log2 != 2)) + 1]++;
    for (size_t b = 1; b < start.size(); b++) start[b] += start[b - 1];
    order.resize(ops.size());
    for (size_t k = 0; k < ops.size(); k++) order[start[(size_t)(2 * level[k] + (ops[k].log2 != 2))]++] = (uint32_t)k;
    thread_local std::vector<IntraOp> tls_out;  // swapped with ops below: buffers circulate, none is freed
    std::vector<IntraOp>& out = tls_out;
    out.clear();
    out.reserve(ops.size());
    span.assign(ops.size(), 0);
    int group = 0;
    size_t first = 0;
    int prev_bucket = -1;
    size_t at = 0;
    for (int b = 0; b < 2 * (maxl + 1); b++) {
        const size_t end = start[(size_t)b];  // start[b] is now the end of bucket b
        for (; at < end; at++) {
            const IntraOp& op = ops[order[at]];
            if (b & 1 || !group4) {  // a step of its own
                span[out.size()] = 1;
                out.push_back(op);
                continue;
            }
            if (b != prev_bucket || group == 4) group = 0;
            prev_bucket = b;
            if (group == 0) first = out.size();
            out.push_back(op);
            span[first] = (uint8_t)(kIntraGroupStep | ++group);
        }
    }
    ops.swap(out);
}

// -- frame preparation --------------------------------------------------------

// Host image of a frame's device block: sections at 256-byte offsets.
// The sections of a frame's device block: recorded first (their sources stay
// alive until the copy), then copied once into host staging memory.
struct BlobWriter {
    struct Part {
        const void* src;
        size_t bytes, off;
    };
    std::vector<Part> parts;
    size_t size = 0;
    size_t put(const void* src, size_t bytes) {
        const size_t off = align256(size);
        parts.push_back({src, bytes, off});
        size = off + bytes;
        return off;
    }
    // right after the previous part (one array in several pieces)
    size_t put_next(const void* src, size_t bytes) {
        const size_t off = size;
        parts.push_back({src, bytes, off});
        size = off + bytes;
        return off;
    }
    void copy_to(uint8_t* dst) const {
        for (const Part& q : parts)
            if (q.bytes) std::memcpy(dst + q.off, q.src, q.bytes);
    }
};

// Work vectors of prepare_frame, one set per host thread, reused from frame
// to frame: a frame's lists are a few MB, and allocating them afresh makes
// every frame fault its pages in (and the allocator return them) on every
// ingest thread at once.
struct PrepScratch {
    std::vector<uint32_t> ctu_ranges, op_begin, ictu, diag_begin, at;
    std::vector<McTile> tiles, tilesb, tiles8;
    std::vector<McBand> sclass[2][4], bands;  // [uni, bi][rows / 4 - 1], then with their destinations
    DecodedMap dmap;
    ResLists res;                          // residual blocks: inter CUs' from validation, then the others
    std::vector<std::vector<IntraOp>> ops_c[3];
    std::vector<uint8_t> ctu_has_intra, op_span, span, scale_m, buf, ictl;
    std::vector<IntraOp> ops;
    std::vector<int> level;
    std::vector<uint32_t> ctl;
    std::vector<hevcr_weight> segw;  // per slice segment and list entry: the weights MC applies
    PicLayout layout;
    BlobWriter w;
};

// This thread's work vectors (reached through a plain reference: each direct
// use of a thread-local object in a shared library is a call to find it).
PrepScratch& prep_scratch() {
    thread_local PrepScratch tls_scratch;
    return tls_scratch;
}

// After validate(st, syn, &prep_scratch().res) on this thread.
int prepare_frame(hevcr_engine* e, Stream& st, int stream_id, const hevcr_frame_syntax* syn, hevcr_frame* f,
                  UploadCtx* up) {
    PrepScratch& W = prep_scratch();
    const double t_start = now_seconds();
    const hevcr_frame_params& p = syn->p;
    f->stream_id = stream_id;
    f->p = p;
    f->fc = make_const(st, p);
    f->bpp = st.bpp;
    f->ncu = syn->num_cus;
    f->npu = syn->num_pus;
    f->ntu = syn->num_tus;
    FrameConst& fcw = f->fc;
    const FrameConst& fc = f->fc;
    const int b = st.bpp;
    const int nctb = fc.wctb * fc.hctb;
    f->nctb = (uint32_t)nctb;
    PicLayout& L = W.layout;
    if (int rc = build_layout(syn, L)) return rc;
    const int nseg = (int)L.segs.size();
    auto seg_at = [&](int x, int y) -> const hevcr_slice& {
        return L.segs[(size_t)L.seg_of((y >> p.ctb_log2) * fc.wctb + (x >> p.ctb_log2))];
    };
    bool any_flags = false;
    for (uint32_t ci = 0; ci < syn->num_cus && !any_flags; ci++) any_flags = syn->cus[ci].flags != 0;
    fcw.pcm_lf_off = syn->pcm_loop_filter_disabled != 0;
    // Slice-segment parameters: deblocking runs if some segment has it on; the
    // motion-compensation weights are per segment where segments weight the
    // lists differently (a table per segment and list entry, and one of each
    // segment's denominators, looked up by the work item's CTB; default
    // weighting is the explicit form with identity weights at denominator 0,
    // bit for bit); the reference pictures of the entries are the picture's
    // lists, built from the first segment (decoder.cpp:108-115).
    fcw.deblock = 0;
    for (const hevcr_slice& q : L.segs) fcw.deblock |= q.deblocking_disabled ? 0 : 1;
    f->seg_weights = false;
    std::vector<hevcr_weight>& segw = W.segw;
    if (nseg > 1) {
        int any_wp = -1;  // a segment with explicit weighting
        for (int k = 0; k < nseg; k++)
            if (L.segs[(size_t)k].weighted_pred && L.segs[(size_t)k].slice_type != HEVCR_SLICE_I) any_wp = k;
        fcw.weighted = any_wp >= 0;
        if (any_wp >= 0) {
            const hevcr_slice& w0 = L.segs[(size_t)any_wp];
            fcw.luma_log2_wd = w0.luma_log2_wd;
            fcw.chroma_log2_wd = w0.chroma_log2_wd;
            f->seg_weights = true;
            // [seg][list][entry] weights, then one denominator record per segment
            const size_t nw = (size_t)nseg * 2 * HEVCR_MAX_REFS;
            segw.assign(nw + (size_t)nseg, hevcr_weight{});
            for (int k = 0; k < nseg; k++) {
                const hevcr_slice& q = L.segs[(size_t)k];
                const bool wp = q.weighted_pred && q.slice_type != HEVCR_SLICE_I;
                hevcr_weight& wd = segw[nw + (size_t)k];
                wd.w[0] = (int16_t)(wp ? q.luma_log2_wd : 0);
                wd.w[1] = (int16_t)(wp ? q.chroma_log2_wd : 0);
                for (int l = 0; l < 2; l++)
                    for (int i = 0; i < HEVCR_MAX_REFS; i++) {
                        hevcr_weight& w = segw[((size_t)k * 2 + l) * HEVCR_MAX_REFS + i];
                        if (wp) {
                            w = q.weights[l][i];
                        } else {
                            w.w[0] = w.w[1] = w.w[2] = 1;
                            w.o[0] = w.o[1] = w.o[2] = 0;
                        }
                    }
            }
            f->nseg_w = nseg;
        }
    }
    // SAO activity: the reference skips the whole pass when no CTB has SAO (sao.cpp:33-39)
    f->sao_active = false;
    if (p.sao_enabled)
        for (int i = 0; i < nctb && !f->sao_active; i++)
            for (int c = 0; c < 3; c++) f->sao_active |= syn->sao[i].type[c] != 0;
    const auto tk = e->tuning.find(std::make_tuple(st.w, st.h, st.bd));
    if (tk != e->tuning.end()) f->tune = tk->second;
    // full runs deblock and apply SAO in one kernel when SAO runs and CTBs are 32x32 or 64x64
    // (it holds up to 2x2 CTBs' SAO parameters); pictures with slice segment or
    // tile boundaries or kept samples take the separate kernels
    f->has_ctl = L.multi || any_flags;
    f->lf_fused = f->tune.lf_fused && f->sao_active && p.ctb_log2 >= 5 && !f->has_ctl;

    // per-CU record ranges, from validation
    ResLists& res = W.res;
    const std::vector<uint32_t>& first_tu = res.first_tu;
    const std::vector<uint32_t>& first_pu = res.first_pu;
    const bool any_intra = res.any_intra;  // intra CUs read which 4x4 cells are decoded (DecodedMap)
    // per-CTB record ranges (records are in decoding order, CTB by CTB); the
    // expansion kernel stages at most 64 CUs, 256 leaves and 256 PUs per CTB
    std::vector<uint32_t>& ctu_ranges = W.ctu_ranges;
    ctu_ranges.assign((size_t)nctb * 6, 0);
    {
        // CTBs in the picture's scan (tile by tile with tiles)
        thread_local std::vector<uint32_t> tls_ts2rs;
        std::vector<uint32_t>& ts2rs = tls_ts2rs;
        ts2rs.resize((size_t)nctb);
        for (int a = 0; a < nctb; a++) ts2rs[L.ts(a)] = (uint32_t)a;
        uint32_t ci = 0;
        for (int ts = 0; ts < nctb; ts++) {
            const int ctb = (int)ts2rs[(size_t)ts];
            const uint32_t c0 = ci;
            while (ci < syn->num_cus &&
                   (syn->cus[ci].y >> p.ctb_log2) * fc.wctb + (syn->cus[ci].x >> p.ctb_log2) == ctb)
                ci++;
            uint32_t* r = &ctu_ranges[(size_t)ctb * 6];
            r[0] = c0;
            r[1] = ci;
            r[2] = first_tu[c0];
            r[3] = first_tu[ci];
            r[4] = first_pu[c0];
            r[5] = first_pu[ci];
            if (r[1] - r[0] > 64 || r[3] - r[2] > 256 || r[5] - r[4] > 256)
                return fail(HEVCR_EUNSUPPORTED, "ctb %d: %u CUs, %u transform leaves, %u PUs", ctb, r[1] - r[0],
                            r[3] - r[2], r[5] - r[4]);
        }
    }

    // Motion compensation work of every PU. Strips (tuning mc_strips): 4 luma
    // columns over up to 16 rows (uni-predicted 8-bit pictures, mc_strips 1)
    // or 8 rows, for the bands of a PU whose windows lie inside the picture
    // in every list; the other bands (the reference's edge clamping) and every
    // PU without strips as tiles of at most 16x16: uni- and bi-predicted ones
    // apart, 8x8 tiles (tuning mc8) in a list of their own.
    std::vector<McTile>& tiles = W.tiles;
    std::vector<McTile>& tilesb = W.tilesb;
    std::vector<McTile>& tiles8 = W.tiles8;
    tiles.clear();
    tilesb.clear();
    tiles8.clear();
    for (auto& l : W.sclass)
        for (auto& c : l) c.clear();
    const bool use_strips = f->tune.mc_strips != 0;
    const int uni_rows = f->tune.mc_strips == 1 && st.bpp == 1 ? 16 : 8;
    std::memset(f->kbytes, 0, sizeof(f->kbytes));
    // algorithmic bytes, summed in integers per kernel (MC, MC8): reference
    // windows read (luma and both chroma, per list) and prediction samples
    // written (1.5 per luma sample)
    uint64_t win[2] = {0, 0}, area[2] = {0, 0};
    const bool mc8_split = f->tune.mc8 && !use_strips;
    for (uint32_t i = 0; i < syn->num_pus; i++) {
        const hevcr_pu& pu = syn->pus[i];
        const bool bi = pu.pred_flags == (HEVCR_PRED_L0 | HEVCR_PRED_L1);
        auto add_tiles = [&](int y0, int h) {
            for (int ty = y0; ty < y0 + h; ty += 16)
                for (int tx = 0; tx < pu.w; tx += 16) {
                    McTile t;
                    std::memset(&t, 0, sizeof(t));
                    t.x = (uint16_t)(pu.x + tx);
                    t.y = (uint16_t)(pu.y + ty);
                    t.w = (uint8_t)std::min(16, pu.w - tx);
                    t.h = (uint8_t)std::min(16, y0 + h - ty);
                    std::memcpy(t.mv, pu.mv, sizeof(t.mv));
                    t.ref[0] = pu.ref_idx[0];
                    t.ref[1] = pu.ref_idx[1];
                    t.flags = pu.pred_flags;
                    (t.w == 8 && t.h == 8 && f->tune.mc8 && !use_strips ? tiles8 : (bi ? tilesb : tiles)).push_back(t);
                }
        };
        if (!use_strips) {
            add_tiles(0, pu.h);
        } else {
            const int rows = bi ? 8 : uni_rows;
            // the band's windows: luma (x - 3 .. x + w + 3) x (y - 3 .. y + th + 3) and
            // chroma (x/2 - 1 .. x/2 + w/2 + 1) x (y/2 - 1 .. y/2 + th/2 + 1), moved by each list's
            // motion; those of the whole PU hold every band's (most PUs: no test per band)
            auto window_inside = [&](int y, int th) {
                bool inside = true;
                for (int l = 0; l < 2; l++) {
                    if (!(pu.pred_flags & (1 << l))) continue;
                    const int mx = pu.mv[l][0], my = pu.mv[l][1];
                    const int xl = pu.x + (mx >> 2) - 3, yl = y + (my >> 2) - 3;
                    const int xc = (pu.x >> 1) + (mx >> 3) - 1, yc = (y >> 1) + (my >> 3) - 1;
                    inside = inside && xl >= 0 && yl >= 0 && xl + pu.w + 7 <= fc.W && yl + th + 7 <= fc.H && xc >= 0 &&
                             yc >= 0 && xc + pu.w / 2 + 3 <= (fc.W >> 1) && yc + th / 2 + 3 <= (fc.H >> 1);
                }
                return inside;
            };
            const bool all_inside = window_inside(pu.y, pu.h);
            for (int ty = 0; ty < pu.h; ty += rows) {
                const int th = std::min(rows, pu.h - ty), y = pu.y + ty;
                if (!all_inside && !window_inside(y, th)) {
                    add_tiles(ty, th);
                    continue;
                }
                // one band record; its pu.w / 4 strips are written on the device
                McBand t;
                t.s.x = pu.x;
                t.s.y = (uint16_t)y;
                std::memcpy(t.s.mv, pu.mv, sizeof(t.s.mv));
                t.s.ref[0] = pu.ref_idx[0];
                t.s.ref[1] = pu.ref_idx[1];
                t.s.flags = pu.pred_flags;
                t.s.h = (uint8_t)th;
                t.n = (uint32_t)pu.w / 4;
                t.dst = 0;
                W.sclass[bi ? 1 : 0][th / 4 - 1].push_back(t);
            }
        }
        const int nlists = (pu.pred_flags & 1) + ((pu.pred_flags >> 1) & 1);
        const int k8 = mc8_split && pu.w == 8 && pu.h == 8;
        win[k8] += (uint64_t)nlists * (uint64_t)((pu.w + 7) * (pu.h + 7) + 2 * (pu.w / 2 + 3) * (pu.h / 2 + 3));
        area[k8] += (uint64_t)(pu.w * pu.h);
    }
    f->kbytes[HEVCR_KERNEL_MC] = b * ((double)win[0] + 1.5 * (double)area[0]);
    f->kbytes[HEVCR_KERNEL_MC8] = b * ((double)win[1] + 1.5 * (double)area[1]);
    const double mc_bytes = f->kbytes[HEVCR_KERNEL_MC] + f->kbytes[HEVCR_KERNEL_MC8];
    // strip lists: height classes 16, 12, 8, 4, each padded to whole waves of
    // 64; uni-predicted strips, then bi-predicted ones, in one device region
    // that the bands are expanded into (the padding as bands of empty strips)
    f->strip_rows = 0;
    std::vector<McBand>& bands = W.bands;
    bands.clear();
    uint32_t nstrips[2] = {0, 0};
    for (int l = 0; l < 2; l++) {
        uint32_t at = l ? nstrips[0] : 0;
        const uint32_t first = at;
        for (int c = 3; c >= 0; c--) {
            std::vector<McBand>& cls = W.sclass[l][c];
            if (cls.empty()) continue;
            if (!l) f->strip_rows = std::max(f->strip_rows, c >= 2 ? 16 : 8);
            for (McBand& bd : cls) {
                bd.dst = at;
                at += bd.n;
                bands.push_back(bd);
            }
            if ((at - first) % 64) {
                McBand pad;
                std::memset(&pad, 0, sizeof(pad));
                pad.s.h = (uint8_t)(4 * (c + 1));
                pad.dst = at;
                pad.n = 64 - (at - first) % 64;
                at += pad.n;
                bands.push_back(pad);
            }
        }
        nstrips[l] = at - first;
    }
    f->nstripw = nstrips[0] / 64;
    f->nstripwb = nstrips[1] / 64;
    f->nbands = (uint32_t)bands.size();
    f->ntiles = (uint32_t)tiles.size();
    f->ntilesb = (uint32_t)tilesb.size();
    f->ntiles8 = (uint32_t)tiles8.size();
    std::memset(f->bytes, 0, sizeof(f->bytes));
    f->bytes[HEVCR_STAGE_MC] = mc_bytes;

    // residual blocks per size; intra operations per CTU and component. Each
    // size's blocks are grouped by plane, decoding order kept within a plane
    // (the three lists of a size go to the device one after the other): a
    // workgroup's blocks then cover one compact region of one plane, and the
    // kernels hand each XCD a contiguous run of them (xcd_remap), so that
    // partly written cache lines fill up in one L2. Those of inter CUs are
    // there already (validation emitted them); intra and PCM CUs' follow.
    std::vector<std::vector<IntraOp>>* ops_c = W.ops_c;
    for (int c = 0; c < 3; c++) {
        if (ops_c[c].size() < (size_t)nctb) ops_c[c].resize((size_t)nctb);
        for (int i = 0; i < nctb; i++) ops_c[c][(size_t)i].clear();
    }
    std::vector<uint8_t>& ctu_has_intra = W.ctu_has_intra;
    ctu_has_intra.assign((size_t)nctb, 0);
    const int qbd = 6 * (p.bit_depth - 8);
    double intra_bytes = 0;
    DecodedMap& dmap = W.dmap;
    dmap.reset(fc.gw, fc.gh);
    // across a CTB boundary a decoded cell is available when it lies in the
    // same slice segment and tile as the block (intra_prediction.cpp:108-124)
    int cur_ctb = 0;
    const int cs = p.ctb_log2 - 2;
    auto cross = [&](int qx, int qy) {
        const int a = (qy >> cs) * fc.wctb + (qx >> cs);
        return L.seg_of(a) == L.seg_of(cur_ctb) && L.tile_of(a) == L.tile_of(cur_ctb);
    };
    const decltype(cross)* crossp = L.multi ? &cross : nullptr;
    for (uint32_t ci = 0; ci < syn->num_cus; ci++) {
      const hevcr_cu& cu = syn->cus[ci];
      const bool intra = cu.pred_mode == HEVCR_MODE_INTRA;
      const hevcr_slice& sq = seg_at(cu.x, cu.y);
      cur_ctb = (cu.y >> p.ctb_log2) * fc.wctb + (cu.x >> p.ctb_log2);
      if (cu.flags & HEVCR_CU_PCM) {  // the samples, written before intra prediction reads them
        const hevcr_tu& t = syn->tus[first_tu[ci]];
        for (int c = 0; c < 3; c++) {
            ResBlock rb;
            rb.coeff_off = t.coeff_off[c];
            rb.x = (uint16_t)(c ? cu.x / 2 : cu.x);
            rb.y = (uint16_t)(c ? cu.y / 2 : cu.y);
            rb.log2 = (uint8_t)(c ? cu.log2_size - 1 : cu.log2_size);
            rb.cidx = (uint8_t)c;
            rb.qp = 0;
            rb.flags = kResPcm;
            const double in = syn->sig ? 4.0 * (1.0 + syn->sig[t.coeff_off[c]]) : 2.0 * (1 << (2 * rb.log2));
            res.add(rb, in + b * (1 << (2 * rb.log2)));
        }
        dmap.mark(cu.x >> 2, cu.y >> 2, 1 << (cu.log2_size - 2));
        continue;
      }
      const bool bypass = cu.flags & HEVCR_CU_BYPASS;
      if (!intra) {  // inter CUs (most of them): their residual blocks came from validation
        if (any_intra) dmap.mark(cu.x >> 2, cu.y >> 2, 1 << (cu.log2_size - 2));
        continue;
      }
      for (uint32_t ti = first_tu[ci]; ti < first_tu[ci + 1]; ti++) {
        const hevcr_tu& t = syn->tus[ti];
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
                rb.qp = (uint8_t)(c ? chroma_qp_prime(p, sq, cu.qp_y, c) : cu.qp_y + qbd);
                rb.flags = (uint8_t)(kResIntra | (c == 0 && l2 == 2 ? kResDst : 0) |
                                     (((t.tskip >> c) & 1) ? kResTskip : 0) | (bypass ? kResBypass : 0));
                // coefficients read (dense block, or the compact list), residual written
                const double in = syn->sig ? 4.0 * (1.0 + syn->sig[t.coeff_off[c]]) : 2.0 * (1 << (2 * l2));
                res.add(rb, in + 2.0 * (1 << (2 * l2)));
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
                op.flags = (uint8_t)(intra_filter_flags(fc, c, l2, op.mode) | (coded ? kIntraCbf : 0));
                op.ang = kIntraAngles.ang[op.mode];
                op.inv = kIntraAngles.inv[op.mode];
                if (!intra_avail(dmap, c, xc, yc, 1 << l2, p.ctb_log2, crossp, op))
                    return fail(HEVCR_EUNSUPPORTED, "intra block at (%d,%d) c%d: neighbour availability pattern", xc, yc,
                                c);
                ops_c[c][(size_t)ctu].push_back(op);
                ctu_has_intra[(size_t)ctu] = 1;
                intra_bytes += (coded ? 2.0 : 0.0) * (1 << (2 * l2));
            }
        }
        dmap.mark(t.x >> 2, t.y >> 2, 1 << (t.log2_size - 2));  // decoded: later blocks may read it
      }
    }
    for (int q = 0; q < 4; q++) {
        f->nblocks[q] = res.count(q);
        f->kbytes[HEVCR_KERNEL_RESIDUAL4 + q] += res.kbytes[q];
    }
    f->bytes[HEVCR_STAGE_RESIDUAL] = res.bytes;

    std::vector<IntraOp>& ops = W.ops;
    std::vector<uint8_t>& op_span = W.op_span;
    std::vector<uint8_t>& span = W.span;
    std::vector<uint32_t>& op_begin = W.op_begin;
    ops.clear();
    op_span.clear();
    op_begin.assign((size_t)nctb * 3 + 1, 0);
    for (int ctu = 0; ctu < nctb; ctu++)
        for (int c = 0; c < 3; c++) {
            std::vector<IntraOp>& list = ops_c[c][(size_t)ctu];
            if (list.empty()) {  // (most CTUs of inter pictures)
                op_begin[(size_t)ctu * 3 + c] = (uint32_t)ops.size();
                continue;
            }
            if (list.size() > 256)  // k_intra stages at most 256 operations per CTU
                return fail(HEVCR_ERROR, "ctu %d: %zu intra blocks", ctu, list.size());
            const int S = (1 << p.ctb_log2) >> (c ? 1 : 0);
            intra_steps(list, (ctu % fc.wctb) * S, (ctu / fc.wctb) * S, S, span, f->tune.intra_group4 != 0);
            for (size_t k = 0; k < list.size();) {  // steps tile the list; 4x4 steps hold 4x4 blocks only
                const size_t cnt = span[k] & 0x7fu;
                if (!cnt || k + cnt > list.size()) return fail(HEVCR_ERROR, "ctu %d: intra step layout", ctu);
                for (size_t q = k; q < k + cnt; q++)
                    if ((span[k] & kIntraGroupStep) ? list[q].log2 != 2 || (q > k && span[q]) : cnt != 1)
                        return fail(HEVCR_ERROR, "ctu %d: intra step layout", ctu);
                k += cnt;
            }
            op_begin[(size_t)ctu * 3 + c] = (uint32_t)ops.size();
            ops.insert(ops.end(), list.begin(), list.end());
            op_span.insert(op_span.end(), span.begin(), span.end());
        }
    op_begin[(size_t)nctb * 3] = (uint32_t)ops.size();
    f->has_intra = !ops.empty();
    // CTUs with intra work, grouped by dependency level. Intra prediction reads
    // samples of the left, above-left, above and above-right CTUs only; those
    // without intra blocks are final before this stage, so a CTU waits only
    // for neighbours that have intra blocks themselves: level = 1 + the
    // highest level among them (0 if none). In an intra picture this is the
    // diagonal x + 2y; in inter pictures with scattered intra CUs it is far
    // fewer launches.
    std::vector<int>& level = W.level;
    level.assign((size_t)nctb, -1);
    int ndiag = 0;
    for (int y = 0; y < fc.hctb; y++)
        for (int x = 0; x < fc.wctb; x++) {
            if (!ctu_has_intra[(size_t)(y * fc.wctb + x)]) continue;
            int lv = 0;
            const int nb[4][2] = {{x - 1, y}, {x - 1, y - 1}, {x, y - 1}, {x + 1, y - 1}};
            for (const auto& q : nb)
                if (q[0] >= 0 && q[1] >= 0 && q[0] < fc.wctb)
                    lv = std::max(lv, level[(size_t)(q[1] * fc.wctb + q[0])] + 1);
            level[(size_t)(y * fc.wctb + x)] = lv;
            ndiag = std::max(ndiag, lv + 1);
        }
    std::vector<uint32_t>& ictu = W.ictu;
    std::vector<uint32_t>& diag_begin = W.diag_begin;
    diag_begin.assign((size_t)ndiag + 1, 0);
    f->diag_count.assign((size_t)ndiag, 0);
    for (int c = 0; c < nctb; c++)
        if (level[(size_t)c] >= 0) f->diag_count[(size_t)level[(size_t)c]]++;
    for (int dg = 0; dg < ndiag; dg++) diag_begin[(size_t)dg + 1] = diag_begin[(size_t)dg] + f->diag_count[(size_t)dg];
    ictu.resize(diag_begin[(size_t)ndiag]);
    {
        std::vector<uint32_t>& at = W.at;
        at.assign(diag_begin.begin(), diag_begin.end() - 1);
        for (int c = 0; c < nctb; c++)
            if (level[(size_t)c] >= 0) ictu[at[(size_t)level[(size_t)c]]++] = (uint32_t)c;
    }
    // persistent intra: per CTB, whether it waits for an intra neighbour
    // above or left (kIctlWaits) and whether one below or right waits for it
    // (kIctlSignals); an item without either skips that synchronisation
    std::vector<uint8_t>& ictl = W.ictl;
    ictl.assign((size_t)nctb, 0);
    for (int y = 0; y < fc.hctb; y++)
        for (int x = 0; x < fc.wctb; x++) {
            if (level[(size_t)(y * fc.wctb + x)] < 0) continue;
            auto intra_at = [&](int qx, int qy) {
                return qx >= 0 && qy >= 0 && qx < fc.wctb && qy < fc.hctb && level[(size_t)(qy * fc.wctb + qx)] >= 0;
            };
            uint8_t v = 0;
            if (intra_at(x - 1, y) || intra_at(x - 1, y - 1) || intra_at(x, y - 1) || intra_at(x + 1, y - 1)) v |= kIctlWaits;
            if (intra_at(x + 1, y) || intra_at(x + 1, y + 1) || intra_at(x, y + 1) || intra_at(x - 1, y + 1))
                v |= kIctlSignals;
            ictl[(size_t)(y * fc.wctb + x)] = v;
        }
    f->intra_dense = 2 * ictu.size() >= (size_t)nctb;
    const int ctbs = 1 << p.ctb_log2;
    intra_bytes += (double)ictu.size() * (2.0 * b + 2.0) * 1.5 * ctbs * ctbs;
    f->bytes[HEVCR_STAGE_INTRA] = intra_bytes;
    const double S = 1.5 * p.width * p.height;
    // deblocking reads the side information and motion of every 4x4 cell
    const double side = (4.0 + (syn->num_pus ? 12.0 : 0.0)) * fc.gw * fc.gh;
    f->bytes[HEVCR_STAGE_DEBLOCK] = !fc.deblock ? 0 : 2 * S * b + side;
    f->bytes[HEVCR_STAGE_SAO] = f->sao_active ? 2 * S * b + 16.0 * nctb : 0;
    f->lf_bytes = f->lf_fused ? 2 * S * b + (!fc.deblock ? 0 : side) + 16.0 * nctb : 0;
    f->bytes[HEVCR_STAGE_EXPAND] = syn->num_cus * sizeof(hevcr_cu) + syn->num_tus * sizeof(uint32_t) +
                                   syn->num_pus * sizeof(hevcr_pu) + 4.0 * fc.gw * fc.gh +
                                   (syn->num_pus ? 12.0 * fc.gw * fc.gh : 0);
    f->kbytes[HEVCR_KERNEL_EXPAND] = f->bytes[HEVCR_STAGE_EXPAND];
    f->kbytes[HEVCR_KERNEL_INTRA] = f->bytes[HEVCR_STAGE_INTRA];
    f->kbytes[HEVCR_KERNEL_DEBLOCK] = f->bytes[HEVCR_STAGE_DEBLOCK];
    f->kbytes[HEVCR_KERNEL_SAO] = f->bytes[HEVCR_STAGE_SAO];

    // per-CTB control words: slice-segment / tile boundaries for deblocking and
    // SAO, the segment's deblocking offsets, CTBs with kept samples
    std::vector<uint32_t>& ctl = W.ctl;
    if (f->has_ctl) {
        ctl.assign((size_t)nctb, 0);
        for (uint32_t ci = 0; ci < syn->num_cus; ci++) {
            const hevcr_cu& cu = syn->cus[ci];
            if (((cu.flags & HEVCR_CU_PCM) && syn->pcm_loop_filter_disabled) || (cu.flags == HEVCR_CU_BYPASS))
                ctl[(size_t)((cu.y >> p.ctb_log2) * fc.wctb + (cu.x >> p.ctb_log2))] |= ctl::kNoFilter;
        }
        for (int y = 0; y < fc.hctb; y++)
            for (int x = 0; x < fc.wctb; x++) {
                const int a = y * fc.wctb + x, sa = L.seg_of(a);
                const hevcr_slice& q = L.segs[(size_t)sa];
                uint32_t v = ctl[(size_t)a];
                v |= (uint32_t)(q.beta_offset_div2 + 6) << info::kBetaShift | (uint32_t)(q.tc_offset_div2 + 6) << info::kTcShift;
                if (q.deblocking_disabled) v |= ctl::kDbkOff;
                // deblocking (deblocking.cpp:69-113): the Q side (this CTB) decides
                auto dbk_cut = [&](int nb) {
                    return (!L.lf_tiles && L.tile_of(nb) != L.tile_of(a)) ||
                           (L.seg_of(nb) != sa && !q.loop_filter_across_slices);
                };
                if (x > 0 && dbk_cut(a - 1)) v |= ctl::kLeftOff;
                if (y > 0 && dbk_cut(a - fc.wctb)) v |= ctl::kTopOff;
                // SAO (sao.cpp:142-185): the later segment's flag, the tile flag
                static const int kNb[8][3] = {{-1, 0, ctl::kNbL},   {1, 0, ctl::kNbR},   {0, -1, ctl::kNbT},
                                              {0, 1, ctl::kNbB},    {-1, -1, ctl::kNbTL}, {1, -1, ctl::kNbTR},
                                              {-1, 1, ctl::kNbBL},  {1, 1, ctl::kNbBR}};
                for (const auto& d : kNb) {
                    const int nx = x + d[0], ny = y + d[1];
                    if (nx < 0 || ny < 0 || nx >= fc.wctb || ny >= fc.hctb) continue;
                    const int nb = ny * fc.wctb + nx, sb = L.seg_of(nb);
                    const bool cut = (sb != sa && !L.segs[(size_t)std::max(sa, sb)].loop_filter_across_slices) ||
                                     (!L.lf_tiles && L.tile_of(nb) != L.tile_of(a));
                    if (cut) v |= 1u << d[2];
                }
                ctl[(size_t)a] = v;
            }
    }

    // one device block: records, coefficients, SAO parameters and the derived lists
    BlobWriter& w = W.w;
    w.parts.clear();
    w.size = 0;
    f->o_cus = w.put(syn->cus, syn->num_cus * sizeof(hevcr_cu));
    f->o_pus = w.put(syn->pus, syn->num_pus * sizeof(hevcr_pu));
    f->o_tus = w.put(res.tuw.data(), res.tuw.size() * sizeof(uint32_t));
    f->compact = syn->sig != nullptr;
    f->o_coeffs = f->compact ? w.put(syn->sig, syn->num_sig * sizeof(uint32_t))
                             : w.put(syn->coeffs, syn->num_coeffs * sizeof(int16_t));
    f->o_sao = w.put(syn->sao, p.sao_enabled ? (size_t)nctb * sizeof(hevcr_sao) : 0);
    f->o_first_tu = w.put(first_tu.data(), first_tu.size() * 4);
    f->o_first_pu = w.put(first_pu.data(), first_pu.size() * 4);
    f->o_ctu_ranges = w.put(ctu_ranges.data(), ctu_ranges.size() * 4);
    f->o_tiles = w.put(tiles.data(), tiles.size() * sizeof(McTile));
    f->o_tilesb = w.put(tilesb.data(), tilesb.size() * sizeof(McTile));
    f->o_tiles8 = w.put(tiles8.data(), tiles8.size() * sizeof(McTile));
    f->o_bands = w.put(bands.data(), bands.size() * sizeof(McBand));
    for (int q = 0; q < 4; q++) {
        f->o_blocks[q] = w.put(res.blocks[q][0].data(), res.n[q][0] * sizeof(ResBlock));
        for (int c = 1; c < 3; c++) w.put_next(res.blocks[q][c].data(), res.n[q][c] * sizeof(ResBlock));
    }
    f->o_ops = w.put(ops.data(), ops.size() * sizeof(IntraOp));
    f->o_op_begin = w.put(op_begin.data(), op_begin.size() * 4);
    f->o_op_span = w.put(op_span.data(), op_span.size());
    f->o_ictu = w.put(ictu.data(), ictu.size() * 4);
    f->o_diag_begin = w.put(diag_begin.data(), diag_begin.size() * 4);
    f->o_ictl = w.put(ictl.data(), ictl.size());
    if (f->has_ctl) f->o_ctl = w.put(ctl.data(), ctl.size() * 4);
    if (f->seg_weights) {
        f->o_segw = w.put(segw.data(), segw.size() * sizeof(hevcr_weight));
        f->o_ctbseg = w.put(syn->ctb_slice, (size_t)nctb);
    }
    f->scaled = syn->scaling != nullptr;
    std::vector<uint8_t>& scale_m = W.scale_m;
    if (f->scaled) {
        scale_m.resize(kScaleBytes);
        scaling_factors(*syn->scaling, scale_m.data());
        f->o_scale = w.put(scale_m.data(), scale_m.size());
    }
    f->blob.pool = e->blobs;
    // the strip lists follow what is uploaded: written on the device from the bands
    f->o_strips = align256(w.size);
    f->o_stripsb = f->o_strips + (size_t)f->nstripw * 64 * sizeof(McStrip);
    const size_t total = f->o_stripsb + (size_t)f->nstripwb * 64 * sizeof(McStrip);
    int rc = e->blobs->get(total, &f->blob.p, &f->blob.cap);
    if (rc) return rc;
    f->upload_bytes = w.size;
    McBand* const dev_bands = f->blob.as<McBand>(f->o_bands);
    McStrip* const dev_strips = f->blob.as<McStrip>(f->o_strips);
    // copied on an upload stream, so queued reconstruction work is not waited for
    if (up) {  // the thread's own stream and pinned staging: other threads' uploads are not waited for
        const double t_lists = now_seconds();
        const int k = up->next;
        if ((rc = up->reserve(k, w.size))) return rc;
        w.copy_to(static_cast<uint8_t*>(up->pinned[k]));
        const double t_staged = now_seconds();
        // a DMA copy: the copy engines move it, leaving the compute units to
        // the steps running meanwhile (a copy kernel reading host memory held
        // wave slots for the whole transfer: 12.8 K against 15.4 K frames/s
        // served with ingest)
        HIPCHK(hipMemcpyAsync(f->blob.p, up->pinned[k], w.size, hipMemcpyHostToDevice, up->s));
        HIPCHK(hipEventRecord(up->ev[k], up->s));
        launch_strip_expand(up->s, dev_bands, f->nbands, dev_strips);
        HIPCHK(hipGetLastError());
        up->next = k ^ 1;
        // not waited for here: the stream that reconstructs the frame waits on
        // the frame's own event (build_batch); the staging buffer is reused
        // after ev[k]
        if (!f->upload_ev) HIPCHK(hipEventCreateWithFlags(&f->upload_ev, hipEventDisableTiming));
        HIPCHK(hipEventRecord(f->upload_ev, up->s));
        f->prep_s[1] = t_lists - t_start;
        f->prep_s[2] = t_staged - t_lists;
        f->prep_s[3] = now_seconds() - t_staged;
    } else {  // complete on return
        std::vector<uint8_t>& buf = W.buf;
        buf.resize(w.size);
        w.copy_to(buf.data());
        HIPCHK(hipMemcpyAsync(f->blob.p, buf.data(), w.size, hipMemcpyHostToDevice, e->upload));
        launch_strip_expand(e->upload, dev_bands, f->nbands, dev_strips);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(e->upload));
    }
    return HEVCR_OK;
}

// -- batches --------------------------------------------------------------------

// Work memory of one frame inside a batch's arena.
struct WorkLayout {
    size_t info4 = 0, mot = 0, resid = 0, iflags = 0, scratch = 0, end = 0;
};

WorkLayout work_layout(const Stream& st, const hevcr_frame* f, bool full, size_t at) {
    WorkLayout L;
    const size_t cells = (size_t)f->fc.gw * f->fc.gh;
    L.info4 = at;
    at = align256(at + cells * 4);
    L.mot = at;
    if (f->npu) at = align256(at + cells * sizeof(Motion4));
    L.resid = at;
    if (f->has_intra) at = align256(at + (st.plane_bytes[0] + st.plane_bytes[1] + st.plane_bytes[2]) / st.bpp * 2);
    L.iflags = at;
    if (f->has_intra) at = align256(at + (size_t)f->nctb * 3 * 4);
    L.scratch = at;
    if (full && f->sao_active) at = align256(at + st.picture_bytes());
    L.end = at;
    return L;
}

void frame_dev(const Stream& st, const hevcr_frame* f, const Picture& slot, bool full, const DevBuf& arena,
               const WorkLayout& L, FrameDev& d) {
    std::memset(&d, 0, sizeof(d));
    d.fc = f->fc;
    const bool use_sao = full && f->sao_active;
    char* scratch = arena.as<char>(L.scratch);
    for (int c = 0; c < 3; c++) {
        size_t off = 0;
        for (int k = 0; k < c; k++) off += st.plane_bytes[k];
        d.rec[c] = use_sao ? scratch + off : slot.plane[c];
        d.out[c] = slot.plane[c];
    }
    if (f->has_intra) {
        int16_t* base = arena.as<int16_t>(L.resid);
        d.res[0] = base;
        d.res[1] = base + st.plane_bytes[0] / st.bpp;
        d.res[2] = d.res[1] + st.plane_bytes[1] / st.bpp;
    }
    for (int l = 0; l < 2; l++)
        for (int i = 0; i < f->p.num_ref[l]; i++)
            for (int c = 0; c < 3; c++) d.ref[l][i][c] = st.slots[(size_t)f->p.ref_slot[l][i]]->plane[c];
    d.info4 = arena.as<uint32_t>(L.info4);
    d.iflags = f->has_intra ? arena.as<uint32_t>(L.iflags) : nullptr;
    d.mot = f->npu ? arena.as<Motion4>(L.mot) : nullptr;
    const PoolBuf& B = f->blob;
    d.cus = B.as<hevcr_cu>(f->o_cus);
    d.tuw = B.as<uint32_t>(f->o_tus);
    d.pus = B.as<hevcr_pu>(f->o_pus);
    if (f->compact)
        d.sig = B.as<uint32_t>(f->o_coeffs);
    else
        d.coeffs = B.as<int16_t>(f->o_coeffs);
    d.sao = B.as<hevcr_sao>(f->o_sao);
    d.cu_first_tu = B.as<uint32_t>(f->o_first_tu);
    d.cu_first_pu = B.as<uint32_t>(f->o_first_pu);
    d.ctu_ranges = B.as<uint32_t>(f->o_ctu_ranges);
    d.tiles = B.as<McTile>(f->o_tiles);
    d.tilesb = B.as<McTile>(f->o_tilesb);
    d.ntiles = f->ntiles;
    d.tiles8 = B.as<McTile>(f->o_tiles8);
    d.ntiles8 = f->ntiles8;
    d.strips = B.as<McStrip>(f->o_strips);
    d.stripsb = B.as<McStrip>(f->o_stripsb);
    for (int q = 0; q < 4; q++) {
        d.blocks[q] = B.as<ResBlock>(f->o_blocks[q]);
        d.nblocks[q] = f->nblocks[q];
    }
    d.ops = B.as<IntraOp>(f->o_ops);
    d.op_begin = B.as<uint32_t>(f->o_op_begin);
    d.op_span = B.as<uint8_t>(f->o_op_span);
    d.ictu = B.as<uint32_t>(f->o_ictu);
    d.diag_begin = B.as<uint32_t>(f->o_diag_begin);
    d.ictl = B.as<uint8_t>(f->o_ictl);
    d.scale_m = f->scaled ? B.as<uint8_t>(f->o_scale) : nullptr;
    d.ctb_ctl = f->has_ctl ? B.as<uint32_t>(f->o_ctl) : nullptr;
    d.ctb_seg = f->seg_weights ? B.as<uint8_t>(f->o_ctbseg) : nullptr;
    d.seg_w = f->seg_weights ? B.as<hevcr_weight>(f->o_segw) : nullptr;
    d.seg_wd = f->seg_weights ? d.seg_w + (size_t)f->nseg_w * 2 * HEVCR_MAX_REFS : nullptr;
}

}  // namespace

namespace hevcr {

int create_stream(hevcr_engine* e, int width, int height, int bit_depth, int dpb_slots, bool allocate, int* stream_id) {
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
    if (allocate) {
        int rc = allocate_dpb(e, *st);
        if (rc) return rc;
        HIPCHK(hipStreamSynchronize(e->stream));
    }
    *stream_id = e->next_stream++;
    e->streams[*stream_id] = std::move(st);
    return HEVCR_OK;
}

// Pictures start zeroed, like the reference's (picture.cpp:35). The clear runs
// on the upload stream and is waited for, so that work on any lane sees it.
int allocate_dpb(hevcr_engine* e, Stream& st) {
    st.slots.clear();
    for (int i = 0; i < st.nslots; i++) {
        st.slots.push_back(std::make_unique<Picture>());
        int rc = alloc_picture(st, *st.slots.back());
        if (rc) {
            st.slots.clear();
            return rc;
        }
        HIPCHK(hipMemsetAsync(st.slots.back()->mem.p, 0, st.slots.back()->mem.cap, e->upload));
    }
    HIPCHK(hipStreamSynchronize(e->upload));
    return HEVCR_OK;
}

int ensure_lanes(hevcr_engine* e, int n) {
    if (n < 1 || n > kMaxLanes) return fail(HEVCR_EINVAL, "lanes %d (1..%d)", n, kMaxLanes);
    while ((int)e->more_lanes.size() < n - 1) {
        Lane l;
        if (hipStreamCreateWithFlags(&l.main, hipStreamNonBlocking) != hipSuccess ||
            hipStreamCreateWithFlags(&l.side, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&l.fork, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&l.join, hipEventDisableTiming) != hipSuccess) {
            if (l.join) (void)hipEventDestroy(l.join);
            if (l.fork) (void)hipEventDestroy(l.fork);
            if (l.side) (void)hipStreamDestroy(l.side);
            if (l.main) (void)hipStreamDestroy(l.main);
            return fail(HEVCR_EDEVICE, "lane %d: hipStreamCreate failed", (int)e->more_lanes.size() + 1);
        }
        e->more_lanes.push_back(l);
    }
    return HEVCR_OK;
}

Lane lane(const hevcr_engine* e, int i) {
    if (i <= 0 || i > (int)e->more_lanes.size()) return Lane{e->stream, e->side, e->fork, e->join};
    return e->more_lanes[(size_t)(i - 1)];
}

int sync_lanes(hevcr_engine* e) {
    HIPCHK(hipStreamSynchronize(e->stream));
    HIPCHK(hipStreamSynchronize(e->side));
    for (auto& l : e->more_lanes) {
        HIPCHK(hipStreamSynchronize(l.main));
        HIPCHK(hipStreamSynchronize(l.side));
    }
    return HEVCR_OK;
}

void run_bytes(const hevcr_frame* f, int stop, double* stage_bytes, double* kernel_bytes) {
    for (int i = 0; i < HEVCR_NUM_STAGES; i++) stage_bytes[i] = f->bytes[i];
    for (int k = 0; k < HEVCR_NUM_KERNELS; k++) kernel_bytes[k] = f->kbytes[k];
    if (stop != HEVCR_RUN_FULL) stage_bytes[HEVCR_STAGE_SAO] = kernel_bytes[HEVCR_KERNEL_SAO] = 0;
    if (stop == HEVCR_RUN_STOP_RECON) stage_bytes[HEVCR_STAGE_DEBLOCK] = kernel_bytes[HEVCR_KERNEL_DEBLOCK] = 0;
    if (stop == HEVCR_RUN_FULL && f->lf_fused) {  // one kernel, counted under the SAO stage
        stage_bytes[HEVCR_STAGE_DEBLOCK] = kernel_bytes[HEVCR_KERNEL_DEBLOCK] = kernel_bytes[HEVCR_KERNEL_SAO] = 0;
        stage_bytes[HEVCR_STAGE_SAO] = kernel_bytes[HEVCR_KERNEL_LOOPFILTER] = f->lf_bytes;
    }
}

int build_batch(hevcr_engine* e, hevcr_frame* const* frames, const int* slots, int n, int stop, hevcr_batch* b,
                int lane_index) {
    if (!e || !frames || !slots || !b || n < 1 || n > 65535) return fail(HEVCR_EINVAL, "batch arguments");
    if (stop < HEVCR_RUN_FULL || stop > HEVCR_RUN_STOP_DEBLOCK) return fail(HEVCR_EINVAL, "stop %d", stop);
    std::set<std::pair<int, int>> written;
    for (int i = 0; i < n; i++) {
        hevcr_frame* f = frames[i];
        if (!f) return fail(HEVCR_EINVAL, "null frame %d", i);
        Stream* st = find_stream(e, f->stream_id);
        if (!st) return fail(HEVCR_EINVAL, "frame %d: stream %d is gone", i, f->stream_id);
        if (!st->allocated()) return fail(HEVCR_EINVAL, "frame %d: stream %d has no DPB yet", i, f->stream_id);
        if (slots[i] < 0 || slots[i] >= st->nslots) return fail(HEVCR_EINVAL, "frame %d: slot %d", i, slots[i]);
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
    if (b->staged) HIPCHK(hipEventSynchronize(b->staged));  // the staging buffer is free again
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->gexec = nullptr;
    b->graph = nullptr;
    b->frames.assign(frames, frames + n);
    b->slots.assign(slots, slots + n);
    b->stop = stop;
    const bool full = stop == HEVCR_RUN_FULL;
    std::memset(b->bytes, 0, sizeof(b->bytes));
    for (int gi = 0; gi < 2; gi++) {
        b->groups[gi] = BatchGroup();
        b->groups[gi].bpp = gi + 1;
    }
    std::memset(b->kbytes, 0, sizeof(b->kbytes));
    for (int i = 0; i < n; i++) {
        b->groups[frames[i]->bpp == 1 ? 0 : 1].members.push_back(i);
        double sb[HEVCR_NUM_STAGES], kb[HEVCR_NUM_KERNELS];
        run_bytes(frames[i], stop, sb, kb);
        for (int s = 0; s < HEVCR_NUM_STAGES; s++) b->bytes[s] += sb[s];
        for (int k = 0; k < HEVCR_NUM_KERNELS; k++) b->kbytes[k] += kb[k];
    }

    // layouts: work arena per frame; FrameDev arrays then prefix arrays per group
    std::vector<WorkLayout> work((size_t)n);
    size_t at = 0;
    for (int i = 0; i < n; i++) {
        work[(size_t)i] = work_layout(*find_stream(e, frames[i]->stream_id), frames[i], full, at);
        at = work[(size_t)i].end;
    }
    b->o_ctl = at;  // the two groups' persistent intra tickets
    const size_t arena_bytes = at + 256;
    size_t meta_bytes = 0;
    for (auto& g : b->groups) {
        size_t ndiag = 0;
        for (int m : g.members) ndiag = std::max(ndiag, frames[m]->diag_count.size());
        g.diag_total.assign(ndiag, 0);
        g.o_fdev = meta_bytes;
        meta_bytes = align256(meta_bytes + g.members.size() * sizeof(FrameDev));
        g.o_index = meta_bytes;
        meta_bytes = align256(meta_bytes + (kPreDiag + ndiag) * (g.members.size() + 1) * 4);
        g.o_wave = meta_bytes;
        meta_bytes = align256(meta_bytes + (ndiag + 1) * 4);
    }
    if (arena_bytes > b->arena.cap || meta_bytes > b->meta.cap) {
        int rc = sync_lanes(e);  // queued work may still read the old blocks
        if (rc) return rc;
        if ((rc = b->arena.reserve(arena_bytes)) || (rc = b->meta.reserve(meta_bytes))) return rc;
    }
    int rc = b->staging.reserve(meta_bytes);
    if (rc) return rc;
    char* host = static_cast<char*>(b->staging.p);
    std::memset(host, 0, meta_bytes);
    for (auto& g : b->groups) {
        const int gn = g.n();
        if (!gn) continue;
        FrameDev* devs = reinterpret_cast<FrameDev*>(host + g.o_fdev);
        g.max_ctb_log2 = 4;
        g.max_nctb = 0;
        g.strip_rows = 0;
        g.intra_persistent = true;
        bool any_dense = false;  // the persistent form pays off for intra pictures (long dependency chains)
        uint32_t* pre[kPreDiag];
        for (int k = 0; k < kPreDiag; k++) pre[k] = reinterpret_cast<uint32_t*>(host + g.prefix_off(k));
        for (int k = 0; k < gn; k++) {
            const int fi = g.members[(size_t)k];
            hevcr_frame* f = frames[fi];
            Stream& st = *find_stream(e, f->stream_id);
            frame_dev(st, f, *st.slots[(size_t)slots[fi]], full, b->arena, work[(size_t)fi], devs[k]);
            const bool fused = full && f->lf_fused;
            const bool dbk = stop != HEVCR_RUN_STOP_RECON && f->fc.deblock && !fused;
            const uint32_t dbk_tiles = (uint32_t)(((f->fc.W + 4 + 63) >> 6) * ((f->fc.H + 4 + 63) >> 6));
            pre[kPreCu][k + 1] = pre[kPreCu][k] + f->nctb;
            g.max_ctb_log2 = std::max(g.max_ctb_log2, f->p.ctb_log2);
            g.max_nctb = std::max(g.max_nctb, (int)f->nctb);
            g.intra_persistent = g.intra_persistent && f->tune.intra_persistent;
            any_dense = any_dense || f->intra_dense;
            pre[kPreTiles][k + 1] = pre[kPreTiles][k] + (f->ntiles + 1) / 2;  // uni tiles in pairs
            pre[kPreTilesB][k + 1] = pre[kPreTilesB][k] + f->ntilesb;
            pre[kPreTiles8][k + 1] = pre[kPreTiles8][k] + (f->ntiles8 + 3) / 4;
            pre[kPreStrips][k + 1] = pre[kPreStrips][k] + f->nstripw;
            pre[kPreStripsB][k + 1] = pre[kPreStripsB][k] + f->nstripwb;
            g.strip_rows = std::max(g.strip_rows, f->strip_rows);
            for (int q = 0; q < 4; q++)
                pre[kPreRes + q][k + 1] =
                    pre[kPreRes + q][k] + (f->nblocks[q] + kResPerGroup[q] - 1) / (uint32_t)kResPerGroup[q];
            pre[kPreDbk][k + 1] = pre[kPreDbk][k] + (dbk ? dbk_tiles : 0);
            pre[kPreSao][k + 1] = pre[kPreSao][k] + (full && f->sao_active && !fused ? f->nctb : 0);
            pre[kPreLf][k + 1] = pre[kPreLf][k] + (fused ? (uint32_t)(((f->fc.W + 63) >> 6) * ((f->fc.H + 63) >> 6)) : 0);
            for (size_t dg = 0; dg < g.diag_total.size(); dg++) {
                uint32_t* row = reinterpret_cast<uint32_t*>(host + g.prefix_off(kPreDiag + (int)dg));
                const uint32_t cnt = dg < f->diag_count.size() ? f->diag_count[dg] : 0;
                row[k + 1] = row[k] + cnt;
                g.diag_total[dg] += cnt;
            }
        }
        g.total_cu = pre[kPreCu][gn];
        g.total_tiles = pre[kPreTiles][gn];
        g.total_tilesb = pre[kPreTilesB][gn];
        g.total_tiles8 = pre[kPreTiles8][gn];
        g.total_stripw = pre[kPreStrips][gn];
        g.total_stripwb = pre[kPreStripsB][gn];
        for (int q = 0; q < 4; q++) g.total_res[q] = pre[kPreRes + q][gn];
        g.total_dbk = pre[kPreDbk][gn];
        g.total_sao = pre[kPreSao][gn];
        g.total_lf = pre[kPreLf][gn];
        uint32_t* ws = reinterpret_cast<uint32_t*>(host + g.o_wave);
        for (size_t dg = 0; dg < g.diag_total.size(); dg++) ws[dg + 1] = ws[dg] + 3 * g.diag_total[dg];
        g.intra_items = ws[g.diag_total.size()];
        // The persistent form pays off where the per-wave launches would leave
        // most of the GPU idle: an intra picture's long dependency chain over
        // waves of fewer than ~1024 items (a few frames per batch). With many
        // intra pictures in a batch the waves fill the GPU by themselves, and
        // the per-item synchronisation (cache write-back and invalidation at
        // device scope) costs more than the launches it saves.
        const size_t nw = g.diag_total.size();
        g.intra_persistent = g.intra_persistent && any_dense && nw && g.intra_items / nw < 1024;
    }
    if (!b->staged) HIPCHK(hipEventCreateWithFlags(&b->staged, hipEventDisableTiming));
    const hipStream_t ls = lane(e, lane_index).main;
    // frames whose upload was left running: the lane's work waits for it
    for (int i = 0; i < n; i++)
        if (frames[i]->upload_ev) HIPCHK(hipStreamWaitEvent(ls, frames[i]->upload_ev, 0));
    HIPCHK(hipMemcpyAsync(b->meta.p, host, meta_bytes, hipMemcpyHostToDevice, ls));
    HIPCHK(hipEventRecord(b->staged, ls));
    return HEVCR_OK;
}

}  // namespace hevcr

namespace {

template <typename T>
void launch_stage(hevcr_engine* e, hevcr_batch* b, BatchGroup& g, int stage, bool timed, hipStream_t s) {
    const FrameDev* fr = b->meta.as<FrameDev>(g.o_fdev);
    auto pre = [&](int k) { return b->meta.as<uint32_t>(g.prefix_off(k)); };
    const int n = g.n();
    switch (stage) {
        case HEVCR_STAGE_EXPAND: {
            StageTimer t(e, kKernel + HEVCR_KERNEL_EXPAND, timed && g.total_cu);
            launch_expand(s, fr, pre(kPreCu), n, g.total_cu, g.max_ctb_log2);
            break;
        }
        case HEVCR_STAGE_MC: {
            {
                StageTimer t(e, kKernel + HEVCR_KERNEL_MC, timed && g.total_stripw);
                launch_mcs<T>(s, fr, pre(kPreStrips), n, g.total_stripw, false, g.strip_rows);
            }
            {
                StageTimer t(e, kKernel + HEVCR_KERNEL_MC, timed && g.total_stripwb);
                launch_mcs<T>(s, fr, pre(kPreStripsB), n, g.total_stripwb, true, 8);
            }
            {
                StageTimer t(e, kKernel + HEVCR_KERNEL_MC, timed && g.total_tiles);
                launch_mc<T>(s, fr, pre(kPreTiles), n, g.total_tiles);
            }
            {
                StageTimer t(e, kKernel + HEVCR_KERNEL_MC, timed && g.total_tilesb);
                launch_mcb<T>(s, fr, pre(kPreTilesB), n, g.total_tilesb);
            }
            StageTimer t(e, kKernel + HEVCR_KERNEL_MC8, timed && g.total_tiles8);
            launch_mc8<T>(s, fr, pre(kPreTiles8), n, g.total_tiles8);
            break;
        }
        case HEVCR_STAGE_RESIDUAL:
            for (int q = 0; q < 4; q++) {
                StageTimer t(e, kKernel + HEVCR_KERNEL_RESIDUAL4 + q, timed && g.total_res[q]);
                const uint32_t* lists[4] = {pre(kPreRes), pre(kPreRes + 1), pre(kPreRes + 2), pre(kPreRes + 3)};
                uint32_t only[4] = {0, 0, 0, 0};
                only[q] = g.total_res[q];
                launch_residual<T>(s, fr, lists, n, only);
            }
            break;
        case HEVCR_STAGE_DEBLOCK: {
            StageTimer t(e, kKernel + HEVCR_KERNEL_DEBLOCK, timed && g.total_dbk);
            launch_deblock<T>(s, fr, pre(kPreDbk), n, g.total_dbk);
            break;
        }
        case HEVCR_STAGE_SAO: {
            {
                StageTimer t(e, kKernel + HEVCR_KERNEL_SAO, timed && g.total_sao);
                launch_sao<T>(s, fr, pre(kPreSao), n, g.total_sao);
            }
            StageTimer t(e, kKernel + HEVCR_KERNEL_LOOPFILTER, timed && g.total_lf);
            launch_loopfilter<T>(s, fr, pre(kPreLf), n, g.total_lf);
            break;
        }
        default: break;
    }
}

template <typename T>
void launch_intra_diag(hevcr_engine* e, hevcr_batch* b, BatchGroup& g, size_t dg, bool timed, hipStream_t s) {
    if (dg >= g.diag_total.size() || !g.diag_total[dg]) return;
    StageTimer t(e, kKernel + HEVCR_KERNEL_INTRA, timed);
    launch_intra<T>(s, b->meta.as<FrameDev>(g.o_fdev), b->meta.as<uint32_t>(g.prefix_off(kPreDiag + (int)dg)),
                    g.n(), (int)dg, g.diag_total[dg], g.max_ctb_log2);
}

bool stage_has_work(const BatchGroup& g, int stage) {
    if (g.members.empty()) return false;
    switch (stage) {
        case HEVCR_STAGE_EXPAND: return g.total_cu > 0;
        case HEVCR_STAGE_MC: return (g.total_tiles | g.total_tilesb | g.total_tiles8 | g.total_stripw | g.total_stripwb) > 0;
        case HEVCR_STAGE_RESIDUAL: return (g.total_res[0] | g.total_res[1] | g.total_res[2] | g.total_res[3]) > 0;
        case HEVCR_STAGE_INTRA: return !g.diag_total.empty();
        case HEVCR_STAGE_DEBLOCK: return g.total_dbk > 0;
        case HEVCR_STAGE_SAO: return (g.total_sao | g.total_lf) > 0;
    }
    return false;
}

// Timed runs record their events on lane 0's stream: they run on lane 0 only.
int launch_batch(hevcr_engine* e, hevcr_batch* b, bool timed, const Lane& L) {
    // The side information the expansion writes is read by the loop filter only:
    // untimed runs expand on the side stream, concurrently with prediction,
    // residual and intra, and join before the loop filter (or at the end).
    bool joined = true;
    for (int stage = 0; stage < HEVCR_NUM_STAGES; stage++) {
        if (!stage_has_work(b->groups[0], stage) && !stage_has_work(b->groups[1], stage)) continue;
        if (stage == HEVCR_STAGE_EXPAND && !timed) {
            HIPCHK(hipEventRecord(L.fork, L.main));
            HIPCHK(hipStreamWaitEvent(L.side, L.fork, 0));
            for (int gi = 0; gi < 2; gi++) {
                if (!stage_has_work(b->groups[gi], stage)) continue;
                if (gi == 0)
                    launch_stage<uint8_t>(e, b, b->groups[0], stage, false, L.side);
                else
                    launch_stage<uint16_t>(e, b, b->groups[1], stage, false, L.side);
            }
            HIPCHK(hipEventRecord(L.join, L.side));
            joined = false;
            continue;
        }
        if (!joined && (stage == HEVCR_STAGE_DEBLOCK || stage == HEVCR_STAGE_SAO)) {
            HIPCHK(hipStreamWaitEvent(L.main, L.join, 0));
            joined = true;
        }
        StageTimer t(e, stage, timed);
        if (stage == HEVCR_STAGE_INTRA) {
            const size_t nd = std::max(b->groups[0].diag_total.size(), b->groups[1].diag_total.size());
            for (int gi = 0; gi < 2; gi++) {  // persistent groups: one launch each
                BatchGroup& g = b->groups[gi];
                if (!stage_has_work(g, stage) || !g.intra_persistent) continue;
                StageTimer kt(e, kKernel + HEVCR_KERNEL_INTRA, timed);
                uint32_t* ticket = b->arena.as<uint32_t>(b->o_ctl + 4 * (size_t)gi);
                const FrameDev* fr = b->meta.as<FrameDev>(g.o_fdev);
                const uint32_t* pre = b->meta.as<uint32_t>(g.prefix_off(kPreDiag));
                const uint32_t* ws = b->meta.as<uint32_t>(g.o_wave);
                if (gi == 0)
                    launch_intra_persistent<uint8_t>(L.main, fr, g.n(), g.max_nctb, pre, ws, (int)g.diag_total.size(),
                                                     g.intra_items, ticket, e->intra_err.as<uint32_t>(),
                                                     e->intra_spin_limit, g.max_ctb_log2);
                else
                    launch_intra_persistent<uint16_t>(L.main, fr, g.n(), g.max_nctb, pre, ws, (int)g.diag_total.size(),
                                                      g.intra_items, ticket, e->intra_err.as<uint32_t>(),
                                                      e->intra_spin_limit, g.max_ctb_log2);
            }
            for (size_t dg = 0; dg < nd; dg++) {  // the others: one launch per dependency wave
                if (!b->groups[0].intra_persistent) launch_intra_diag<uint8_t>(e, b, b->groups[0], dg, timed, L.main);
                if (!b->groups[1].intra_persistent) launch_intra_diag<uint16_t>(e, b, b->groups[1], dg, timed, L.main);
            }
        } else {
            if (stage_has_work(b->groups[0], stage)) launch_stage<uint8_t>(e, b, b->groups[0], stage, timed, L.main);
            if (stage_has_work(b->groups[1], stage)) launch_stage<uint16_t>(e, b, b->groups[1], stage, timed, L.main);
        }
    }
    if (!joined) HIPCHK(hipStreamWaitEvent(L.main, L.join, 0));
    return last_launch_error();
}

}  // namespace

int hevcr::run_batch(hevcr_engine* e, hevcr_batch* b, bool graph, int lane_index) {
    HIPCHK(hipSetDevice(e->device));
    for (size_t i = 0; i < b->frames.size(); i++)
        if (!e->streams.count(b->frames[i]->stream_id)) return fail(HEVCR_EINVAL, "a stream of this batch is gone");
    const Lane L = lane(e, lane_index);
    if (!graph) return launch_batch(e, b, e->profiling && L.main == e->stream, L);
    if (!b->gexec) {
        HIPCHK(hipStreamBeginCapture(L.main, hipStreamCaptureModeThreadLocal));
        const int rc = launch_batch(e, b, false, L);
        hipGraph_t gr = nullptr;
        const hipError_t ec = hipStreamEndCapture(L.main, &gr);
        if (rc) {
            if (gr) (void)hipGraphDestroy(gr);
            return rc;
        }
        if (ec != hipSuccess) return fail(HEVCR_EDEVICE, "hipStreamEndCapture: %s", hipGetErrorString(ec));
        b->graph = gr;
        HIPCHK(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
    }
    HIPCHK(hipGraphLaunch(b->gexec, L.main));
    return HEVCR_OK;
}

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------

int hevcr::prepare_checked(hevcr_engine* e, Stream& st, int stream_id, const hevcr_frame_syntax* syn, hevcr_frame* f,
                           UploadCtx* up) {
    const double t0 = now_seconds();
    int rc = validate(st, syn, &prep_scratch().res);
    if (rc) return rc;
    f->prep_s[0] = now_seconds() - t0;
    HIPCHK(hipSetDevice(e->device));
    return prepare_frame(e, st, stream_id, syn, f, up);
}

extern "C" {

const char* hevcr_last_error(void) { return hevcr::last_error(); }

int hevcr_engine_create(int device, hevcr_engine** out) {
    if (!out) return fail(HEVCR_EINVAL, "null out");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(HEVCR_EDEVICE, "no GPU visible");
    if (device < 0 || device >= n) return fail(HEVCR_EINVAL, "device %d of %d", device, n);
    HIPCHK(hipSetDevice(device));
    void* scratch = nullptr;
    if (int rc = noalloc_init(device, &scratch)) return rc;
    auto* e = new (std::nothrow) hevcr_engine();
    if (!e) return fail(HEVCR_ENOMEM, "engine allocation");
    e->device = device;
    e->scratch = scratch;
    e->blobs = std::make_shared<BlobPool>();
    e->blobs->device = device;
    if (e->intra_err.reserve(256) || hipMemset(e->intra_err.p, 0, 256) != hipSuccess) {
        hevcr_engine_destroy(e);
        return fail(HEVCR_EDEVICE, "engine error word");
    }
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) {
        delete e;
        return fail(HEVCR_EDEVICE, "hipStreamCreate failed");
    }
    if (hipStreamCreateWithFlags(&e->upload, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&e->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&e->fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&e->join, hipEventDisableTiming) != hipSuccess) {
        if (e->join) (void)hipEventDestroy(e->join);
        if (e->fork) (void)hipEventDestroy(e->fork);
        if (e->side) (void)hipStreamDestroy(e->side);
        if (e->upload) (void)hipStreamDestroy(e->upload);
        (void)hipStreamDestroy(e->stream);
        delete e;
        return fail(HEVCR_EDEVICE, "hipStreamCreate failed");
    }
    *out = e;
    return HEVCR_OK;
}

void hevcr_engine_destroy(hevcr_engine* e) {
    if (!e) return;
    (void)hipSetDevice(e->device);
    (void)sync_lanes(e);
    collect(e);
    for (auto ev : e->free_events) (void)hipEventDestroy(ev);
    e->streams.clear();
    (void)hipStreamSynchronize(e->upload);
    (void)hipStreamSynchronize(e->side);
    for (auto& l : e->more_lanes) {
        (void)hipStreamSynchronize(l.main);
        (void)hipStreamSynchronize(l.side);
        (void)hipEventDestroy(l.join);
        (void)hipEventDestroy(l.fork);
        (void)hipStreamDestroy(l.side);
        (void)hipStreamDestroy(l.main);
    }
    (void)hipEventDestroy(e->join);
    (void)hipEventDestroy(e->fork);
    (void)hipStreamDestroy(e->side);
    (void)hipStreamDestroy(e->upload);
    (void)hipStreamDestroy(e->stream);
    delete e;
}

int hevcr_engine_sync(hevcr_engine* e) {
    if (!e) return fail(HEVCR_EINVAL, "null engine");
    HIPCHK(hipSetDevice(e->device));
    int rc = sync_lanes(e);
    if (rc) return rc;
    uint32_t err = 0;
    HIPCHK(hipMemcpy(&err, e->intra_err.p, 4, hipMemcpyDeviceToHost));
    if (err) {
        HIPCHK(hipMemset(e->intra_err.p, 0, 4));
        return fail(HEVCR_EDEVICE, "a persistent intra launch gave up waiting for a neighbour CTU");
    }
    return HEVCR_OK;
}

int hevcr_stream_create(hevcr_engine* e, int width, int height, int bit_depth, int dpb_slots, int* stream_id) {
    return create_stream(e, width, height, bit_depth, dpb_slots, true, stream_id);
}

int hevcr_stream_destroy(hevcr_engine* e, int stream_id) {
    if (!e) return fail(HEVCR_EINVAL, "null engine");
    auto it = e->streams.find(stream_id);
    if (it == e->streams.end()) return fail(HEVCR_EINVAL, "stream %d", stream_id);
    HIPCHK(hipSetDevice(e->device));
    int rc = sync_lanes(e);
    if (rc) return rc;
    e->streams.erase(it);
    return HEVCR_OK;
}

static int picture_upload(hevcr_engine* e, int stream_id, int slot, const uint16_t* const* host) {
    Stream* st = e ? find_stream(e, stream_id) : nullptr;
    if (!st) return fail(HEVCR_EINVAL, "stream %d", stream_id);
    if (!st->allocated()) return fail(HEVCR_EINVAL, "stream %d has no DPB yet", stream_id);
    if (slot < 0 || slot >= st->nslots) return fail(HEVCR_EINVAL, "slot %d", slot);
    for (int c = 0; c < 3; c++)
        if (!host[c]) return fail(HEVCR_EINVAL, "null plane");
    HIPCHK(hipSetDevice(e->device));
    Picture& pic = *st->slots[(size_t)slot];
    for (int c = 0; c < 3; c++) {
        const int w = c ? st->w / 2 : st->w, h = c ? st->h / 2 : st->h;
        std::vector<uint8_t> tmp((size_t)st->stride[c] * h * st->bpp, 0);
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
    }
    return HEVCR_OK;
}

int hevcr_picture_upload(hevcr_engine* e, int stream_id, int slot, const uint16_t* y, const uint16_t* cb,
                         const uint16_t* cr) {
    const uint16_t* h[3] = {y, cb, cr};
    return picture_upload(e, stream_id, slot, h);
}

}  // extern "C"

namespace {

// Output plane sizes of one picture in `format` (bytes per sample in *bps).
int egress_planes(int width, int height, int bit_depth, int format, const int* crop, int w[3], int h[3], int* bps) {
    if (width <= 0 || height <= 0 || (width & 1) || (height & 1)) return fail(HEVCR_EINVAL, "picture %dx%d", width, height);
    if (bit_depth < 8 || bit_depth > 16) return fail(HEVCR_EINVAL, "bit depth %d", bit_depth);
    int l = 0, r = 0, t = 0, b = 0;
    if (format == HEVCR_EGRESS_YUV) {
        if (crop) {
            l = crop[0], r = crop[1], t = crop[2], b = crop[3];
            // 4:2:0 conformance window offsets are in chroma units of 2 luma samples
            if (l < 0 || r < 0 || t < 0 || b < 0 || ((l | r | t | b) & 1) || l + r >= width || t + b >= height)
                return fail(HEVCR_EINVAL, "crop %d %d %d %d of a %dx%d picture", l, r, t, b, width, height);
        }
        *bps = bit_depth > 8 ? 2 : 1;
    } else if (format == HEVCR_EGRESS_U16) {
        *bps = 2;
    } else {
        return fail(HEVCR_EINVAL, "egress format %d", format);
    }
    for (int c = 0; c < 3; c++) {
        const int s = c ? 2 : 1;  // picture.cpp:54-57: chroma offsets are the luma ones / 2
        w[c] = width / s - l / s - r / s;
        h[c] = height / s - t / s - b / s;
    }
    return HEVCR_OK;
}

struct EgressItem {
    const Stream* st;
    const Picture* pic;
    int crop[4];
    uint64_t offset;  // into the destination
};

// Queues the egress of `items` into device-visible memory at `dst` on the
// engine's stream.
int egress_launch(hevcr_engine* e, const std::vector<EgressItem>& items, int format, char* dst) {
    std::vector<EgressPlane> planes[3];  // (1 -> 1), (1 -> 2), (2 -> 2) bytes per sample
    uint64_t most[3] = {0, 0, 0};
    for (const EgressItem& it : items) {
        int w[3], h[3], bps = 0;
        int rc = egress_planes(it.st->w, it.st->h, it.st->bd, format, it.crop, w, h, &bps);
        if (rc) return rc;
        const int v = it.st->bpp == 1 ? (bps == 1 ? 0 : 1) : 2;
        uint64_t at = it.offset;
        for (int c = 0; c < 3; c++) {
            const int s = c ? 2 : 1;
            const int x0 = format == HEVCR_EGRESS_YUV ? it.crop[0] / s : 0, y0 = format == HEVCR_EGRESS_YUV ? it.crop[2] / s : 0;
            EgressPlane pl;
            pl.src = static_cast<const char*>(it.pic->plane[c]) +
                     ((size_t)y0 * (size_t)it.st->stride[c] + (size_t)x0) * (size_t)it.st->bpp;
            pl.dst = dst + at;
            pl.w = w[c];
            pl.h = h[c];
            pl.stride = it.st->stride[c];
            pl.pad = 0;
            planes[v].push_back(pl);
            most[v] = std::max(most[v], (uint64_t)w[c] * (uint64_t)h[c]);
            at += (uint64_t)w[c] * (uint64_t)h[c] * (uint64_t)bps;
        }
    }
    const size_t np = planes[0].size() + planes[1].size() + planes[2].size();
    int rc = e->egress_desc.reserve(np * sizeof(EgressPlane));
    if (rc) return rc;
    std::vector<EgressPlane> all;
    for (auto& v : planes) all.insert(all.end(), v.begin(), v.end());
    EgressPlane* desc = e->egress_desc.as<EgressPlane>();
    HIPCHK(hipMemcpyAsync(desc, all.data(), np * sizeof(EgressPlane), hipMemcpyHostToDevice, e->stream));
    launch_egress(e->stream, desc, (int)planes[0].size(), 1, 1, most[0]);
    launch_egress(e->stream, desc + planes[0].size(), (int)planes[1].size(), 1, 2, most[1]);
    launch_egress(e->stream, desc + planes[0].size() + planes[1].size(), (int)planes[2].size(), 2, 2, most[2]);
    HIPCHK(hipGetLastError());
    return HEVCR_OK;
}

int egress_item(hevcr_engine* e, int stream_id, int slot, const int* crop, EgressItem* it) {
    const Stream* st = find_stream(e, stream_id);
    if (!st) return fail(HEVCR_EINVAL, "stream %d", stream_id);
    if (!st->allocated()) return fail(HEVCR_EINVAL, "stream %d has no DPB yet", stream_id);
    if (slot < 0 || slot >= st->nslots) return fail(HEVCR_EINVAL, "slot %d", slot);
    it->st = st;
    it->pic = st->slots[(size_t)slot].get();
    for (int k = 0; k < 4; k++) it->crop[k] = crop ? crop[k] : 0;
    return HEVCR_OK;
}

}  // namespace

extern "C" {

int hevcr_picture_download(hevcr_engine* e, int stream_id, int slot, uint16_t* y, uint16_t* cb, uint16_t* cr) {
    if (!e || !y || !cb || !cr) return fail(HEVCR_EINVAL, "null argument");
    std::vector<EgressItem> items(1);
    int rc = egress_item(e, stream_id, slot, nullptr, &items[0]);
    if (rc) return rc;
    items[0].offset = 0;
    HIPCHK(hipSetDevice(e->device));
    const Stream* st = items[0].st;
    const size_t ny = (size_t)st->w * st->h, nc = ny / 4;
    if ((rc = e->egress_stage.reserve((ny + 2 * nc) * 2))) return rc;
    char* stage = e->egress_stage.as<char>();
    if ((rc = egress_launch(e, items, HEVCR_EGRESS_U16, stage))) return rc;
    HIPCHK(hipMemcpyAsync(y, stage, ny * 2, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipMemcpyAsync(cb, stage + ny * 2, nc * 2, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipMemcpyAsync(cr, stage + (ny + nc) * 2, nc * 2, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return HEVCR_OK;
}

int hevcr_egress_size(int width, int height, int bit_depth, int format, const int* crop, uint64_t* bytes) {
    if (!bytes) return fail(HEVCR_EINVAL, "null argument");
    int w[3], h[3], bps = 0;
    int rc = egress_planes(width, height, bit_depth, format, crop, w, h, &bps);
    if (rc) return rc;
    *bytes = 0;
    for (int c = 0; c < 3; c++) *bytes += (uint64_t)w[c] * (uint64_t)h[c] * (uint64_t)bps;
    return HEVCR_OK;
}

int hevcr_picture_egress(hevcr_engine* e, int n, const int* stream_ids, const int* slots, const int* crops, int format,
                         void* dst, uint64_t dst_bytes) {
    if (!e || n < 0 || (n && (!stream_ids || !slots || !dst))) return fail(HEVCR_EINVAL, "null argument");
    if (!n) return HEVCR_OK;
    std::vector<EgressItem> items((size_t)n);
    uint64_t total = 0;
    for (int i = 0; i < n; i++) {
        const int* crop = crops ? crops + 4 * i : nullptr;
        int rc = egress_item(e, stream_ids[i], slots[i], crop, &items[(size_t)i]);
        uint64_t sz = 0;
        if (!rc) rc = hevcr_egress_size(items[(size_t)i].st->w, items[(size_t)i].st->h, items[(size_t)i].st->bd, format, crop, &sz);
        if (rc) return rc;
        items[(size_t)i].offset = total;
        total += sz;
    }
    if (total > dst_bytes) return fail(HEVCR_EINVAL, "egress of %d pictures needs %llu bytes, dst holds %llu", n,
                                       (unsigned long long)total, (unsigned long long)dst_bytes);
    HIPCHK(hipSetDevice(e->device));
    // device memory and pinned host memory are written by the kernels; other
    // host memory through a device staging buffer
    hipPointerAttribute_t attr;
    std::memset(&attr, 0, sizeof(attr));
    char* direct = nullptr;
    if (hipPointerGetAttributes(&attr, dst) == hipSuccess &&
        (attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeHost) && attr.devicePointer) {
        if (attr.type == hipMemoryTypeDevice && attr.device != e->device)
            return fail(HEVCR_EINVAL, "dst is memory of device %d, the engine runs on device %d", attr.device, e->device);
        direct = static_cast<char*>(attr.devicePointer);
    } else {
        (void)hipGetLastError();  // not HIP memory: clear the query's error
    }
    int rc;
    if (direct) {
        if ((rc = egress_launch(e, items, format, direct))) return rc;
    } else {
        if ((rc = e->egress_stage.reserve(total))) return rc;
        if ((rc = egress_launch(e, items, format, e->egress_stage.as<char>()))) return rc;
        HIPCHK(hipMemcpyAsync(dst, e->egress_stage.p, total, hipMemcpyDeviceToHost, e->stream));
    }
    HIPCHK(hipStreamSynchronize(e->stream));
    return HEVCR_OK;
}

int hevcr_host_alloc(uint64_t bytes, void** out) {
    if (!out) return fail(HEVCR_EINVAL, "null argument");
    *out = nullptr;
    if (!bytes) return fail(HEVCR_EINVAL, "0 bytes");
    if (hipHostMalloc(out, bytes, hipHostMallocMapped | hipHostMallocPortable) != hipSuccess) {
        *out = nullptr;
        return fail(HEVCR_ENOMEM, "hipHostMalloc of %llu bytes", (unsigned long long)bytes);
    }
    return HEVCR_OK;
}

void hevcr_host_free(void* p) {
    if (p) (void)hipHostFree(p);
}

int hevcr_frame_prepare(hevcr_engine* e, int stream_id, const hevcr_frame_syntax* syn, hevcr_frame** out) {
    if (!e || !syn || !out) return fail(HEVCR_EINVAL, "null argument");
    *out = nullptr;
    Stream* st = find_stream(e, stream_id);
    if (!st) return fail(HEVCR_EINVAL, "stream %d", stream_id);
    auto f = std::make_unique<hevcr_frame>();
    int rc = prepare_checked(e, *st, stream_id, syn, f.get(), nullptr);
    if (rc) return rc;
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
        if (!f->plain) f->plain = new (std::nothrow) hevcr_batch();
        if (!f->plain) return fail(HEVCR_ENOMEM, "batch allocation");
        hevcr_frame* one[1] = {f};
        f->plain_slot = -1;
        int rc = build_batch(e, one, &out_slot, 1, stop, f->plain);
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
        (void)sync_lanes(e);
    }
    delete f;
}

int hevcr_frame_bytes(const hevcr_frame* f, double* bytes) {
    if (!f || !bytes) return fail(HEVCR_EINVAL, "null argument");
    double kb[HEVCR_NUM_KERNELS];
    run_bytes(f, HEVCR_RUN_FULL, bytes, kb);
    return HEVCR_OK;
}

int hevcr_frame_kernel_bytes(const hevcr_frame* f, double* bytes) {
    if (!f || !bytes) return fail(HEVCR_EINVAL, "null argument");
    double sb[HEVCR_NUM_STAGES];
    run_bytes(f, HEVCR_RUN_FULL, sb, bytes);
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
    if (!e || !out) return fail(HEVCR_EINVAL, "null argument");
    HIPCHK(hipSetDevice(e->device));
    auto b = std::unique_ptr<hevcr_batch>(new (std::nothrow) hevcr_batch());
    if (!b) return fail(HEVCR_ENOMEM, "batch allocation");
    int rc = build_batch(e, frames, out_slots, n, stop, b.get());
    if (rc) return rc;
    *out = b.release();
    return HEVCR_OK;
}

int hevcr_batch_run(hevcr_engine* e, hevcr_batch* b, int use_graph) {
    if (!e || !b) return fail(HEVCR_EINVAL, "null argument");
    return run_batch(e, b, use_graph != 0);
}

void hevcr_batch_release(hevcr_engine* e, hevcr_batch* b) {
    if (!b) return;
    if (e) {
        (void)hipSetDevice(e->device);
        (void)sync_lanes(e);
    }
    delete b;
}

int hevcr_tuning_set(hevcr_engine* e, int width, int height, int bit_depth, const hevcr_tuning* t) {
    if (!e || !t) return fail(HEVCR_EINVAL, "null argument");
    for (int v : {t->mc8, t->lf_fused, t->intra_group4, t->intra_persistent})
        if (v != 0 && v != 1) return fail(HEVCR_EINVAL, "tuning fields are 0 or 1 (mc_strips 0, 1 or 2)");
    if (t->mc_strips < 0 || t->mc_strips > 2) return fail(HEVCR_EINVAL, "tuning fields are 0 or 1 (mc_strips 0, 1 or 2)");
    e->tuning[std::make_tuple(width, height, bit_depth)] = *t;
    return HEVCR_OK;
}

int hevcr_tuning_get(hevcr_engine* e, int width, int height, int bit_depth, hevcr_tuning* t) {
    if (!e || !t) return fail(HEVCR_EINVAL, "null argument");
    const auto it = e->tuning.find(std::make_tuple(width, height, bit_depth));
    *t = it != e->tuning.end() ? it->second : hevcr_tuning{1, 1, 1, 1, 1};
    return HEVCR_OK;
}

int hevcr_autotune(hevcr_engine* e, int n, const int* stream_ids, const hevcr_frame_syntax* const* syn,
                   const int* out_slots, int reps, hevcr_tuning* best, double* ms) {
    if (!e || n <= 0 || !stream_ids || !syn || !out_slots || !best || !ms) return fail(HEVCR_EINVAL, "null argument");
    if (reps < 1 || reps > 1000) return fail(HEVCR_EINVAL, "reps %d", reps);
    const Stream* s0 = find_stream(e, stream_ids[0]);
    if (!s0) return fail(HEVCR_EINVAL, "stream %d", stream_ids[0]);
    const auto key = std::make_tuple(s0->w, s0->h, s0->bd);
    for (int i = 0; i < n; i++) {
        const Stream* st = find_stream(e, stream_ids[i]);
        if (!st) return fail(HEVCR_EINVAL, "stream %d", stream_ids[i]);
        if (std::make_tuple(st->w, st->h, st->bd) != key) return fail(HEVCR_EINVAL, "frames of more than one size");
    }
    HIPCHK(hipSetDevice(e->device));
    const auto had = e->tuning.find(key);
    const bool had_tuning = had != e->tuning.end();
    const hevcr_tuning before = had_tuning ? had->second : hevcr_tuning{1, 1, 1, 1, 1};
    std::vector<uint64_t> want;
    hipEvent_t ev[2] = {nullptr, nullptr};
    int rc = HEVCR_OK, pick = -1;
    double pick_ms = 0;
    for (int k = 0; k < 2 && !rc; k++)
        if (hipEventCreate(&ev[k]) != hipSuccess) rc = fail(HEVCR_EDEVICE, "hipEventCreate failed");
    for (int cand = 0; cand < HEVCR_TUNE_CANDIDATES && !rc; cand++) {
        const hevcr_tuning t{cand & 1, (cand >> 1) & 1, (cand >> 2) & 1, (cand >> 3) & 1, (cand >> 4) & 1};
        e->tuning[key] = t;
        std::vector<hevcr_frame*> frames((size_t)n, nullptr);
        hevcr_batch* b = nullptr;
        for (int i = 0; i < n && !rc; i++) rc = hevcr_frame_prepare(e, stream_ids[i], syn[i], &frames[(size_t)i]);
        if (!rc) rc = hevcr_batch_create(e, frames.data(), out_slots, n, HEVCR_RUN_FULL, &b);
        if (!rc) rc = hevcr_batch_run(e, b, 1);  // captures the graph
        std::vector<uint64_t> got((size_t)n * 3);
        if (!rc) rc = hevcr_picture_digest(e, n, stream_ids, out_slots, got.data());
        if (!rc) {
            if (want.empty())
                want = got;
            else if (got != want)
                rc = fail(HEVCR_ERROR, "tuning candidate %d changes the pictures", cand);
        }
        // HIP failures inside the loop set rc (no early return), so that the
        // candidate's batch and frames are released and the tuning restored
        auto hip = [&](hipError_t err, const char* what) {
            if (!rc && err != hipSuccess) rc = fail(HEVCR_EDEVICE, "%s: %s", what, hipGetErrorString(err));
        };
        float t_ms = 0.f;
        if (!rc) hip(hipEventRecord(ev[0], e->stream), "hipEventRecord");
        for (int r = 0; r < reps && !rc; r++) rc = hevcr_batch_run(e, b, 1);
        if (!rc) hip(hipEventRecord(ev[1], e->stream), "hipEventRecord");
        if (!rc) hip(hipEventSynchronize(ev[1]), "hipEventSynchronize");
        if (!rc) hip(hipEventElapsedTime(&t_ms, ev[0], ev[1]), "hipEventElapsedTime");
        if (!rc) {
            ms[cand] = (double)t_ms / reps;
            if (pick < 0 || ms[cand] < pick_ms) {
                pick = cand;
                pick_ms = ms[cand];
            }
        }
        hevcr_batch_release(e, b);
        for (hevcr_frame* f : frames) hevcr_frame_release(e, f);
    }
    for (auto x : ev)
        if (x) (void)hipEventDestroy(x);
    if (rc) {  // leave the tuning as it was
        if (had_tuning)
            e->tuning[key] = before;
        else
            e->tuning.erase(key);
        return rc;
    }
    *best = hevcr_tuning{pick & 1, (pick >> 1) & 1, (pick >> 2) & 1, (pick >> 3) & 1, (pick >> 4) & 1};
    e->tuning[key] = *best;
    return HEVCR_OK;
}

int hevcr_batch_bytes(const hevcr_batch* b, double* bytes) {
    if (!b || !bytes) return fail(HEVCR_EINVAL, "null argument");
    for (int i = 0; i < HEVCR_NUM_STAGES; i++) bytes[i] = b->bytes[i];
    return HEVCR_OK;
}

int hevcr_batch_kernel_bytes(const hevcr_batch* b, double* bytes) {
    if (!b || !bytes) return fail(HEVCR_EINVAL, "null argument");
    for (int i = 0; i < HEVCR_NUM_KERNELS; i++) bytes[i] = b->kbytes[i];
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
        if (ms) ms[i] = e->timer_ms[i];
        if (launches) launches[i] = e->timer_n[i];
    }
    return HEVCR_OK;
}

int hevcr_kernel_times(hevcr_engine* e, double* ms, uint64_t* launches) {
    if (!e) return fail(HEVCR_EINVAL, "null engine");
    HIPCHK(hipSetDevice(e->device));
    int rc = collect(e);
    if (rc) return rc;
    for (int i = 0; i < HEVCR_NUM_KERNELS; i++) {
        if (ms) ms[i] = e->timer_ms[kKernel + i];
        if (launches) launches[i] = e->timer_n[kKernel + i];
    }
    return HEVCR_OK;
}

const char* hevcr_kernel_name(int kernel) {
    static const char* const names[HEVCR_NUM_KERNELS] = {"k_expand", "k_mc", "k_mc8", "k_residual<4x4>",
                                                         "k_residual<8x8>", "k_residual<16x16>", "k_residual<32x32>",
                                                         "k_intra", "k_deblock", "k_sao", "k_loopfilter"};
    return kernel >= 0 && kernel < HEVCR_NUM_KERNELS ? names[kernel] : "";
}

int hevcr_reset_stage_times(hevcr_engine* e) {
    if (!e) return fail(HEVCR_EINVAL, "null engine");
    HIPCHK(hipSetDevice(e->device));
    int rc = collect(e);
    std::memset(e->timer_ms, 0, sizeof(e->timer_ms));
    std::memset(e->timer_n, 0, sizeof(e->timer_n));
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
        lists[log2[i] - 2].push_back(rb);
    }
    HIPCHK(hipSetDevice(e->device));
    hipStream_t s = e->stream;
    HIPCHK(hipStreamSynchronize(s));  // the buffers below may still be read by a previous call
    int rc;
    if ((rc = e->kat_coeffs.reserve(std::max<uint64_t>(num_coeffs, 1) * 2)) ||
        (rc = e->kat_out.reserve((size_t)std::max(rows, 1u) * 32 * 2)))
        return rc;
    size_t total = 0;
    for (auto& l : lists) total += l.size();
    if ((rc = e->kat_blocks.reserve(std::max<size_t>(total, 1) * sizeof(ResBlock)))) return rc;
    if (num_coeffs) HIPCHK(hipMemcpyAsync(e->kat_coeffs.p, coeffs, num_coeffs * 2, hipMemcpyHostToDevice, s));
    // one frame: a descriptor, then one two-entry prefix array per block size
    struct Meta {
        FrameDev d;
        uint32_t pre[4][2];
    } m;
    std::memset(&m, 0, sizeof(m));
    m.d.fc.bd = bit_depth;
    m.d.fc.maxv = (1 << bit_depth) - 1;
    m.d.fc.stride[0] = 32;
    m.d.coeffs = e->kat_coeffs.as<int16_t>();
    m.d.res[0] = e->kat_out.as<int16_t>();
    // the kernel reads the prediction under every block before it knows whether the block
    // adds to it; point it at valid memory (the values are not used: no block adds here)
    for (int c = 0; c < 3; c++) m.d.rec[c] = e->kat_out.p;
    size_t at = 0;
    uint32_t wgs[4];
    for (int k = 0; k < 4; k++) {
        m.d.blocks[k] = e->kat_blocks.as<ResBlock>(at * sizeof(ResBlock));
        m.d.nblocks[k] = (uint32_t)lists[k].size();
        if (!lists[k].empty())
            HIPCHK(hipMemcpyAsync(e->kat_blocks.as<ResBlock>(at * sizeof(ResBlock)), lists[k].data(),
                                  lists[k].size() * sizeof(ResBlock), hipMemcpyHostToDevice, s));
        at += lists[k].size();
        wgs[k] = (m.d.nblocks[k] + kResPerGroup[k] - 1) / (uint32_t)kResPerGroup[k];
        m.pre[k][1] = wgs[k];
    }
    if ((rc = e->kat_meta.reserve(sizeof(m)))) return rc;
    HIPCHK(hipMemcpyAsync(e->kat_meta.p, &m, sizeof(m), hipMemcpyHostToDevice, s));
    const uint32_t* pre[4];
    for (int k = 0; k < 4; k++) pre[k] = reinterpret_cast<const uint32_t*>(e->kat_meta.as<char>(offsetof(Meta, pre)) + k * 8);
    launch_residual<uint16_t>(s, e->kat_meta.as<FrameDev>(), pre, 1, wgs);
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

int hevcr_syntax_compact(const hevcr_frame_syntax* syn, uint32_t* sig, uint64_t cap, uint64_t* num_sig,
                         uint32_t* tu_coeff_off) {
    if (!syn || !num_sig || (sig && !tu_coeff_off)) return fail(HEVCR_EINVAL, "null argument");
    if (syn->sig) return fail(HEVCR_EINVAL, "the syntax is compact already");
    if (syn->num_tus && !syn->tus) return fail(HEVCR_EINVAL, "null syntax array");
    uint64_t n = 0;
    for (uint32_t ti = 0; ti < syn->num_tus; ti++) {
        const hevcr_tu& t = syn->tus[ti];
        const bool chroma_here = t.log2_size > 2 || t.blk_idx == 3;
        const int nc = t.log2_size > 2 ? 1 << (t.log2_size - 1) : 4;
        for (int c = 0; c < 3; c++) {
            if (sig) tu_coeff_off[3 * ti + c] = 0;
            if (!((t.cbf >> c) & 1) || (c && !chroma_here)) continue;
            const int bn = c ? nc : 1 << t.log2_size;
            if (t.log2_size < 2 || t.log2_size > 5 || (uint64_t)t.coeff_off[c] + (uint64_t)bn * bn > syn->num_coeffs ||
                !syn->coeffs)
                return fail(HEVCR_EINVAL, "tu %u coefficients out of range", ti);
            const int16_t* blk = syn->coeffs + t.coeff_off[c];
            uint32_t k = 0;
            for (int i = 0; i < bn * bn; i++) k += blk[i] != 0;
            if (sig) {
                if (n + 1 + k > cap) return fail(HEVCR_EINVAL, "compact buffer holds %llu words", (unsigned long long)cap);
                tu_coeff_off[3 * ti + c] = (uint32_t)n;
                sig[n] = k;
                uint64_t at = n + 1;
                for (int i = 0; i < bn * bn; i++)
                    if (blk[i]) sig[at++] = (uint32_t)i | ((uint32_t)(uint16_t)blk[i] << 16);
            }
            n += 1 + k;
            if (n > 0xffffffffull) return fail(HEVCR_EINVAL, "compact form exceeds 2^32 words");
        }
    }
    *num_sig = n;
    return HEVCR_OK;
}

int hevcr_picture_digest(hevcr_engine* e, int n, const int* stream_ids, const int* slots, uint64_t* digests) {
    if (!e || n < 0 || (n && (!stream_ids || !slots || !digests))) return fail(HEVCR_EINVAL, "null argument");
    if (!n) return HEVCR_OK;
    std::vector<PlaneRef> refs[2];  // per pixel size
    std::vector<int> where[2];
    for (int i = 0; i < n; i++) {
        Stream* st = find_stream(e, stream_ids[i]);
        if (!st) return fail(HEVCR_EINVAL, "stream %d", stream_ids[i]);
        if (!st->allocated()) return fail(HEVCR_EINVAL, "stream %d has no DPB yet", stream_ids[i]);
        if (slots[i] < 0 || slots[i] >= st->nslots) return fail(HEVCR_EINVAL, "slot %d", slots[i]);
        const Picture& pic = *st->slots[(size_t)slots[i]];
        const int g = st->bpp == 1 ? 0 : 1;
        for (int c = 0; c < 3; c++) {
            PlaneRef r;
            r.p = pic.plane[c];
            r.w = c ? st->w / 2 : st->w;
            r.h = c ? st->h / 2 : st->h;
            r.stride = st->stride[c];
            r.pad = 0;
            r.key = 0x5851F42D4C957F2Dull * (uint64_t)(c + 1);
            refs[g].push_back(r);
            where[g].push_back(3 * i + c);
        }
    }
    HIPCHK(hipSetDevice(e->device));
    hipStream_t s = e->stream;
    HIPCHK(hipStreamSynchronize(s));
    const size_t np = refs[0].size() + refs[1].size();
    int rc = e->digest_buf.reserve(np * (sizeof(PlaneRef) + 8));
    if (rc) return rc;
    PlaneRef* dref = e->digest_buf.as<PlaneRef>();
    uint64_t* dout = e->digest_buf.as<uint64_t>(np * sizeof(PlaneRef));
    std::vector<PlaneRef> all(refs[0]);
    all.insert(all.end(), refs[1].begin(), refs[1].end());
    HIPCHK(hipMemcpyAsync(dref, all.data(), np * sizeof(PlaneRef), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemsetAsync(dout, 0, np * 8, s));
    launch_digest(s, dref, (int)refs[0].size(), 1, dout);
    launch_digest(s, dref + refs[0].size(), (int)refs[1].size(), 2, dout + refs[0].size());
    HIPCHK(hipGetLastError());
    std::vector<uint64_t> host(np);
    HIPCHK(hipMemcpyAsync(host.data(), dout, np * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    size_t k = 0;
    for (int g = 0; g < 2; g++)
        for (int w : where[g]) digests[w] = host[k++];
    return HEVCR_OK;
}

int hevcr_picture_sha256(hevcr_engine* e, int n, const int* stream_ids, const int* slots, const int* crops,
                         uint8_t* digests) {
    if (!e || n < 0 || (n && (!stream_ids || !slots || !digests))) return fail(HEVCR_EINVAL, "null argument");
    if (!n) return HEVCR_OK;
    std::vector<ShaPlane> planes[2];  // per sample size
    std::vector<uint64_t> length[2];  // message bytes of each plane
    for (int i = 0; i < n; i++) {
        EgressItem it;
        int rc = egress_item(e, stream_ids[i], slots[i], crops ? crops + 4 * i : nullptr, &it);
        int w[3], h[3], bps = 0;
        if (!rc) rc = egress_planes(it.st->w, it.st->h, it.st->bd, HEVCR_EGRESS_YUV, it.crop, w, h, &bps);
        if (rc) return rc;
        const int g = it.st->bpp == 1 ? 0 : 1;
        for (int c = 0; c < 3; c++) {
            const int sub = c ? 2 : 1;
            ShaPlane pl;
            pl.src = static_cast<const char*>(it.pic->plane[c]) +
                     ((size_t)(it.crop[2] / sub) * (size_t)it.st->stride[c] + (size_t)(it.crop[0] / sub)) *
                         (size_t)it.st->bpp;
            pl.w = w[c];
            pl.h = h[c];
            pl.stride = it.st->stride[c];
            pl.out = 3 * i + c;
            planes[g].push_back(pl);
            length[g].push_back((uint64_t)w[c] * (uint64_t)h[c]);
        }
    }
    // longest first, so that the lanes of a wavefront hash messages of similar length
    for (int g = 0; g < 2; g++) {
        std::vector<size_t> order(planes[g].size());
        for (size_t k = 0; k < order.size(); k++) order[k] = k;
        std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return length[g][a] > length[g][b]; });
        std::vector<ShaPlane> sorted;
        for (size_t k : order) sorted.push_back(planes[g][k]);
        planes[g].swap(sorted);
    }
    HIPCHK(hipSetDevice(e->device));
    hipStream_t s = e->stream;
    const size_t np = planes[0].size() + planes[1].size();
    int rc = e->sha_buf.reserve(np * sizeof(ShaPlane) + (size_t)n * 3 * 32);
    if (rc) return rc;
    ShaPlane* dplanes = e->sha_buf.as<ShaPlane>();
    uint32_t* dout = e->sha_buf.as<uint32_t>(np * sizeof(ShaPlane));
    std::vector<ShaPlane> all(planes[0]);
    all.insert(all.end(), planes[1].begin(), planes[1].end());
    HIPCHK(hipMemcpyAsync(dplanes, all.data(), np * sizeof(ShaPlane), hipMemcpyHostToDevice, s));
    launch_sha256(s, dplanes, (int)planes[0].size(), 1, dout);
    launch_sha256(s, dplanes + planes[0].size(), (int)planes[1].size(), 2, dout);
    HIPCHK(hipGetLastError());
    std::vector<uint32_t> words((size_t)n * 3 * 8);
    HIPCHK(hipMemcpyAsync(words.data(), dout, words.size() * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    for (size_t k = 0; k < words.size(); k++)  // big-endian words (FIPS 180-4 6.2.2 step 4)
        for (int j = 0; j < 4; j++) digests[4 * k + (size_t)j] = (uint8_t)(words[k] >> (24 - 8 * j));
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
